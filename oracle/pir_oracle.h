/*
 * CPU oracle for the ReinsPIRe online answer path — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, on the CPU, the reference algorithms of
 * google/InsPIRe's PIR server answer path (PirServer::ProcessResponse and
 * everything it calls), plus the offline preprocessing and the client that
 * produce and consume its inputs, so that tests can (1) build decryptable
 * seeded inputs, (2) check the HIP product path bit-for-bit, and (3) time a
 * CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library (inspire_amd) never does.
 *
 * Every function cites the reference file:line it restates. Coefficients are
 * held in uint64_t; orc_params.coeff_bits selects the reference's CoeffType
 * (uint64_t, or uint32_t: every coefficient then wraps mod 2^32). log_q is
 * the bit width of the power-of-two modulus (64, or 32 for uint32_t, means
 * q = 0, i.e. arithmetic mod 2^width, as in RlweParams::LogModulus,
 * crypto/encryption.h:46-49).
 *
 * Return convention: 0 = OK, otherwise an absl::StatusCode number
 * (3 INVALID_ARGUMENT, 9 FAILED_PRECONDITION, 13 INTERNAL) with a message
 * written to err (if err != NULL).
 */
#ifndef PIR_ORACLE_H_
#define PIR_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors PirParams<uint64_t> (crypto/pir/pir_params.h:36-105) and
 * RlweParams (crypto/encryption.h:29-92). */
typedef struct {
  int log_degree;       /* d = 2^log_degree */
  int log_q;            /* 1..64; 64 <=> modulus 0 (Z_{2^64}) */
  int log_q1;           /* LogModulus1AfterSwitch */
  int log_q2;           /* LogModulus2AfterSwitch */
  int variance;         /* error-sampler variance parameter */
  uint64_t plaintext_modulus;
  int pack_log_digit, pack_num_digits;
  int eval_log_digit, eval_num_digits;
  int num_entries;      /* n */
  int interp_degree;    /* t */
  int entry_size_multiple; /* ell */
  int db_bytes;         /* 1 (uint8 DB) or 2 (uint16 DB) */
  int mat_bytes;        /* 4 (int32, condensed) or 2 (int16, plain) */
  int coeff_bits;       /* CoeffType: 64 (or 0) uint64_t, 32 uint32_t */
} orc_params;

/* Derived sizes (crypto/pir/pir_params.h:71-83, preprocessing_server.h:127-146). */
typedef struct {
  int d, t, ell, k, e;
  int r;            /* InterpolatedRows = n / t */
  int S;            /* NumFirstDimSamples */
  int K;            /* DB columns after SIMD padding (32 for u16, 64 for u8) */
  int M;            /* ell * t * d (DB / pack rows) */
  int db_words_per_row;   /* K / 4 (u16) or K / 8 (u8) */
  int pack_cols;          /* k * d */
  int pack_words_per_row; /* k*d/2 (int32 condensed) ; 0 for int16 plain */
} orc_sizes;

int orc_get_sizes(const orc_params* P, orc_sizes* out, char* err, size_t errlen);

/* ---- Polynomial coefficient ops (crypto/polynomial.cc) ---- */
void orc_low_bits(const uint64_t* in, int len, int amount, uint64_t* out);
void orc_right_shift(const uint64_t* in, int len, int amount, uint64_t* out);
void orc_rescale(const uint64_t* in, int len, int log_cur, int log_new,
                 uint64_t* out);
int orc_gadget_inv(const uint64_t* in, int len, int log_modulus, int log_digit,
                   int num_digits, uint64_t* out /* num_digits*len */,
                   char* err, size_t errlen);
int orc_signed_gadget_inv(const uint64_t* in, int len, int log_modulus,
                          int log_digit, int num_digits, uint64_t* out,
                          char* err, size_t errlen);
int orc_automorph(const uint64_t* in, int len, int power, uint64_t* out,
                  char* err, size_t errlen);

/* ---- FFT product (crypto/polynomial_fft.cc:68-197, crypto/fft.h) ----
 * u, v: n polynomials of length d each (row-major n x d).
 * max_round_err (optional) receives max |x - round(x)| over every rounded
 * value, the precision margin of the double-precision product. */
int orc_inner_product_fft(const uint64_t* u, const uint64_t* v, int n, int d,
                          int u_bits, int v_bits, int chunk_bits,
                          int u_is_signed, uint64_t* out, double* max_round_err,
                          char* err, size_t errlen);
/* The values InnerProductFft rounds, per accumulated chunk pair (a, b) with
 * chunk_bits*(a+b) < 64, a-major: pre[p*2d + j] = backward_out[j].real() / 2d
 * (polynomial_fft.cc:176-184), before round(), shift and fold; pair_a/pair_b
 * receive (a, b); *npairs their number (pre holds u_chunks*v_chunks*2d
 * doubles). fused = 1 evaluates every complex product as a build with FMA
 * contraction may (fma(xr, wr, -(xi*wi)), fma(xr, wi, xi*wr)). Rounding
 * analysis only. */
int orc_inner_product_fft_pairs(const uint64_t* u, const uint64_t* v, int n, int d,
                                int u_bits, int v_bits, int chunk_bits, int u_is_signed,
                                int fused, double* pre, int* pair_a, int* pair_b,
                                int* npairs, char* err, size_t errlen);
/* Process-wide: fused = 1 makes InnerProductFft and everything built on it
 * (ExternalProduct, FinalizeMatrixPack, ProcessResponse) evaluate complex
 * products with FMA contraction, as above; 0 (default) restores the
 * reference's uncontracted arithmetic. Rounding analysis only. */
void orc_set_fft_contraction(int fused);
/* Exact negacyclic inner product mod 2^64 with the same operand
 * interpretation (chunk truncation / sign extension) as InnerProductFft, by
 * 128-bit schoolbook. A closed-form checker for orc_inner_product_fft. */
int orc_inner_product_exact(const uint64_t* u, const uint64_t* v, int n, int d,
                            int u_bits, int v_bits, int chunk_bits,
                            int u_is_signed, uint64_t* out, char* err,
                            size_t errlen);

/* Width-generic variants: width 32 restates Polynomial<uint32_t>. */
int orc_inner_product_fft_w(const uint64_t* u, const uint64_t* v, int n, int d,
                            int u_bits, int v_bits, int chunk_bits,
                            int u_is_signed, int width, uint64_t* out,
                            double* max_round_err, char* err, size_t errlen);
int orc_inner_product_exact_w(const uint64_t* u, const uint64_t* v, int n, int d,
                              int u_bits, int v_bits, int chunk_bits,
                              int u_is_signed, int width, uint64_t* out,
                              char* err, size_t errlen);
void orc_mt19937_64(uint64_t seed, uint64_t* out, size_t n);
void orc_fft_twiddles(int n, double* fwd, double* bwd);

/* ---- Matrix (crypto/matrix.cc) ---- */
int orc_condense_u16(const uint16_t* data, int rows, int cols, uint64_t* out,
                     char* err, size_t errlen);
int orc_condense_u8(const uint8_t* data, int rows, int cols, uint64_t* out,
                    char* err, size_t errlen);
int orc_condense_i32(const int32_t* data, int rows, int cols, uint64_t* out,
                     char* err, size_t errlen);
/* lanes = Highway lane count of the CPU target (the reference's divisibility
 * check depends on it: matrix.cc:269-272). 8 for AVX-512. */
int orc_condensed_multiply_u16(const uint64_t* cdata, int rows, int cols,
                               const uint64_t* vec, int lanes, uint64_t* out,
                               char* err, size_t errlen);
int orc_condensed_multiply_u8(const uint64_t* cdata, int rows, int cols,
                              const uint64_t* vec, int lanes, uint64_t* out,
                              char* err, size_t errlen);
int orc_condensed_multiply_i32(const uint64_t* cdata, int rows, int cols,
                               const uint64_t* vec, int lanes, uint64_t* out,
                               char* err, size_t errlen);
void orc_multiply_i16(const int16_t* data, int rows, int cols,
                      const uint64_t* vec, uint64_t* out);
/* Row subset of the u16 / i32 products (rows [row0, row0+nrows)); used to time
 * a bounded CPU sample. */
void orc_condensed_multiply_u16_rows(const uint64_t* cdata, int cols, int row0,
                                     int nrows, const uint64_t* vec,
                                     uint64_t* out);
void orc_condensed_multiply_i32_rows(const uint64_t* cdata, int cols, int row0,
                                     int nrows, const uint64_t* vec,
                                     uint64_t* out);

/* ---- Packing (crypto/lwes_to_rlwe.cc) ---- */
int orc_finalize_matrix_pack(const orc_params* P, const uint64_t* b,
                             const uint64_t* b_agg_partial,
                             const uint64_t* y_vec_h /* k x d */,
                             const uint64_t* t_vec_h /* k x d */,
                             const uint64_t* a_tilde_agg /* d */,
                             uint64_t* out_a, uint64_t* out_b,
                             double* max_round_err, char* err, size_t errlen);
/* MatrixPack (lwes_to_rlwe.cc:481-508) with an int32 condensed or int16
 * matrix (d x k*d). */
int orc_matrix_pack(const orc_params* P, const uint64_t* b,
                    const uint64_t* y_vec_g, const uint64_t* y_vec_h,
                    const uint64_t* pack_matrix, const uint64_t* t_vec_h,
                    const uint64_t* a_tilde_agg, uint64_t* out_a,
                    uint64_t* out_b, char* err, size_t errlen);

/* ---- External product / EvalPoly (crypto/external_product.cc,
 *      crypto/pir/server/pir_server.cc:126-150) ----
 * rgsw layout: 2e "a" polys (u.a then v.a) followed by 2e "b" polys. */
int orc_external_product(const orc_params* P, const uint64_t* ct_a,
                         const uint64_t* ct_b, const uint64_t* rgsw_a,
                         const uint64_t* rgsw_b, uint64_t* out_a,
                         uint64_t* out_b, double* max_round_err, char* err,
                         size_t errlen);

/* Threads for the DB and pack products of orc_process_response (default 1,
 * the reference's single-threaded loop); rows split in contiguous slices,
 * results unchanged. */
void orc_set_matrix_threads(int n);

/* ---- Server (crypto/pir/server/pir_server.cc) ----
 * Material layout (concatenated exactly as PirServer::Create does):
 *   db_cdata:   M rows x db_words_per_row condensed words (shard-major)
 *   pack:       M rows x pack_words_per_row words (int32 condensed) or
 *               M rows x k*d int16 values (mat_bytes == 2)
 *   a_tilde:    (ell*t) x d
 *   t_vec_h:    (ell*t) x k x d
 *   second_dim_a: 2e x d (ignored when t == 1)
 * Request: first_dim (len fd_len), second_dim (sd_polys x d),
 *          packing_key (pk_polys x d).
 * Response: ell x (a[d], b[d]).
 * stage_ms (optional, 3 doubles): db-mult, packing, poly-eval wall time. */
int orc_create_check(const orc_params* P, int num_db_matrices, int db_rows,
                     int db_cols, int num_outputs, int outputs_len,
                     int second_dim_a_len, char* err, size_t errlen);
int orc_process_response(const orc_params* P, const uint64_t* db_cdata,
                         int db_cols, const void* pack, const uint64_t* a_tilde,
                         const uint64_t* t_vec_h, const uint64_t* second_dim_a,
                         const uint64_t* first_dim, int fd_len,
                         const uint64_t* second_dim, int sd_polys,
                         const uint64_t* packing_key, int pk_polys,
                         uint64_t* out, double* stage_ms, double* max_round_err,
                         char* err, size_t errlen);

/* ---- Offline preprocessing (crypto/pir/server/preprocessing_server.h,
 *      crypto/db_row_interpolation.cc, crypto/hint_computation.cc,
 *      crypto/lwes_to_rlwe.cc:163-446) ---- */
int orc_interpolate(const uint64_t* row, int d, int t, uint64_t p,
                    uint64_t* out, char* err, size_t errlen);
int orc_preprocess(const orc_params* P, const uint16_t* db_raw,
                   uint64_t public_seed, uint64_t* db_cdata, void* pack,
                   uint64_t* a_tilde, uint64_t* t_vec_h,
                   uint64_t* second_dim_a, char* err, size_t errlen);
/* ComputeHintMatrix<uint64_t, S, uint64_t> (hint_computation.cc:78-160) of a
 * rows x cols DB (values given as uint16, db_bits = 8 * sizeof(S)) against n
 * a-components (n x d); out rows x d. */
int orc_compute_hint_matrix(const uint16_t* db, int rows, int cols, int db_bits,
                            const uint64_t* a, int n, int d, uint64_t* out,
                            char* err, size_t errlen);
/* The a-components orc_preprocess draws from public_seed, in its order:
 * first_dim_a (S x d), second_dim_a (2e x d, only when t > 1), pack_a (2k x d). */
int orc_sample_a_components(const orc_params* P, uint64_t public_seed,
                            uint64_t* first_dim_a, uint64_t* second_dim_a,
                            uint64_t* pack_a, char* err, size_t errlen);
/* PreprocessMatrixPack on explicit LWE a-vectors (d x d) and w polys
 * (2k x d). pack is d x k*d int32 (condensed words) or int16. */
int orc_preprocess_matrix_pack(const orc_params* P, const uint64_t* a_is,
                               const uint64_t* w, void* pack,
                               uint64_t* a_tilde, uint64_t* t_vec_h,
                               char* err, size_t errlen);
/* Pack (lwes_to_rlwe.cc:370-419): the automorphism-based online packing, used
 * to check that BuildPreprocessMatrix is the same linear map. */
int orc_pack_reference(const orc_params* P, const uint64_t* a_is,
                       const uint64_t* w, const uint64_t* b,
                       const uint64_t* y_vec_g, const uint64_t* y_vec_h,
                       uint64_t* out_a, uint64_t* out_b, char* err,
                       size_t errlen);

/* ---- Client (crypto/pir/client/pir_client.cc, crypto/encryption.cc) ----
 * A client is identified by (public_seed, secret_seed); create_request
 * regenerates its secret key from secret_seed. */
int orc_client_request(const orc_params* P, uint64_t public_seed,
                       uint64_t secret_seed, int index, uint64_t* first_dim,
                       uint64_t* second_dim, uint64_t* packing_key, char* err,
                       size_t errlen);
int orc_client_decrypt(const orc_params* P, uint64_t secret_seed,
                       const uint64_t* response, uint64_t* out_msg, char* err,
                       size_t errlen);
/* Raw LWE/RLWE helpers for packing tests. */
int orc_client_secret_key(const orc_params* P, uint64_t secret_seed,
                          uint64_t* out);

/* ---- Bit-packed formats ----
 * Polynomial::ToProto / CreateFromProto (crypto/polynomial.cc:562-590,
 * 208-252): the low log_modulus bits of each coefficient, LSB first, one byte
 * string per polynomial. BitPackCoefficients / BitUnpackCoefficients and
 * BitPackSigned / BitUnpackSigned (crypto/pir/server/pir_server.cc:309-429)
 * use the same bit order; the signed unpack sign-extends. */
int orc_poly_to_proto(const uint64_t* coeffs, int len, int log_modulus,
                      uint8_t* out, size_t out_cap, size_t* out_len, char* err,
                      size_t errlen);
int orc_poly_from_proto(const uint8_t* bytes, size_t nbytes, int len,
                        int log_modulus, uint64_t* out, char* err,
                        size_t errlen);
/* values: n entries; is_signed selects BitPackSigned (entries are the
 * sign-extended values of a type of value_bits bits) or BitPackCoefficients
 * (value_bits = 64 for uint64 coefficients). Returns the byte count. */
size_t orc_bit_pack(const uint64_t* values, size_t n, int bits, int is_signed,
                    int value_bits, uint8_t* out);
void orc_bit_unpack(const uint8_t* data, size_t n, int bits, int is_signed,
                    int value_bits, uint64_t* out);

/* PirServer::SaveToFile's byte layout (pir_server.cc:433-592) for the combined
 * material (db: M x db_words condensed words; pack: int32 condensed words or
 * int16 entries; polys as in orc_process_response). out == NULL returns the
 * size only. */
size_t orc_save_server(const orc_params* P, const uint64_t* db_cdata,
                       int db_cols, const void* pack, const uint64_t* a_tilde,
                       const uint64_t* t_vec_h, const uint64_t* second_dim_a,
                       uint8_t* out);
/* PirServer::LoadFromBuffer (pir_server.cc:594-793) back into the combined
 * material; fails on a layout that does not match P. */
int orc_load_server(const orc_params* P, const uint8_t* buf, size_t size,
                    uint64_t* db_cdata, void* pack, uint64_t* a_tilde,
                    uint64_t* t_vec_h, uint64_t* second_dim_a, char* err,
                    size_t errlen);

/* ---- vPIR online path ----
 * VpirServer::ProcessResponse (crypto/vpir/server/vpir_server.cc:101-195):
 * db (ell*d x N condensed), combined pack (ell*d x k*d), a_tilde (ell x d),
 * t_vec_h (ell x k x d); first_dim: fd_polys x d; packing_key: pk_polys x d.
 * out: ell x d b polynomials. P->interp_degree must be 1. */
int orc_vpir_process(const orc_params* P, int local_finalize,
                     const uint64_t* db_cdata, int db_cols, const void* pack,
                     const uint64_t* a_tilde, const uint64_t* t_vec_h,
                     const uint64_t* first_dim, int fd_polys,
                     const uint64_t* packing_key, int pk_polys, uint64_t* out,
                     char* err, size_t errlen);
/* ProcessResponseWithMatMul (vpir_server.cc:197-287) with the H' matrix of
 * CreateHPrimeMatrix (:320-375): the t_vec_h inner product as an exact
 * negacyclic matrix product. */
int orc_vpir_process_matmul(const orc_params* P, int local_finalize,
                            const uint64_t* db_cdata, int db_cols,
                            const void* pack, const uint64_t* t_vec_h,
                            const uint64_t* first_dim, int fd_polys,
                            const uint64_t* packing_key, int pk_polys,
                            uint64_t* out, char* err, size_t errlen);

/* ---- Seeded generators (documented SplitMix64 stream) ---- */
void orc_fill_u64(uint64_t seed, uint64_t* out, size_t n, uint64_t mask);
void orc_fill_mod(uint64_t seed, uint64_t* out, size_t n, uint64_t modulus);
void orc_fill_db_u16(uint64_t seed, uint16_t* out, size_t n, uint64_t p);

#ifdef __cplusplus
}
#endif

#endif  /* PIR_ORACLE_H_ */
