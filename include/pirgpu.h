/*
 * pirgpu — MI355X (gfx950) implementation of ReinsPIRe's online PIR answer
 * path, exposed as a plain C ABI.
 *
 * This header is the drop-in boundary. It replaces the server API of
 * google/InsPIRe's PirServer<uint16_t|uint8_t, uint64_t, int32_t|int16_t>
 * (crypto/pir/server/pir_server.h:41-111). Every entry point names the
 * reference declaration it stands for. No absl, C++ or torch types appear
 * here: only plain pointers, sizes and int status codes.
 *
 * Status codes are numerically equal to absl::StatusCode:
 *   0 OK, 3 INVALID_ARGUMENT, 8 RESOURCE_EXHAUSTED, 9 FAILED_PRECONDITION,
 *   13 INTERNAL, 14 UNAVAILABLE (no usable GPU).
 * Every function that can fail takes (char* err, size_t errlen) and writes the
 * same message text the reference returns for the same condition.
 *
 * Threading: a pirg_server is externally synchronised (like the reference,
 * whose ProcessResponse is not safe to call concurrently), except for the
 * serving-loop calls pirg_serve_submit(_batch, _records) / pirg_serve_poll
 * (_batch) / pirg_serve_stats / pirg_serve_stop, which are thread-safe and
 * may run beside each other and beside the other calls (pirg_server_destroy
 * excepted), and pirg_host_alloc / pirg_host_free / pirg_request_record,
 * which are thread-safe.
 */
#ifndef PIRGPU_H_
#define PIRGPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIRG_OK 0
#define PIRG_INVALID_ARGUMENT 3
#define PIRG_DEADLINE_EXCEEDED 4
#define PIRG_RESOURCE_EXHAUSTED 8
#define PIRG_FAILED_PRECONDITION 9
#define PIRG_INTERNAL 13
#define PIRG_UNAVAILABLE 14

/* PirParams<uint64_t> + RlweParams + the two GadgetParams
 * (crypto/pir/pir_params.h:36-105, crypto/encryption.h:29-98), with the
 * compile-time types of the reference's template instantiation made runtime:
 * db_bytes 2 = uint16_t DB, 1 = uint8_t DB; mat_bytes 4 = int32_t pack
 * matrix (condensed; a plain uint32_t matrix has the same bytes and, with
 * 32-bit coefficients, the same products), 2 = int16_t pack matrix (plain),
 * 8 = uint64_t pack matrix (plain; entries must lie in [-2^31, 2^31) as
 * two's complement, which BuildPreprocessMatrix's centering gives);
 * coeff_bits 64 (or 0)
 * = CoeffType uint64_t, 32 = uint32_t (pir_server.cc:806-818). Coefficients
 * cross the ABI as uint64_t either way; with uint32_t they are the 32-bit
 * values zero-extended, and the answer wraps where Polynomial<uint32_t>
 * does. log_q is LogModulus() (64, or 32 for uint32_t, means modulus 0,
 * i.e. Z_{2^width}). */
typedef struct {
  int log_degree;
  int log_q;
  int log_q1;
  int log_q2;
  int variance;
  uint64_t plaintext_modulus;
  int pack_log_digit, pack_num_digits;
  int eval_log_digit, eval_num_digits;
  int num_entries;
  int interp_degree;
  int entry_size_multiple;
  int db_bytes;
  int mat_bytes;
  int coeff_bits;
} pirg_params;

/* View of PirPreprocessedData<Db, uint64_t, Mat>
 * (crypto/pir/server/pir_preprocessed_data.h:38-50): the per-shard DB
 * matrices (condensed words, crypto/matrix.cc:76-160), the [shard][t]
 * PreprocessMatrixPackOutput entries (crypto/lwes_to_rlwe.h:42-47) and
 * second_dim_a. All pointers are host pointers, borrowed for the call. */
typedef struct {
  int num_db_matrices;               /* db_matrices.size() */
  const uint64_t* const* db_cdata;   /* [num_db_matrices] rows x words */
  const int* db_rows;                /* [num_db_matrices] Rows() */
  const int* db_cols;                /* [num_db_matrices] Cols() (original) */
  int num_outputs;                   /* preprocessed_outputs.size() */
  const int* outputs_len;            /* [num_outputs] preprocessed_outputs[k].size() */
  /* [num_outputs][t] (row-major over shard then i): */
  const void* const* pack;           /* d x k*d: int32 condensed words or int16 */
  const uint64_t* const* a_tilde;    /* d coefficients */
  const uint64_t* const* t_vec_h;    /* k x d coefficients */
  int second_dim_a_len;              /* second_dim_a.size() */
  const uint64_t* second_dim_a;      /* second_dim_a_len x d */
} pirg_preprocessed_view;

typedef struct pirg_server pirg_server;

/* PirServer::Create (crypto/pir/server/pir_server.cc:52-124). Validates with
 * the reference's messages, then moves the material into HBM in the
 * layouts the kernels stream. */
int pirg_server_create(const pirg_params* params,
                       const pirg_preprocessed_view* data,
                       pirg_server** out, char* err, size_t errlen);

/* The same, from the already-concatenated form PirServer keeps after Create
 * (combined_db_matrix_ / combined_pack_matrix_, pir_server.cc:100-118):
 *   db_cdata: M x (db_cols/4) words (uint16 DB) or M x (db_cols/8) (uint8),
 *             M = ell * t * d, shard-major;
 *   pack:     M rows, (shard, i)-major, each d x k*d (int32 condensed words
 *             or int16 values);
 *   a_tilde:  (ell*t) x d;  t_vec_h: (ell*t) x k x d;
 *   second_dim_a: 2e x d (may be NULL when t == 1). */
int pirg_server_create_flat(const pirg_params* params, const uint64_t* db_cdata,
                            int db_cols, const void* pack,
                            const uint64_t* a_tilde, const uint64_t* t_vec_h,
                            const uint64_t* second_dim_a, pirg_server** out,
                            char* err, size_t errlen);

void pirg_server_destroy(pirg_server* server);

/* PirServer::ProcessResponse (crypto/pir/server/pir_server.cc:152-306).
 * Request = PirRequest<uint64_t> (crypto/pir/pir_params.h:196-208):
 *   first_dim[first_dim_len], second_dim[second_dim_polys x d],
 *   packing_key[packing_key_polys x d].
 * Response = PirResponse<uint64_t> (pir_params.h:211-216) written to
 *   out[ell][2][d] as (a_0, b_0, a_1, b_1, ...). */
int pirg_process(pirg_server* server, const uint64_t* first_dim,
                 size_t first_dim_len, const uint64_t* second_dim,
                 size_t second_dim_polys, const uint64_t* packing_key,
                 size_t packing_key_polys, uint64_t* out, char* err,
                 size_t errlen);

/* One PirRequest, borrowed. */
typedef struct {
  const uint64_t* first_dim;
  size_t first_dim_len;
  const uint64_t* second_dim;
  size_t second_dim_polys;
  const uint64_t* packing_key;
  size_t packing_key_polys;
} pirg_request_view;

/* Many PirServer::ProcessResponse calls answered by one batched pass: the DB
 * and pack products become int8-limb MFMA matrix products over the batch.
 * statuses[i] receives request i's status (its validation mirrors
 * ProcessResponse); outs[i] (ell x 2 x d) is written when statuses[i] == 0.
 * No reference counterpart: the reference answers one request per call. */
int pirg_process_batch(pirg_server* server, int nreq,
                       const pirg_request_view* reqs, uint64_t* const* outs,
                       int* statuses, char* err, size_t errlen);

/* Device-resident batch, for callers that keep requests in HBM:
 *   pirg_batch_stage copies nreq requests into the server's HBM workspace;
 *   pirg_batch_run answers the staged batch (all inputs already in HBM);
 *   pirg_batch_fetch copies the nreq responses back.
 * All requests of a staged batch must have first_dim_len == r (InterpolatedRows)
 * and the full second_dim / packing_key; pirg_batch_stage validates them. */
int pirg_batch_stage(pirg_server* server, int nreq,
                     const pirg_request_view* reqs, char* err, size_t errlen);
int pirg_batch_run(pirg_server* server, int nreq, char* err, size_t errlen);
int pirg_batch_fetch(pirg_server* server, int nreq, uint64_t* const* outs,
                     char* err, size_t errlen);

/* Autotuner: for each batch size in widths, times every candidate
 * configuration of the DB and pack scans (narrow scan with 1/2/4 tiles per
 * group, wide scan with 2/4/7 tiles per wave, two K-splits) on the resident
 * matrices and keeps the fastest for that batch size; other sizes keep the
 * built-in choice. `report` (may be NULL) receives one line per scan and
 * width with every candidate's time. Not while serving. No reference
 * counterpart. */
int pirg_server_autotune(pirg_server* server, const int* widths, int nwidths,
                         char* report, size_t report_len);

/* Whether pirg_process / pirg_process_batch / pirg_batch_run replay one
 * hipGraph per batch size (captured on first use: the whole answer pipeline
 * as one graph launch) or launch the kernels one by one. On by default; both
 * add to the stage times below (graphs carry event-record nodes at the stage
 * boundaries). No reference counterpart. */
int pirg_server_use_graphs(pirg_server* server, int enable);

/* Graphs captured from now on bracket the DB and pack scans with event-record
 * nodes, so that graph replays (not only eager runs) add the scan-kernel
 * times that pirg_server_scan_times reports. Changing the mode drops the
 * graphs captured so far. Off by default. No reference counterpart. */
int pirg_server_time_graph_scans(pirg_server* server, int enable);

/* ---- Serving loop (no reference counterpart: the reference answers one
 * request per call). The DB and preprocessed material stay resident; a
 * worker thread keeps `slots` execution slots busy, each answering up to
 * `max_batch` requests per graph launch, and every request waiting when a
 * slot frees joins that slot's next batch (continuous batching). Batches run
 * at fixed sizes (powers of two up to 16, then multiples of 16, padded),
 * whose graphs are captured at start; a batch with an MFMA-bound DB scan
 * runs its kernels after the previous batch's, its copies overlapping them,
 * and smaller batches overlap freely. Slots and batch capacity are reduced
 * at start to fit 80% of the free HBM, or the limit set with
 * pirg_serve_set_memory_limit when lower; start fails with
 * PIRG_RESOURCE_EXHAUSTED when not even one batch fits. A submit is refused
 * with PIRG_RESOURCE_EXHAUSTED while max_outstanding requests are submitted
 * but not yet polled. */
/* The device memory a later pirg_serve_start may give its slots (0, the
 * default: 80% of the free memory at start), and the bytes one slot takes
 * for batches of `batch` requests. */
int pirg_serve_set_memory_limit(pirg_server* server, size_t bytes);
size_t pirg_serve_workspace_bytes(const pirg_server* server, int batch);
int pirg_serve_start(pirg_server* server, int max_batch, int slots,
                     int max_outstanding, char* err, size_t errlen);
/* Validates like ProcessResponse (same messages) and queues the request; the
 * buffers are copied before return. Thread-safe. */
int pirg_serve_submit(pirg_server* server, const pirg_request_view* req,
                      uint64_t tag, char* err, size_t errlen);
/* Waits up to timeout_ms for a finished request: returns 0 and fills tag,
 * status (its ProcessResponse status), latency_ms (submit to completion) and
 * out (ell x 2 x d, when status == 0); returns PIRG_DEADLINE_EXCEEDED when
 * none finished in time, and PIRG_FAILED_PRECONDITION when the loop is not
 * running or was stopped while the call waited. Thread-safe. */
int pirg_serve_poll(pirg_server* server, int timeout_ms, uint64_t* tag,
                    uint64_t* out, int* status, double* latency_ms, char* err,
                    size_t errlen);
/* n submits in one call (tags[i] for reqs[i]): statuses[i] receives each
 * request's validation status (0 = queued; err holds the first failure's
 * message). The valid requests are copied (by several threads for large
 * calls) and queued together; when they do not all fit under
 * max_outstanding, none is queued and PIRG_RESOURCE_EXHAUSTED is returned.
 * Thread-safe. */
int pirg_serve_submit_batch(pirg_server* server, int n, const pirg_request_view* reqs,
                            const uint64_t* tags, int* statuses, char* err, size_t errlen);
/* Up to max finished requests in one call: waits up to timeout_ms for the
 * first, then takes whatever else is finished. *count receives how many;
 * for each i < *count, tags[i], statuses[i], latency_ms[i] (may be NULL) and,
 * when statuses[i] == 0, outs[i] (ell x 2 x d) are filled. Return codes as
 * pirg_serve_poll. Thread-safe. */
int pirg_serve_poll_batch(pirg_server* server, int max, int timeout_ms, uint64_t* tags,
                          uint64_t* const* outs, int* statuses, double* latency_ms, int* count,
                          char* err, size_t errlen);
/* Zero-copy submission (no counterpart in the reference: a serving form of
 * PirServer::ProcessResponse, crypto/pir/server/pir_server.cc:152-306, for
 * callers that receive requests straight into pinned host memory).
 *
 * A request record holds one request in pirg_request_record_words(server)
 * words: the first dimension zero-padded to the DB's padded column count,
 * the second dimension (t > 1), the packing key. pirg_request_record
 * validates a request as ProcessResponse does (pir_server.cc:161-194) and
 * writes its record. Records in memory from pirg_host_alloc go to
 * pirg_serve_submit_records, which queues them without a host copy (records
 * adjacent in memory travel to HBM as one copy); each must stay allocated
 * and unchanged until its response is polled. pirg_serve_submit_records
 * returns PIRG_INVALID_ARGUMENT for a record outside such memory and
 * PIRG_RESOURCE_EXHAUSTED past max_outstanding, queuing none; responses come
 * back through pirg_serve_poll / pirg_serve_poll_batch. pirg_host_free takes
 * what pirg_host_alloc returned. Thread-safe. */
int pirg_host_alloc(size_t bytes, void** out, char* err, size_t errlen);
void pirg_host_free(void* p);
size_t pirg_request_record_words(const pirg_server* server);
int pirg_request_record(const pirg_server* server, const pirg_request_view* req,
                        uint64_t* record, char* err, size_t errlen);
int pirg_serve_submit_records(pirg_server* server, int n, const uint64_t* const* records,
                              const uint64_t* tags, char* err, size_t errlen);
typedef struct {
  long long submitted, answered, batches;
  double mean_batch;            /* requests per launched batch */
  double p50_ms, p99_ms, max_ms; /* submit-to-completion latency, over the
                                    latest 2^20 answers */
  double elapsed_ms;            /* since start or the last reset */
  int slots, max_batch;         /* after admission by free HBM */
} pirg_serve_stats_t;
int pirg_serve_stats(pirg_server* server, pirg_serve_stats_t* stats, int reset);
/* Answers everything submitted, then stops the worker; threads waiting in
 * pirg_serve_poll return (with a result still undelivered, or
 * PIRG_FAILED_PRECONDITION), and results nobody polls are dropped. Calls
 * already inside the loop finish safely: the loop is freed when the last of
 * them returns. pirg_server_destroy stops a running loop itself. */
int pirg_serve_stop(pirg_server* server);

/* ---- Batched LWE-to-RLWE packing. MatrixPack (crypto/lwes_to_rlwe.cc:481-508,
 * declared at crypto/lwes_to_rlwe.h:126-133) for njobs independent jobs,
 * each with its own PreprocessMatrixPackOutput (lwes_to_rlwe.h:42-47), made
 * resident in HBM:
 *   pack[j]:    d x k*d matrix (int32 condensed words, or int16 values when
 *               params->mat_bytes == 2);
 *   a_tilde[j]: a_tilde_agg, d coefficients;  t_vec_h[j]: k x d.
 * params supplies log_degree, log_q, pack_log_digit/num_digits, mat_bytes.
 * The reference packs exactly d LWEs per call (lwes_to_rlwe.cc:428); a job
 * with m < d LWEs passes its m b-values and is zero-padded. */
typedef struct pirg_packer pirg_packer;
int pirg_packer_create(const pirg_params* params, int njobs, const void* const* pack,
                       const uint64_t* const* a_tilde, const uint64_t* const* t_vec_h,
                       pirg_packer** out, char* err, size_t errlen);
/* One MatrixPack per job, all jobs in one launch (a hipGraph, captured on the
 * first call): b[j] holds m[j] <= d LWE b-values, keys[j] the packing key
 * y_vec_g || y_vec_h (2k x d; jobs may share one key pointer), out[j]
 * receives the RLWE ciphertext (a, b) as 2 x d. Host pointers. */
int pirg_packer_run(pirg_packer* packer, const uint64_t* const* b, const int* m,
                    const uint64_t* const* keys, uint64_t* const* out, char* err,
                    size_t errlen);
/* Graph replay (default) or eager launches, which also time the scan. */
int pirg_packer_use_graph(pirg_packer* packer, int enable);
/* Cumulative batch time and (eager runs) pack-scan kernel time, in ms. */
int pirg_packer_times(pirg_packer* packer, double* total_ms, double* scan_ms,
                      long long* runs, int reset);
void pirg_packer_destroy(pirg_packer* packer);

/* PirServer::GetDbMultiplyTime / GetPackingTime / GetPolyEvalTime
 * (pir_server.h:70-72): cumulative GPU time, in milliseconds, of the DB
 * product, the packing (pack product + FinalizeMatrixPack) and EvalPoly
 * stages over every call so far. */
int pirg_server_stage_times(pirg_server* server, double* db_mult_ms,
                            double* packing_ms, double* poly_eval_ms);

/* HIP-event durations of the int8-limb scan kernels alone (DB product and
 * pack product), summed over `launches` batches since the last reset. No
 * reference counterpart: used to place the scans against the HBM and MFMA
 * rooflines. */
int pirg_server_scan_times(pirg_server* server, double* db_scan_ms,
                           double* pack_scan_ms, long long* launches,
                           int reset);

/* *fused = 1 when a batch of `batch` requests runs the DB and pack scans as
 * one kernel launch (batches whose DB scan takes a few rounds of workgroups:
 * the pack product's workgroups fill the CUs the DB product's last round
 * leaves idle); pirg_server_scan_times then reports that launch's time as the
 * DB scan's and none for the pack scan. 0 when they are two launches. No
 * reference counterpart. */
int pirg_server_scans_fused(pirg_server* server, int batch, int* fused);

/* Largest |x - round(x)| over every value the FFT products rounded since the
 * last reset, the precision margin of the double-precision products. */
double pirg_server_fft_margin(pirg_server* server, int reset);

/* Matrix<uint16_t>::Multiply<uint64_t> on a condensed matrix
 * (crypto/matrix.h:165-195 -> crypto/matrix.cc:261-318), full 64-bit result,
 * computed by the same int8-limb kernel as the DB product. Host pointers. */
int pirg_matrix_multiply_u16(const uint64_t* cdata, int rows, int cols,
                             const uint64_t* vec, uint64_t* out, char* err,
                             size_t errlen);
/* Matrix<int32_t>::Multiply<uint64_t> on a condensed matrix
 * (crypto/matrix.cc:320-374). */
int pirg_matrix_multiply_i32(const uint64_t* cdata, int rows, int cols,
                             const uint64_t* vec, uint64_t* out, char* err,
                             size_t errlen);

/* Polynomial::InnerProductFft (crypto/polynomial_fft.cc:68-197): n
 * polynomials u, v of length d (n x d each), on the GPU with the reference's
 * FFT arithmetic. */
int pirg_inner_product_fft(const uint64_t* u, const uint64_t* v, int n, int d,
                           int u_bits, int v_bits, int chunk_bits,
                           int u_is_signed, uint64_t* out, char* err,
                           size_t errlen);

/* The same product on device memory, with no device allocation: u, v (n x d)
 * and out (d) are device pointers, and the work is queued on `stream` (a
 * hipStream_t; NULL for the default stream) without waiting for it. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream capture, serving loops that forbid
 * hipMalloc). Its spectra and partial results live in one static 32 MiB
 * device array per GPU, so n * (u chunks + v chunks) * 2d * 16 +
 * pairs * d * 8 bytes may not exceed about 31.5 MiB (INVALID_ARGUMENT
 * otherwise); calls that overlap on one GPU (other streams, other threads)
 * race on that array. The first call on each device uploads the twiddle
 * tables and may not run under stream capture (FAILED_PRECONDITION). Errors
 * and messages otherwise as pirg_inner_product_fft. */
int pirg_inner_product_fft_noalloc(const uint64_t* u, const uint64_t* v, int n, int d,
                                   int u_bits, int v_bits, int chunk_bits, int u_is_signed,
                                   uint64_t* out, void* stream, char* err, size_t errlen);

/* The twiddle table the FFT kernels use for an n-point transform, as the
 * interleaved (re, im) pairs of FftPlan's twiddles_forward_ /
 * twiddles_backward_ (crypto/fft.h:71-80); n/2 pairs each. Host only. */
int pirg_fft_twiddles(int n, double* fwd, double* bwd);

/* Version / device info string (for logs). */
/* ---- Bit-packed formats ----
 * Polynomial::ToProto (crypto/polynomial.cc:562-590) for npolys polynomials
 * of len coefficients each (coeffs: npolys x len). out receives each
 * polynomial's encoded_coeffs back to back, ceil(len * log_modulus / 8)
 * bytes apiece: the low log_modulus bits of every coefficient, LSB first.
 * Errors: "Invalid log_modulus." (log_modulus outside [1, 64]). */
int pirg_poly_to_proto(const uint64_t* coeffs, int npolys, int len, int log_modulus,
                       uint8_t* out, char* err, size_t errlen);
/* Polynomial::CreateFromProto (crypto/polynomial.cc:208-252) for npolys
 * byte strings of poly_bytes bytes each, with the reference's checks and
 * messages: "Invalid length.", "Invalid log_modulus.", "Empty serialized
 * polynomial.", "Unexpected serialized size.". out: npolys x len. */
int pirg_poly_from_proto(const uint8_t* bytes, size_t poly_bytes, int npolys, int len,
                         int log_modulus, uint64_t* out, char* err, size_t errlen);
/* The same two conversions on device memory, queued on `stream` (a
 * hipStream_t; NULL for the default stream). A polynomial's stream is
 * ceil(len * log_modulus / 64) 64-bit words at a pitch of pitch_words words
 * (bytes in the little-endian order of ToProto's string); words past a
 * stream's end are written as zero by the encoder and ignored by the
 * decoder. No counterpart in the reference: the device forms are for callers
 * that keep requests and responses in HBM. */
int pirg_poly_to_proto_device(const uint64_t* d_coeffs, int npolys, int len, int log_modulus,
                              uint64_t* d_words, long long pitch_words, void* stream,
                              char* err, size_t errlen);
int pirg_poly_from_proto_device(const uint64_t* d_words, long long pitch_words, int npolys,
                                int len, int log_modulus, uint64_t* d_coeffs, void* stream,
                                char* err, size_t errlen);

/* PirServer::LoadFromBuffer (crypto/pir/server/pir_server.cc:594-793): builds
 * a server from the byte image PirServer::SaveToFile writes (:433-592) --
 * combined DB matrix (raw condensed words), combined pack matrix (signed
 * entries bit-packed at pack_log_digit + log2(d) bits), the preprocessed
 * outputs and second_dim_a (bit-packed at log q bits). The pack matrix and
 * the polynomials are unpacked on the GPU straight into the resident
 * layouts. The image's shapes must match params (INVALID_ARGUMENT otherwise);
 * a record cut short fails with "Error reading from buffer" (INTERNAL), as
 * in the reference. */
int pirg_server_load_from_buffer(const pirg_params* params, const uint8_t* buffer,
                                 size_t size, pirg_server** out, char* err, size_t errlen);

/* PirServer::LoadFromFile (pir_server.cc:794-804): LoadFromBuffer on the
 * file's bytes (mapped, not copied). An unreadable file fails with "Failed to
 * read file." (INTERNAL), as in the reference. */
int pirg_server_load_from_file(const pirg_params* params, const char* file_path,
                               pirg_server** out, char* err, size_t errlen);

/* PirServer::SaveToFile (pir_server.cc:433-592): the image LoadFromBuffer
 * reads, byte for byte as the reference writes it for the same material,
 * written from the resident server (DB words and pack entries rebuilt from
 * their HBM tiles and bit-packed on the GPU). Errors: "Failed to open file
 * for writing: <path>", "Error writing to file: <path>" (INTERNAL). A
 * server given a plain uint32_t pack matrix under mat_bytes 4 is written
 * with the condensed int32_t header (5 bytes differ from the reference's
 * Matrix<uint32_t> image; the entries are the same). */
int pirg_server_save_to_file(pirg_server* server, const char* file_path, char* err,
                             size_t errlen);
/* The same image into a caller's buffer: *size receives its length; with
 * buffer == NULL nothing else happens (no device work), otherwise capacity
 * must be at least that length (INVALID_ARGUMENT). No reference counterpart:
 * for callers that keep the image in memory. */
int pirg_server_save_to_buffer(pirg_server* server, uint8_t* buffer, size_t capacity,
                               size_t* size, char* err, size_t errlen);

/* DeserializePirPreprocessedData + PirServer::Create
 * (crypto/pir/server/pir_preprocessed_data.h:93-148, pir_server.cc:52-124) on
 * the bytes of a serialized proto::PirPreprocessedData
 * (crypto/proto/pir_server.proto:54-66), i.e. the contents of the file
 * SavePirPreprocessedDataToFile writes. Matrices are taken as stored;
 * polynomials are decoded on the GPU. Errors: "Failed to parse
 * PirPreprocessedData from file.", Matrix::FromProto's and CreateFromProto's
 * messages, then Create's. */
int pirg_server_create_from_proto(const pirg_params* params, const uint8_t* bytes, size_t size,
                                  pirg_server** out, char* err, size_t errlen);
/* LoadPirPreprocessedDataFromDirectory + PirServer::Create
 * (pir_preprocessed_data.h:226-320): metadata.bin and shard_<i>.bin under
 * dir_path, for i < entry_size_multiple. */
int pirg_server_create_from_directory(const pirg_params* params, const char* dir_path,
                                      pirg_server** out, char* err, size_t errlen);

/* ---- vPIR online path (crypto/vpir/server/vpir_server.h:47-121) ----
 * VpirServer's answer is PirServer's with one interpolation slice (t = 1)
 * and no modulus switch: b = FinalizeMatrixPack(DB*q, pack*y_g, y_h) per
 * shard, or LowBits(DB*q + pack*y_g) with local finalize. The proof (z-PIR)
 * material is offline and not part of this path. */
typedef struct pirg_vpir_server pirg_vpir_server;

/* VpirServer::Create (vpir_server.cc:45-99) on the online material: db
 * (ell*d rows x N columns, condensed words, crypto/matrix.cc:76-160), the
 * combined pack matrix (ell*d x k*d: int32 condensed words or int16 entries),
 * and the num_outputs = ell preprocessed outputs' a_tilde_agg (ell x d) and
 * t_vec_h (ell x k x d). params: interp_degree must be 1; log_q1, log_q2 and
 * the eval gadget are unused. local_finalize: VpirParams::LocalFinalize.
 * Errors (vpir_server.cc:58-69): "DB matrix must have ell * d rows.",
 * "DB matrix must have N columns.", "preprocessed_outputs size mismatch." */
int pirg_vpir_server_create(const pirg_params* params, int local_finalize,
                            const uint64_t* db_cdata, int db_rows, int db_cols,
                            const void* pack, int num_outputs, const uint64_t* a_tilde,
                            const uint64_t* t_vec_h, pirg_vpir_server** out, char* err,
                            size_t errlen);
void pirg_vpir_server_destroy(pirg_vpir_server* server);

/* VpirRequest (crypto/vpir/vpir_params.h:160-164): first_dimension_query as
 * first_dim_polys polynomials of d coefficients (contiguous), packing_key as
 * packing_key_polys polynomials (k = y_vec_g, then k of y_vec_h unless local
 * finalize). */
typedef struct {
  const uint64_t* first_dim;
  size_t first_dim_polys;
  const uint64_t* packing_key;
  size_t packing_key_polys;
} pirg_vpir_request_view;

/* VpirServer::ProcessResponse (vpir_server.cc:101-195): out receives the ell
 * b_responses (ell x d). Errors, in the reference's order: "Packing key must
 * contain <n> polynomials.", "Flattened first dimension query must match DB
 * matrix columns (N)." */
int pirg_vpir_process(pirg_vpir_server* server, const pirg_vpir_request_view* req,
                      uint64_t* out, char* err, size_t errlen);
/* Many requests in one batched pass (per-request statuses, as
 * pirg_process_batch). No reference counterpart. */
int pirg_vpir_process_batch(pirg_vpir_server* server, int nreq,
                            const pirg_vpir_request_view* reqs, uint64_t* const* outs,
                            int* statuses, char* err, size_t errlen);
/* Device-resident batches (as pirg_batch_stage/run/fetch) and the scan
 * autotuner (as pirg_server_autotune) of the underlying server. */
int pirg_vpir_batch_stage(pirg_vpir_server* server, int nreq,
                          const pirg_vpir_request_view* reqs, char* err, size_t errlen);
int pirg_vpir_batch_run(pirg_vpir_server* server, int nreq, char* err, size_t errlen);
int pirg_vpir_batch_fetch(pirg_vpir_server* server, int nreq, uint64_t* const* outs,
                          char* err, size_t errlen);
int pirg_vpir_autotune(pirg_vpir_server* server, const int* widths, int nwidths,
                       char* report, size_t report_len, char* err, size_t errlen);
/* As pirg_server_use_graphs / pirg_server_time_graph_scans /
 * pirg_server_scan_times, for the underlying server. */
int pirg_vpir_use_graphs(pirg_vpir_server* server, int enable);
int pirg_vpir_time_graph_scans(pirg_vpir_server* server, int enable);
int pirg_vpir_scan_times(pirg_vpir_server* server, double* db_scan_ms, double* pack_scan_ms,
                         long long* launches, int reset);

/* ---- Offline preprocessing (crypto/pir/server/preprocessing_server.h:68-273) ----
 * The reference's PreprocessingServer samples the first-dimension, second-
 * dimension and packing a-components from its ChaCha PRNG (:98-111); here the
 * caller passes them. Everything else is computed on the GPU, bit-identical to
 * the reference. */

/* Interpolate (crypto/db_row_interpolation.cc:115-228) of nrows rows of t*d
 * DB values each (uint16 or uint8 per db_bytes, host memory); out has the
 * same shape. t == 1 copies the rows. Errors as the reference: "ring_degree
 * must be a power of two.", "interpolation_degree must be a power of two.",
 * "interpolation_degree must be less than or equal to 2 * ring_degree.",
 * "plaintext_modulus must be odd." */
int pirg_interpolate(const void* rows, int db_bytes, long long nrows, int d, int t,
                     uint32_t plaintext_modulus, void* out, char* err, size_t errlen);

/* ComputeHintMatrix<uint64_t, S, uint64_t> (crypto/hint_computation.cc:78-160):
 * db is rows x cols (S = uint16 / uint8 per db_bytes), a_components n x d,
 * out rows x d (host memory). Errors: "Database matrix cannot be empty.",
 * "a_components cannot be empty.", "Database columns must match d * n where d
 * is the polynomial length and n is the number of polynomials." */
int pirg_compute_hint_matrix(const void* db, int db_bytes, int rows, int cols,
                             const uint64_t* a_components, int n, int d, uint64_t* out,
                             char* err, size_t errlen);

/* PreprocessMatrixPack (crypto/lwes_to_rlwe.cc:421-446) for nblocks
 * independent blocks at once: block b's LWE a-vectors are rows
 * [b*d, b*d + d) of a_is (d words each); w is w_vec_g then w_vec_h (2k x d).
 * Outputs: pack rows [b*d, b*d + d) (d x k*d, int32 condensed words or int16
 * values per mat_bytes), a_tilde nblocks x d, t_vec_h nblocks x k x d. */
int pirg_preprocess_matrix_pack(const pirg_params* params, int nblocks, const uint64_t* a_is,
                                const uint64_t* w, void* pack, uint64_t* a_tilde,
                                uint64_t* t_vec_h, char* err, size_t errlen);

/* PreprocessingServer::Preprocess: db_raw holds num_entries * d * ell values
 * (record-major, uint16 / uint8), first_dim_a ceil((n/t)/d) x d, pack_a
 * 2k x d (w_vec_g then w_vec_h). Outputs (host, each may be NULL) in the
 * layouts pirg_server_create_flat takes: db_cdata M x K/4 (or K/8) words with
 * *db_cols = K, pack, a_tilde, t_vec_h. stage_ms (4 entries, may be NULL)
 * receives the interpolation, hint, pack-preprocessing and total GPU times
 * (GetInterpolateTime / GetComputeHintsTime / GetPreprocessPackTime,
 * preprocessing_server.h:165-168), from the moment the raw DB is in device
 * memory. */
int pirg_preprocess(const pirg_params* params, const void* db_raw, const uint64_t* first_dim_a,
                    const uint64_t* pack_a, uint64_t* db_cdata, int* db_cols, void* pack,
                    uint64_t* a_tilde, uint64_t* t_vec_h, double* stage_ms, char* err,
                    size_t errlen);

/* Preprocess, then PirServer::Create on the device-resident result (no host
 * round trip of the material). second_dim_a: 2e x d, or NULL when t == 1. */
int pirg_server_create_preprocessed(const pirg_params* params, const void* db_raw,
                                    const uint64_t* first_dim_a, const uint64_t* second_dim_a,
                                    const uint64_t* pack_a, double* stage_ms, pirg_server** out,
                                    char* err, size_t errlen);

const char* pirg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PIRGPU_H_ */
