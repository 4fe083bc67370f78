// Kernel launchers of the offline preprocessing (preprocess_kernels.hip and
// the FFT-side kernels in fft_kernels.hip), used by preprocess.cpp.
#ifndef INSPIRE_AMD_PREPROCESS_INTERNAL_H_
#define INSPIRE_AMD_PREPROCESS_INTERNAL_H_

#include "pirg_internal.h"

namespace pirg {

// One butterfly stage (length `len`) of Interpolate's inverse FFT over
// Z_p[X]/(X^d + 1) (crypto/db_row_interpolation.cc:183-210) on `groups` rows
// of t polynomials each. Element type: 1 or 2 bytes.
//   first: the source is the raw DB (record g*t + x of shard `shard`, at
//          raw + ((g*t + x)*ell + shard)*d, reduced mod p), read in the
//          bit-reversed polynomial order of :165-176; otherwise `src` holds
//          the previous stage, [g][t][d].
//   t_inv: non-zero on the last stage, which multiplies by t^-1 mod p.
struct InterpStage {
  const void* src;
  void* dst;  // [g][t][d]
  int elem_bytes;
  long long groups;
  int d, log_d, t, log_t, len;
  int first, ell, shard;
  uint32_t p, t_inv;
};
hipError_t LaunchInterpStage(const InterpStage& A, hipStream_t s);
// All stages at once for t in {2, 4, 8, 16} (first and t_inv as for the
// first and last stage above; len is unused).
hipError_t LaunchInterpFused(const InterpStage& A, hipStream_t s);

// dst[x][g] = src[g * src_stride + x] for x < cols, g < rows (zero for
// rows <= g < dst_cols): InterpolateAndTranspose's transpose
// (preprocessing_server.h:198-206) written straight into the condensed layout
// (row x of dst has dst_cols elements).
hipError_t LaunchTranspose(const void* src, long long src_stride, int rows, int cols,
                           int elem_bytes, void* dst, long long dst_cols, hipStream_t s);

// One stage of EvaluateNttInPlace (crypto/lwes_to_rlwe.cc:88-124) on `batches`
// arrays of d ring elements (each d coefficients, word type u32 or u64 by
// `word_bytes`): out[i+j] = in[i+j] + X^(2j*d/len) in[i+j+len/2],
// out[i+j+len/2] = in[i+j] - X^(2j*d/len) in[i+j+len/2]. Out of place.
hipError_t LaunchRingNttStage(const void* in, void* out, int word_bytes, long long batches,
                              int log_d, int len, hipStream_t s);

// PreprocessPack steps 1-3 before the transform (lwes_to_rlwe.cc:247-299):
// for block b (rows b*d .. b*d+d-1 of `hint`, d words each), element k of the
// NTT input is X^k * sum_j a_tilde_j[k] X^j with a_tilde_j the LowBits /
// RightShift(log2 d) / LowBits image of a_j read as sum_i a_j[i] X^-i. It is
// stored at position bit_rev(k) (the permutation EvaluateNttInPlace starts
// with): A[b][bit_rev(k)][r].
hipError_t LaunchPackNttInit(const u64* hint, int nblocks, int d, int log_d, int shift,
                             u64 qmask, u64* A, hipStream_t s);

// Chain update of PreprocessPack's reduction loop (:313-346): for chain
// c = 2*b + type, st[c] = A[b][type ? aidx_gh : aidx_g] + sum_p part[c][p]
// (wraparound, the reference's Polynomial::Add).
hipError_t LaunchChainCombine(const u64* A, int d, int aidx_g, int aidx_gh, const u64* part,
                              int npairs, int nchains, u64* st, hipStream_t s);
// a_tilde[b] = LowBits(st[2b] + sum_p part[b][p], log_q) (:355-361).
hipError_t LaunchChainFinal(const u64* st, const u64* part, int npairs, int nblocks, int d,
                            u64 qmask, u64* a_tilde, hipStream_t s);

// BuildPreprocessMatrix's output loop (:195-226) for nblocks x k transformed
// arrays T[b][l][c][r] (low 32 bits): entry (b*d + r, l*d + c) of the matrix
// is coefficient r of X^c * T[b][l][c], stored as int32 (mat_bytes 4) or
// int16 (mat_bytes 2), row-major with k*d entries per row.
hipError_t LaunchMatrixOut(const uint32_t* T, int nblocks, int k, int d, int mat_bytes,
                           void* out, hipStream_t s);

// out[x][m] = sum_q part[x][q][m] (wraparound) for x < rows, m < d.
hipError_t LaunchSumParts(const u64* part, int nparts, int d, long long rows, u64* out,
                          hipStream_t s);

// ---------------------------------------------------------------------------
// FFT side (fft_kernels.hip).

// Forward spectra of DB row polynomials for ComputeHintMatrix
// (hint_computation.cc:120-152): polynomial k of row x is columns
// [k*d, k*d + d) of row x of `db` (elements of elem_bytes, row stride
// `stride` elements), zero at columns >= valid_cols. dst[(x*polys + k)*nu + a]
// is the N-point spectrum of chunk a.
hipError_t LaunchDbSpectra(const void* db, int elem_bytes, long long stride, long long valid_cols,
                           int rows, int polys, const ChunkRule& rule, const FftDims& F,
                           const double2* tw_fwd, double2* dst, hipStream_t s);

// Spectra of one step of PreprocessPack's reduction loop (:313-346).
//   Digit jobs: for chain ch < n_dig (state row dig_first + ch * dig_step),
//   digit l < k of SignedGadgetInv(LowBits(state, log_q)) (polynomial.cc:376-427)
//   and chunk a of ur: uspec[(ch*k + l)*nu + a]. The digits are also stored:
//   as the low 32 bits at digits32[((chain/2)*k + l)*d + slot][.] (slot =
//   slot_g or slot_gh by chain parity) and/or in full at digits64[(ch*k + l)][.].
//   Automorphism jobs: for type ty < n_wtypes, w polynomial m < k and chunk b
//   of vr: vspec[(ty*k + m)*nv + b] = spectrum of Automorph(w[m], power_ty),
//   given as the inverse power pinv_ty mod 2d (polynomial.cc:429-456).
struct ChainSpectraArgs {
  const u64* state;
  int n_dig, dig_first, dig_step;
  int k, log_digit, to_shift;
  u64 qmask;
  ChunkRule ur;
  double2* uspec;
  uint32_t* digits32;
  int slot_g, slot_gh;
  u64* digits64;
  const u64* w;
  int n_wtypes;
  int pinv[2];
  ChunkRule vr;
  double2* vspec;
};
hipError_t LaunchChainSpectra(const ChainSpectraArgs& A, const FftDims& F,
                              const double2* tw_fwd, hipStream_t s);

}  // namespace pirg

#endif  // INSPIRE_AMD_PREPROCESS_INTERNAL_H_
