// Internal declarations shared by the host orchestration (server.cpp) and the
// HIP kernels (limb_kernels.hip, fft_kernels.hip) of libpirgpu.
#ifndef INSPIRE_AMD_PIRG_INTERNAL_H_
#define INSPIRE_AMD_PIRG_INTERNAL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace pirg {

using u64 = uint64_t;
using i64 = int64_t;

// ---------------------------------------------------------------------------
// Chunk extraction of Polynomial::InnerProductFft
// (crypto/polynomial_fft.cc:98-152): chunk c of x is (x >> bits*c) & mask;
// the top chunk of a signed operand is sign-extended from its remaining bits.
struct ChunkRule {
  int chunk_bits;
  int num_chunks;
  int bits;        // u_bits / v_bits argument
  int is_signed;
};

__host__ __device__ inline double ChunkValue(const ChunkRule& r, u64 x, int c) {
  const u64 mask = (1ULL << r.chunk_bits) - 1;
  const u64 chunk = (x >> (r.chunk_bits * c)) & mask;
  if (r.is_signed && c == r.num_chunks - 1) {
    const int bits_in_chunk = r.bits - r.chunk_bits * c;
    const int shift = 64 - bits_in_chunk;
    const i64 sc = static_cast<i64>(chunk << shift) >> shift;
    return static_cast<double>(sc);
  }
  return static_cast<double>(chunk);
}

// The (a, b) chunk pairs InnerProductFft accumulates (it skips pairs with
// chunk_bits*(a+b) >= 8*sizeof(CoeffType), polynomial_fft.cc:154-173).
struct PairList {
  int count;
  int a[64];
  int b[64];
  int shift[64];
};

inline PairList MakePairs(const ChunkRule& u, const ChunkRule& v, int width = 64) {
  PairList p{};
  for (int a = 0; a < u.num_chunks; ++a) {
    for (int b = 0; b < v.num_chunks; ++b) {
      if (u.chunk_bits * (a + b) >= width) continue;
      p.a[p.count] = a;
      p.b[p.count] = b;
      p.shift[p.count] = u.chunk_bits * (a + b);
      ++p.count;
    }
  }
  return p;
}

// ---------------------------------------------------------------------------
// Int8-limb operand layout.
//
// A matrix A (rows x cols) of small integers is stored as NP int8 "planes":
// A = C_off + sum_p 2^(8p) plane_p, so that A * y = sum_p 2^(8p) (plane_p * y)
// + C_off * sum(y). Each plane is tiled for v_mfma_i32_16x16x64_i8: a tile is
// 16 rows x 64 columns = 64 lanes x 16 bytes; lane l holds row (l % 16),
// columns 16*(l/16) .. +15. Tiles are ordered [strip][kblock][plane].
// A vector y (uint64) is split into nl balanced int8 digits
// y = sum_c 2^(8c) digit_c (mod 2^(8 nl)); the B operand tile for k-block kb
// holds 16 columns = (query, digit) pairs, qpt queries per tile.
struct LimbLayout {
  int rows;      // logical rows
  int cols;      // logical columns
  int strips;    // ceil(rows / 16)
  int kblocks;   // ceil(cols / 64)
  int planes;    // NP
  u64 c_off;     // constant offset of the plane decomposition
};

constexpr int kTileBytes = 1024;

// Tiles of 16 consecutive strips are stored together per k-block:
// tile (strip, kb, p) is at index ((strip / 16 * kblocks + kb) * 16 +
// strip % 16) * planes + p. A workgroup's strips for one k-block are then one
// contiguous run, so the scans stream HBM sequentially (measured 1.24x over a
// strip-major layout for the single-query scan).
constexpr int kStripGroup = 16;

__host__ __device__ inline long long TileIndex(long long strip, long long kb, int p,
                                               long long kblocks, int planes) {
  return ((strip / kStripGroup * kblocks + kb) * kStripGroup + strip % kStripGroup) * planes + p;
}

// Bytes of a tile buffer (strips rounded up to whole groups).
inline size_t TileBufferBytes(const LimbLayout& L) {
  const size_t strips = (static_cast<size_t>(L.strips) + kStripGroup - 1) / kStripGroup * kStripGroup;
  return strips * L.kblocks * L.planes * kTileBytes;
}

// ---------------------------------------------------------------------------
// Kernel launchers (limb_kernels.hip).

// Converts condensed uint16 / uint8 DB words (rows x cols/4 or cols/8, device
// memory) for rows [row0, row0 + nrows) into planes. row0 must be a multiple
// of 16.
hipError_t LaunchDbToTiles(const u64* cdata, int db_bytes, int rows, int cols,
                           int row0, int nrows, const LimbLayout& L,
                           int8_t* tiles, hipStream_t s);
// Converts pack rows (int32 condensed words or int16 values, device memory).
hipError_t LaunchPackToTiles(const void* pack, int mat_bytes, int rows,
                             int cols, int row0, int nrows,
                             const LimbLayout& L, int8_t* tiles,
                             hipStream_t s);
// The inverse for SaveToFile: rows [row0, row0 + nrows) (row0 + nrows <=
// L.rows) back to row-major values, nrows x L.cols u64 (the planes' sum;
// sign-extended when signed_top, as for pack matrices; ANDed with mask).
hipError_t LaunchTilesToValues(const int8_t* tiles, const LimbLayout& L, int row0, int nrows,
                              bool signed_top, u64 mask, u64* out, hipStream_t s);
// Copies `words` u64 from src to dst with a kernel: either side may be
// pinned host memory, so a small transfer runs in line with the kernels
// around it instead of on a copy engine that must synchronise with them.
hipError_t LaunchCopyWords(const u64* src, u64* dst, size_t words, hipStream_t s);
// How one limb product over nq query vectors is launched. Columns of the
// int8 product are (query, shift s) pairs, s < ns = ceil(out_bits / 8): column
// (q, s) collects plane p times digit s - p of query q. Queries are grouped,
// qpg per group of ctg 16-column tiles; a wave of the scan owns one group.
struct ScanPlan {
  int wide;          // 0: NarrowScan (A-stream bound), 1: WideScan (MFMA bound)
  int nq;            // query vectors
  int ns;            // shifts per query
  int ctg;           // tiles per group
  int qpg;           // queries per group
  // B column order. 0 (query-major): column j of a group is (query j / ns,
  // shift j % ns) and plane p's operand is plane 0's shifted p columns.
  // 1 (digit-major, WideScan with ctg == ns only): tile s of a group holds
  // digit s of its 16 queries, so plane p's operand for output tile s is
  // plane 0's tile s - p, unchanged.
  int digit_major;
  // Digit-major with 7 digits: 1 runs the streaming schedule with its widest
  // tile (two A planes: WideScanTallKernel, 6 strips per wave, 12 per
  // workgroup; four: WideScanStreamKernel, 5 and 10, two-stage rings) and 2,
  // for four planes, WideScanStreamKernel with 4 strips per wave and a
  // three-stage A ring; 0 the pipelined (two planes) or plain WideScan kernel
  // with 4 and 8.
  int tall;
  int ngp;           // groups in the B layout (padded for WideScan)
  int ksplit;        // k-ranges (exact int32 sums and occupancy)
  int kb_per_split;
  size_t btile_bytes;
  // Grouped product (NarrowScan only): row block b of group_rows rows is
  // multiplied by query b alone, and partial is [ksplit][rows]. 0: every row
  // by every query.
  int group_rows;
};

// wide_min_queries: batches of at least this many queries use WideScan even
// when they would fit NarrowScan. allow_tall: digit-major scans of two-plane
// matrices use the tall kernel; four_plane_stream: the `tall` value of
// digit-major scans of four-plane matrices (0: the plain WideScan; 7 digit
// tiles: 1 and 2 the streaming scans with 10 and 8 strips per workgroup; 4
// digit tiles: 1 and 2 those with 6 and 8 strips and two query groups per
// wave; -1: the default for the shape and batch size).
ScanPlan MakeScanPlan(const LimbLayout& L, int nq, int out_bits, int wide_min_queries,
                      bool allow_digit_major = true, bool allow_tall = true,
                      int four_plane_stream = -1);
// The configurations an autotuner may time for nq queries (kind, tile width,
// K-split).
std::vector<ScanPlan> ScanPlanCandidates(const LimbLayout& L, int nq, int out_bits);
// Query groups of 16 that one wide workgroup covers (ngp is a multiple of it).
int WideGroupsPerWorkgroup(const LimbLayout& L, const ScanPlan& P);
// Grouped plan: ngroups blocks of group_rows rows (a multiple of 64), block b
// times vector b.
ScanPlan MakeGroupedScanPlan(const LimbLayout& L, int ngroups, int group_rows, int out_bits);
// Sets the launch attributes the scan kernels need; call once per process
// (and device) before the first scan, outside any stream capture.
hipError_t ConfigureScanKernels();
// Plane-0 digit tiles of nq vectors (vec + q*stride, length len each, zero
// beyond len): layout [kb][group][tile][64 lanes][16 bytes].
hipError_t LaunchQueryDigits(const u64* vec, int stride, int len, const LimbLayout& L,
                             const ScanPlan& P, int8_t* btiles, hipStream_t s);
// corr[q] = c_off * sum(vec_q) mod 2^64; scratch holds
// VectorSumScratchWords(nq) words.
constexpr int kVectorSumChunks = 64;
int VectorSumScratchWords(int nq);
hipError_t LaunchVectorSums(const u64* vec, int nq, int stride, int len, u64 c_off,
                            u64* scratch, u64* corr, hipStream_t s);
// LaunchQueryDigits and LaunchVectorSums of the same vectors (vec rows of
// P.nq queries); digit-major plans read the values once for both. scratch:
// P.nq * max(kVectorSumChunks, DigitSumChunks(L)) words.
int DigitSumChunks(const LimbLayout& L);
hipError_t LaunchQueryDigitsSums(const u64* vec, int stride, int len, const LimbLayout& L,
                                 const ScanPlan& P, int8_t* btiles, u64 c_off, u64* scratch,
                                 u64* corr, hipStream_t s);
// partial[ks][q][row] = sum over the ks-th k-range of the limb product,
// combined over (plane, digit) pairs with shift < 8 * ns.
hipError_t LaunchScan(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                      const int8_t* btiles, u64* partial, hipStream_t s);
// The DB product's tall scan and the pack product's digit-major four-plane
// WideScan as one launch, when both plans are of those kinds
// (ScanPairFusable): the pack product's workgroups follow the DB product's
// in launch order and fill the CUs its last round leaves idle. Same results
// as the two LaunchScan calls.
bool ScanPairFusable(const LimbLayout& Ld, const ScanPlan& Pd, const LimbLayout& Lp,
                     const ScanPlan& Pp);
// Workgroups of the DB product's scan under ScanPairFusable's plans.
long long ScanWorkgroups(const LimbLayout& L, const ScanPlan& P);
hipError_t LaunchScanPair(const int8_t* db_tiles, const LimbLayout& Ld, const ScanPlan& Pd,
                          const int8_t* db_btiles, u64* db_partial, const int8_t* pk_tiles,
                          const LimbLayout& Lp, const ScanPlan& Pp, const int8_t* pk_btiles,
                          u64* pk_partial, hipStream_t s);

// ---------------------------------------------------------------------------
// FFT-side launchers (fft_kernels.hip). N = 2d, twiddles = N/2 complex.
struct FftDims {
  int d, log_d, N, log_N;
};

hipError_t ConfigureFftKernels(int N);
// dst[(p*num_chunks + c)*N] = FFT of chunk c of poly p, polys addressed as
// src + (p / per_group) * group_stride + (p % per_group) * d.
hipError_t LaunchSpectra(const u64* src, int npolys, int per_group,
                         long long group_stride, const ChunkRule& rule,
                         const FftDims& F, const double2* tw_fwd,
                         double2* dst, hipStream_t s);

// Stage arguments of ProcessResponse, batched over nq requests.
struct FinalizeArgs {
  const u64* db_partial;   // [ksplit_db][nq][M]
  int ksplit_db;
  const u64* corr_db;      // [nq]
  const u64* pk_partial;   // [ksplit_pk][nq][M]
  int ksplit_pk;
  const u64* corr_pk;      // [nq], or [M / corr_rows] when corr_rows > 0
  int corr_rows;           // > 0: pack correction per block of corr_rows rows
  const u64* fin_partial;  // [nq][ell*t][npairs][d]
  int npairs;
  const u64* a_tilde;      // [ell*t][d]
  u64* ct_b;               // [nq][ell*t][d]
  u64* st_a;               // [nq][ell][d]
  u64* st_b;               // [nq][ell][d]
  u64* out;                // [nq][ell][2][d] (written when t == 1)
  int nq, M, d, t, ell, log_q, log_q1, log_q2;
  u64 cmask;               // 2^width - 1 of CoeffType
};

struct DigitsArgs {
  const u64* st_a;   // [nq][ell][d]
  const u64* st_b;
  int nq, ell, d, e, log_q, log_digit;
  ChunkRule rule;    // u rule of the external product (digits)
  double2* dspec;    // [nq][ell][2e][nu][N]
  // prev_step >= 0: the Horner state is not in st_a / st_b but is the
  // previous step's result, formed here as ExtCombineKernel would: the sum of
  // its pair partials plus c_prev_step (a_tilde / ct_b).
  int prev_step;
  const u64* ext_partial;  // [nq][ell][2][npairs][d]
  int npairs, t;
  const u64* a_tilde;      // [ell*t][d]
  const u64* ct_b;         // [nq][ell*t][d]
};

struct ExtCombineArgs {
  const u64* ext_partial;  // [nq][ell][2][npairs][d]
  int npairs;
  const u64* a_tilde;      // [ell*t][d]
  const u64* ct_b;         // [nq][ell*t][d]
  u64* st_a;               // [nq][ell][d]
  u64* st_b;
  u64* out;                // [nq][ell][2][d]
  int nq, ell, t, d, step_i, log_q, log_q1, log_q2;
  u64 cmask;               // 2^width - 1 of CoeffType
};

// outs_per_group: consecutive outputs that share their u spectra (one
// request's outputs); their workgroups are placed on one XCD. acc (optional,
// kIpSplitMaxJobs * N complex values): launches of at most that many
// (output, pair) jobs first accumulate the spectra across the whole GPU into
// acc, then invert one job per workgroup. tile_outs > 0 (n_out a multiple of
// outs_per_group): tiles of tile_requests requests x tile_outs of their
// outputs take one XCD instead, so that the tile's u and v spectra share its
// L2.
constexpr int kIpSplitMaxJobs = 64;
hipError_t LaunchIp(const double2* const* uptr, const double2* const* vptr,
                    u64* const* optr, int n_out, int outs_per_group, int nterms, int nu,
                    int nv, const PairList& pairs, const FftDims& F,
                    const double2* tw_bwd, unsigned long long* margin,
                    hipStream_t s, double2* acc = nullptr, int tile_requests = 1,
                    int tile_outs = 0);
hipError_t LaunchFinalizeCombine(const FinalizeArgs& A, hipStream_t s);
hipError_t LaunchDigitsSpectra(const DigitsArgs& A, const FftDims& F,
                               const double2* tw_fwd, hipStream_t s);
hipError_t LaunchExtCombine(const ExtCombineArgs& A, hipStream_t s);
hipError_t LaunchSumPairs(const u64* partial, int npairs, int d, u64* out,
                          hipStream_t s);

}  // namespace pirg

#endif  // INSPIRE_AMD_PIRG_INTERNAL_H_
