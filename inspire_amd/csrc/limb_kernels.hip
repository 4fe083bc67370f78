// Int8-limb matrix x vector(s) products on CDNA4 MFMA.
//
// Both products of PirServer::ProcessResponse that stream resident matrices
// -- the DB product (Matrix<uint16_t|uint8_t>::Multiply<uint64_t>,
// crypto/matrix.cc:261-318 / :376-462) and the pack product
// (Matrix<int32_t|int16_t>::Multiply<uint64_t>, crypto/matrix.cc:320-374,
// crypto/matrix.h:183-194) -- compute out[i] = sum_j A[i][j] * y[j] mod 2^64
// with small-integer A and 64-bit y. Only out mod 2^log_q is observable
// downstream (FinalizeMatrixPack masks it at crypto/lwes_to_rlwe.cc:469).
//
// Here A is kept in HBM as int8 planes (A = C_off + sum_p 2^(8p) plane_p) and
// y is split into balanced int8 digits, so every (plane, digit) pair is an
// exact int8 x int8 -> int32 product on v_mfma_i32_16x16x64_i8; the pairs
// are recombined with 64-bit shifts. A k-range of at most 65536 columns keeps
// every int32 sum exact (|sum| <= 65536 * 2^14 = 2^30).
#include <algorithm>
#include <vector>

#include <type_traits>

#include "pirg_internal.h"

namespace pirg {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kMaxColsPerSplitNP1 = 65536;  // 2^16 / NP columns keep int32 sums exact

__device__ inline int8_t OffsetByte(unsigned x) {
  return static_cast<int8_t>(static_cast<int>(x & 255u) - 128);
}

// One thread per (strip, kblock, lane): builds 16 bytes of every plane.
__global__ void DbToTilesKernel(const u64* __restrict__ cdata, int db_bytes,
                                int rows, int cols, int row0, int nrows,
                                int strips, int kblocks, int planes,
                                int8_t* __restrict__ tiles) {
  const long long strip_lo = row0 / 16;
  const long long strip_hi = (row0 + nrows + 15) / 16;
  const long long total = (strip_hi - strip_lo) * kblocks * 64;
  const int wpr = (db_bytes == 2) ? cols / 4 : cols / 8;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int lane = static_cast<int>(g & 63);
    const long long t = g >> 6;
    const int kb = static_cast<int>(t % kblocks);
    const long long strip = strip_lo + t / kblocks;
    if (strip >= strips) continue;
    const long long row = strip * 16 + (lane & 15);
    const int kbase = kb * 64 + 16 * (lane >> 4);
    unsigned vals[16];
    const bool row_ok = row < rows && row >= row0 && row < row0 + nrows;
    // A strip shared with a neighbouring chunk keeps that chunk's rows.
    if (!row_ok && row < rows) continue;
    for (int b = 0; b < 16; ++b) {
      const int k = kbase + b;
      unsigned x = 0;
      if (row_ok && k < cols) {
        const u64* rp = cdata + (row - row0) * (long long)wpr;
        if (db_bytes == 2) {
          x = static_cast<unsigned>((rp[k >> 2] >> (16 * (k & 3))) & 0xFFFF);
        } else {
          x = static_cast<unsigned>((rp[k >> 3] >> (8 * (k & 7))) & 0xFF);
        }
      }
      vals[b] = x;
    }
    for (int p = 0; p < planes; ++p) {
      int8_t out[16];
      for (int b = 0; b < 16; ++b) out[b] = OffsetByte(vals[b] >> (8 * p));
      int8_t* dst = tiles + (TileIndex(strip, kb, p, kblocks, planes) * 64 + lane) * 16;
      *reinterpret_cast<int4*>(dst) = *reinterpret_cast<const int4*>(out);
    }
  }
}

// Pack matrix: signed values; planes 0..NP-2 hold offset unsigned bytes, the
// top plane the arithmetic-shifted signed top byte.
__global__ void PackToTilesKernel(const void* __restrict__ pack, int mat_bytes,
                                  int rows, int cols, int row0, int nrows,
                                  int strips, int kblocks, int planes,
                                  int8_t* __restrict__ tiles) {
  const long long strip_lo = row0 / 16;
  const long long strip_hi = (row0 + nrows + 15) / 16;
  const long long total = (strip_hi - strip_lo) * kblocks * 64;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int lane = static_cast<int>(g & 63);
    const long long t = g >> 6;
    const int kb = static_cast<int>(t % kblocks);
    const long long strip = strip_lo + t / kblocks;
    if (strip >= strips) continue;
    const long long row = strip * 16 + (lane & 15);
    const int kbase = kb * 64 + 16 * (lane >> 4);
    int vals[16];
    const bool row_ok = row < rows && row >= row0 && row < row0 + nrows;
    if (!row_ok && row < rows) continue;
    for (int b = 0; b < 16; ++b) {
      const int k = kbase + b;
      int x = 0;
      if (row_ok && k < cols) {
        if (mat_bytes == 4) {
          const u64* rp = static_cast<const u64*>(pack) + (row - row0) * (long long)(cols / 2);
          x = static_cast<int>(static_cast<uint32_t>(rp[k >> 1] >> (32 * (k & 1))));
        } else {
          const int16_t* rp = static_cast<const int16_t*>(pack) + (row - row0) * (long long)cols;
          x = rp[k];
        }
      }
      vals[b] = x;
    }
    for (int p = 0; p < planes; ++p) {
      int8_t out[16];
      for (int b = 0; b < 16; ++b) {
        if (p == planes - 1) {
          out[b] = static_cast<int8_t>(vals[b] >> (8 * p));
        } else {
          out[b] = OffsetByte(static_cast<unsigned>(vals[b]) >> (8 * p));
        }
      }
      int8_t* dst = tiles + (TileIndex(strip, kb, p, kblocks, planes) * 64 + lane) * 16;
      *reinterpret_cast<int4*>(dst) = *reinterpret_cast<const int4*>(out);
    }
  }
}

// The inverse of the two conversions above, for SaveToFile: rows
// [row0, row0 + nrows) of a tiled matrix back to row-major values (out is
// nrows x cols, sign-extended to 64 bits when the top plane is signed, then
// ANDed with mask). One thread per (row, k-block, 16-column group) reads that
// group's 16 bytes in every plane.
__global__ void TilesToValuesKernel(const int8_t* __restrict__ tiles, int cols, int row0,
                                    int nrows, int kblocks, int planes, int signed_top,
                                    u64 mask, u64* __restrict__ out) {
  const long long total = static_cast<long long>(nrows) * kblocks * 4;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int lg = static_cast<int>(g & 3);
    const long long t = g >> 2;
    const int kb = static_cast<int>(t % kblocks);
    const int rr = static_cast<int>(t / kblocks);
    const long long row = static_cast<long long>(row0) + rr;
    const long long strip = row / 16;
    const int lane = static_cast<int>(row % 16) + 16 * lg;
    long long vals[16];
    for (int b = 0; b < 16; ++b) vals[b] = 0;
    for (int p = 0; p < planes; ++p) {
      int8_t in[16];
      *reinterpret_cast<int4*>(in) = *reinterpret_cast<const int4*>(
          tiles + (TileIndex(strip, kb, p, kblocks, planes) * 64 + lane) * 16);
      const bool top = signed_top && p == planes - 1;
      for (int b = 0; b < 16; ++b) {
        const long long byte = top ? static_cast<long long>(in[b])
                                   : static_cast<long long>(static_cast<int>(in[b]) + 128);
        vals[b] += byte * (1LL << (8 * p));
      }
    }
    const int c0 = kb * 64 + 16 * lg;
    u64* dst = out + static_cast<long long>(rr) * cols;
    for (int b = 0; b < 16; ++b) {
      if (c0 + b < cols) dst[c0 + b] = static_cast<u64>(vals[b]) & mask;
    }
  }
}

__device__ inline int BalancedDigit(u64 y, int c) {
  u64 v = y;
  int dig = 0;
  for (int cc = 0; cc <= c; ++cc) {
    const unsigned byte = static_cast<unsigned>(v & 255u);
    const bool hi = byte >= 128u;
    dig = hi ? static_cast<int>(byte) - 256 : static_cast<int>(byte);
    v = (v >> 8) + (hi ? 1 : 0);
  }
  return dig;
}

// B operand tiles (plane-0 digits only). Layout [kb][group][tile][64 lanes]
// [16 bytes]. Group g holds queries [g*qpg, (g+1)*qpg). Query-major: its
// column j is (query g*qpg + j / ns, shift s = j % ns) for j < qpg*ns and zero
// above. Digit-major (ctg == ns, qpg == 16): column c of tile s is (query
// g*16 + c, shift s). The column holds balanced digit s of the query. Lane l
// of a tile holds column l % 16 for the 16 k's 16*(l/16) .. +15 of the
// k-block. One thread per (kb, group, tile, lane).
__global__ void QueryDigitsKernel(const u64* __restrict__ vec, int nq, int stride,
                                  int len, int kblocks, int ns, int qpg, int ngp,
                                  int ctg, int digit_major, int8_t* __restrict__ btiles) {
  const long long total = (long long)kblocks * ngp * ctg * 64;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int lane = static_cast<int>(g & 63);
    long long t = g >> 6;
    const int ct = static_cast<int>(t % ctg);
    t /= ctg;
    const int grp = static_cast<int>(t % ngp);
    const int kb = static_cast<int>(t / ngp);
    const int j = ct * 16 + (lane & 15);
    const int q = digit_major ? grp * qpg + (lane & 15) : grp * qpg + j / ns;
    const int s = digit_major ? ct : j % ns;
    const bool valid = (digit_major || j < qpg * ns) && q < nq;
    const int kbase = kb * 64 + 16 * (lane >> 4);
    int8_t out[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const int k = kbase + b;
      const u64 y = (valid && k < len) ? vec[(long long)q * stride + k] : 0ULL;
      out[b] = static_cast<int8_t>(BalancedDigit(y, s));
    }
    int8_t* dst = btiles + (((((long long)kb * ngp + grp) * ctg + ct) * 64) + lane) * 16;
    *reinterpret_cast<int4*>(dst) = *reinterpret_cast<const int4*>(out);
  }
}

// Digit-major B tiles (ctg == ns, qpg == 16), one thread per (kb, group,
// lane): the lane's query q = group*16 + lane%16 and k's 16*(lane/16) .. +15
// are read once and all ns balanced digits are produced in one carry pass
// (the recurrence of BalancedDigit), one 16-byte store per digit tile.
__global__ void QueryDigitsDigitMajorKernel(const u64* __restrict__ vec, int nq, int stride,
                                            int len, int kblocks, int ns, int ngp,
                                            int8_t* __restrict__ btiles) {
  const long long total = (long long)kblocks * ngp * 64;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int lane = static_cast<int>(g & 63);
    const long long t = g >> 6;
    const int grp = static_cast<int>(t % ngp);
    const int kb = static_cast<int>(t / ngp);
    const int q = grp * 16 + (lane & 15);
    const int kbase = kb * 64 + 16 * (lane >> 4);
    u64 v[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const int k = kbase + b;
      v[b] = (q < nq && k < len) ? vec[(long long)q * stride + k] : 0ULL;
    }
    int8_t* dst = btiles + ((((long long)kb * ngp + grp) * ns) * 64 + lane) * 16;
    for (int sft = 0; sft < ns; ++sft) {
      int8_t out[16];
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const unsigned byte = static_cast<unsigned>(v[b] & 255u);
        const bool hi = byte >= 128u;
        out[b] = static_cast<int8_t>(hi ? static_cast<int>(byte) - 256 : static_cast<int>(byte));
        v[b] = (v[b] >> 8) + (hi ? 1 : 0);
      }
      *reinterpret_cast<int4*>(dst + (long long)sft * 64 * 16) =
          *reinterpret_cast<const int4*>(out);
    }
  }
}

// Digit-major B tiles and the column sums of the same query values in one
// pass. A workgroup owns one group of 16 queries over a chunk of k-blocks,
// each wave one k-block at a time: the wave reads the 16 queries' 64 values
// of the k-block row by row (512 contiguous bytes per load), transposes them
// through its LDS region so that lane l holds query l % 16 at k's
// 16 * (l / 16) .. +15 (the B operand layout), and emits the ns balanced
// digit tiles as QueryDigitsDigitMajorKernel does. The lanes' sums of their
// values reduce, per query, into part[q * nchunks + chunk] (raw sums mod
// 2^64; VectorSumFinalKernel multiplies by C_off).
constexpr int kDigitSumWaves = 4;
constexpr int kDigitSumRow = 65;  // u64 per LDS row: 64 values + 1 of padding

__global__ __launch_bounds__(256) void QueryDigitsSumsKernel(
    const u64* __restrict__ vec, int nq, int stride, int len, int kblocks, int ns, int ngp,
    int kb_per_chunk, int nchunks, int8_t* __restrict__ btiles, u64* __restrict__ part) {
  __shared__ u64 rows[kDigitSumWaves][16 * kDigitSumRow];
  __shared__ u64 qsum[kDigitSumWaves][16];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int chunk = blockIdx.x;
  const int grp = blockIdx.y;
  const int kb_lo = chunk * kb_per_chunk;
  const int kb_hi = min(kblocks, kb_lo + kb_per_chunk);
  const int qi = lane & 15;
  const int kc = lane >> 4;
  u64* const mine = rows[w];
  u64 acc = 0;
  for (int kb = kb_lo + w; kb < kb_hi; kb += kDigitSumWaves) {
    // Two rows per load instruction: lanes 0-31 take row i's 64 values in
    // pairs, lanes 32-63 row i + 1's (16 bytes per lane).
    const int k = kb * 64 + 2 * (lane & 31);
    const int half = lane >> 5;
#pragma unroll
    for (int i = 0; i < 16; i += 2) {
      const int q = grp * 16 + i + half;
      const u64* src = vec + (long long)q * stride + k;
      u64 x0 = 0, x1 = 0;
      if (q < nq) {
        if (k + 1 < len && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
          const ulonglong2 p2 = *reinterpret_cast<const ulonglong2*>(src);
          x0 = p2.x;
          x1 = p2.y;
        } else {
          x0 = k < len ? src[0] : 0ULL;
          x1 = k + 1 < len ? src[1] : 0ULL;
        }
      }
      mine[(i + half) * kDigitSumRow + 2 * (lane & 31)] = x0;
      mine[(i + half) * kDigitSumRow + 2 * (lane & 31) + 1] = x1;
    }
    // The wave's own LDS region: its stores are visible to its loads after
    // the wave-wide fence.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    u64 v[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      v[b] = mine[qi * kDigitSumRow + kc * 16 + b];
      acc += v[b];
    }
    // Every lane has read the region before the next k-block overwrites it.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int8_t* dst = btiles + ((((long long)kb * ngp + grp) * ns) * 64 + lane) * 16;
    for (int sft = 0; sft < ns; ++sft) {
      int8_t out[16];
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const unsigned byte = static_cast<unsigned>(v[b] & 255u);
        const bool hi = byte >= 128u;
        out[b] = static_cast<int8_t>(hi ? static_cast<int>(byte) - 256 : static_cast<int>(byte));
        v[b] = (v[b] >> 8) + (hi ? 1 : 0);
      }
      *reinterpret_cast<int4*>(dst + (long long)sft * 64 * 16) =
          *reinterpret_cast<const int4*>(out);
    }
  }
  // Lanes qi, qi + 16, qi + 32, qi + 48 hold query qi's partial sums.
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  if (lane < 16) qsum[w][lane] = acc;
  __syncthreads();
  if (threadIdx.x < 16) {
    const int q = grp * 16 + threadIdx.x;
    u64 t = 0;
#pragma unroll
    for (int ww = 0; ww < kDigitSumWaves; ++ww) t += qsum[ww][threadIdx.x];
    if (q < nq) part[(long long)q * nchunks + chunk] = t;
  }
}

// corr[q] = c_off * sum of query q's nchunks partial sums: one wave per query.
__global__ __launch_bounds__(64) void ChunkSumFinalKernel(const u64* __restrict__ part,
                                                          int nchunks, u64 c_off,
                                                          u64* __restrict__ corr) {
  const int q = blockIdx.x;
  u64 t = 0;
  for (int c = threadIdx.x; c < nchunks; c += 64) t += part[(long long)q * nchunks + c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
  if (threadIdx.x == 0) corr[q] = c_off * t;
}

// corr[q] = c_off * sum_k vec[q][k] mod 2^64, in two passes: partial sums of
// chunks, then one thread per query.
__global__ void VectorPartialSumKernel(const u64* __restrict__ vec, int stride,
                                       int len, int chunk, int nchunks,
                                       u64* __restrict__ part) {
  __shared__ u64 red[256];
  const int q = blockIdx.y;
  const int c = blockIdx.x;
  const int k0 = c * chunk;
  const int k1 = min(len, k0 + chunk);
  u64 s = 0;
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) s += vec[(long long)q * stride + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)q * nchunks + c] = red[0];
}

__global__ void VectorSumFinalKernel(const u64* __restrict__ part, int nchunks, int nq,
                                     u64 c_off, u64* __restrict__ corr) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  u64 s = 0;
  for (int c = 0; c < nchunks; ++c) s += part[(long long)q * nchunks + c];
  corr[q] = c_off * s;
}

// Plane p of the B operand is plane 0 with every column moved p places
// right: column (query, s) of plane p holds digit s - p of the query. Within
// a 16-column tile that is a DPP shift inside each 16-lane row (the 16 lanes
// of a row are the 16 columns); the first p columns come from the previous
// tile (row rotate). Columns with s < p belong to the previous query and are
// zeroed by the caller's mask.
template <int P>
__device__ __forceinline__ v4i ShiftColumns(const v4i cur, const v4i prev) {
  if constexpr (P == 0) {
    (void)prev;
    return cur;
  } else {
    v4i out;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int rot = __builtin_amdgcn_update_dpp(0, prev[e], 0x120 + P, 0xF, 0xF, false);
      out[e] = __builtin_amdgcn_update_dpp(rot, cur[e], 0x110 + P, 0xF, 0xF, false);
    }
    return out;
  }
}

__device__ __forceinline__ v4i Masked(const v4i x, bool keep) {
  return keep ? x : v4i{0, 0, 0, 0};
}

// Column-validity bits of a lane: bit (t * 4 + p) is set when column
// t*16 + lane%16 of a group has shift >= p (or lies in the padding).
__device__ __forceinline__ unsigned ColumnMask(int ntiles, int np, int ns, int lane) {
  unsigned m = 0;
  for (int t = 0; t < ntiles; ++t) {
    const int s = (t * 16 + (lane & 15)) % ns;
    for (int p = 0; p < np; ++p) {
      if (s >= p) m |= 1u << (t * 4 + p);
    }
  }
  return m;
}

constexpr int kQMax = 16;  // queries per group (bounds the epilogue buffer)

// Epilogue of a wave: acc[r][t] (16x16 MFMA layout: lane l, element v holds
// row 4*(l/16) + v of strip r and column t*16 + l%16 of the wave's group)
// -> partial[ks][q][row] = sum over shifts s of acc << 8s. The wave reduces
// through its own LDS region red (nstrips*16 rows x kQMax queries).
template <int NS, int NT>
__device__ __forceinline__ void WaveEpilogue(const v4i (&acc)[NS][NT], u64* red, int lane,
                                             int ns, int qpg, int grp, int nq, int rows,
                                             long long row0, int ks, u64* partial) {
  constexpr int kRows = NS * 16;
  for (int i = lane; i < kRows * kQMax; i += 64) red[i] = 0;
#pragma unroll
  for (int r = 0; r < NS; ++r) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int j = t * 16 + (lane & 15);
      if (j < qpg * ns) {
        const int q = j / ns;
        const int sh = 8 * (j - q * ns);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rl = r * 16 + 4 * (lane >> 4) + v;
          atomicAdd(reinterpret_cast<unsigned long long*>(&red[rl * kQMax + q]),
                    static_cast<unsigned long long>(static_cast<u64>(static_cast<i64>(acc[r][t][v])) << sh));
        }
      }
    }
  }
  for (int i = lane; i < kRows * qpg; i += 64) {
    const int q = i / kRows;
    const int rl = i % kRows;
    const int qg = grp * qpg + q;
    const long long row = row0 + rl;
    if (qg < nq && row < rows) {
      partial[((long long)ks * nq + qg) * rows + row] = red[rl * kQMax + q];
    }
  }
}

// Digit-major epilogue: acc[r][s] lane l, element v holds digit-s partial of
// query grp*16 + l%16 for row 4*(l/16) + v of strip r, so a lane combines its
// own NT tiles (sum of acc << 8s) and the wave stages the values in red
// (nstrips*16 rows x 16 queries) for row-contiguous stores.
template <int NS, int NT>
__device__ __forceinline__ void WaveEpilogueDigitMajor(const v4i (&acc)[NS][NT], u64* red,
                                                       int lane, int grp, int nq, int rows,
                                                       long long row0, int ks, u64* partial) {
  constexpr int kRows = NS * 16;
  const int ql = lane & 15;
#pragma unroll
  for (int r = 0; r < NS; ++r) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      u64 sum = 0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        sum += static_cast<u64>(static_cast<i64>(acc[r][t][v])) << (8 * t);
      }
      red[(r * 16 + 4 * (lane >> 4) + v) * kQMax + ql] = sum;
    }
  }
  __builtin_amdgcn_wave_barrier();
  for (int i = lane; i < kRows * 16; i += 64) {
    const int q = i / kRows;
    const int rl = i % kRows;
    const int qg = grp * 16 + q;
    const long long row = row0 + rl;
    if (qg < nq && row < rows) {
      partial[((long long)ks * nq + qg) * rows + row] = red[rl * kQMax + q];
    }
  }
}

// ---------------------------------------------------------------------------
// NarrowScan: few queries, bound by the A stream. Each of the 4 waves owns R
// strips and the workgroup's group of T tiles; the waves never synchronise.
// A and B fragments are prefetched D-1 k-blocks ahead into rotating register
// sets (the loop is unrolled by D so every set is a fixed register range).
template <int NP, int R, int T, int D>
__global__ __launch_bounds__(256) void NarrowScanKernel(
    const int8_t* __restrict__ tiles, int strips, int rows, int kblocks, int kbps,
    const int8_t* __restrict__ btiles, int ngp, int nq, int ns, int qpg, int group_rows,
    u64* __restrict__ partial) {
  __shared__ u64 red[4][R * 16 * kQMax];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int strip0 = (blockIdx.x * 4 + wave) * R;
  const int ks = blockIdx.y;
  // Grouped product: the wave's rows (R strips, inside one block of
  // group_rows rows) use their block's group and write the single-query
  // layout.
  const int grp = group_rows > 0 ? min(strip0 * 16, rows - 1) / group_rows : blockIdx.z;
  const int kb0 = ks * kbps;
  const int nkb = min(kblocks, kb0 + kbps) - kb0;
  const v4i* A = reinterpret_cast<const v4i*>(tiles);
  const v4i* Bg = reinterpret_cast<const v4i*>(btiles) + (long long)grp * T * 64 + lane;
  const v4i* Ar[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    Ar[r] = A + TileIndex(min(strip0 + r, strips - 1), 0, 0, kblocks, NP) * 64 + lane;
  }
  const unsigned cmask = ColumnMask(T, NP, ns, lane);
  v4i acc[R][T];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int t = 0; t < T; ++t) acc[r][t] = v4i{0, 0, 0, 0};
  v4i a[D][R][NP];
  v4i b[D][T];
  // Steps past the k-range read A's last k-block and B's all-zero block
  // `kblocks`, so they add nothing.
  auto load = [&](int i, v4i (&as)[R][NP], v4i (&bs)[T]) {
    const int kb = kb0 + min(i, nkb - 1);
    const int kbb = i < nkb ? kb0 + i : kblocks;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        // Streamed once: non-temporal loads (measured +11% bandwidth).
        as[r][p] = __builtin_nontemporal_load(Ar[r] + ((long long)kb * kStripGroup * NP + p) * 64);
      }
#pragma unroll
    for (int t = 0; t < T; ++t) bs[t] = Bg[((long long)kbb * ngp * T + t) * 64];
  };
  auto compute = [&](const v4i (&as)[R][NP], const v4i (&b0)[T]) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      v4i bp[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const v4i prev = t > 0 ? b0[t > 0 ? t - 1 : 0] : v4i{0, 0, 0, 0};
        v4i x;
        if (p == 0) x = b0[t];
        else if (p == 1) x = ShiftColumns<1>(b0[t], prev);
        else if (p == 2) x = ShiftColumns<2>(b0[t], prev);
        else x = ShiftColumns<3>(b0[t], prev);
        bp[t] = p == 0 ? x : Masked(x, (cmask >> (t * 4 + p)) & 1u);
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int t = 0; t < T; ++t)
          acc[r][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[r][p], bp[t], acc[r][t], 0, 0, 0);
    }
  };
#pragma unroll
  for (int d = 0; d < D - 1; ++d) load(d, a[d], b[d]);
  for (int i = 0; i < nkb; i += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int nx = (d + D - 1) % D;
      load(i + d + D - 1, a[nx], b[nx]);
      compute(a[d], b[d]);
    }
  }
  WaveEpilogue<R, T>(acc, red[wave], lane, ns, qpg, group_rows > 0 ? 0 : grp,
                     group_rows > 0 ? 1 : nq, rows, (long long)strip0 * 16, ks, partial);
}

// ---------------------------------------------------------------------------
// WideScan: batches, bound by MFMA. A workgroup of 8 waves owns RS = 8 strips
// (128 rows) and 4 consecutive groups (CT tiles = up to 16 queries each).
// Wave (h, g) computes strips [4h, 4h + 4) of group g: 4 x CT x NP MFMAs per
// k-block, two waves per SIMD so one computes while the other waits.
//  * Every global load is an LDS-DMA (global_load_lds_dwordx4) into an
//    S-stage LDS ring: the k-block's A fragments (2*NP per wave) and the
//    groups' B tiles (a group's CT tiles split between its two waves). No
//    load has a register destination, so nothing waits on the loads but the
//    explicit vmcnt wait before each step's barrier (DMA ops complete in
//    issue order).
//  * The barrier is s_barrier after an LDS-only wait: HIP's __syncthreads
//    also waits for every outstanding load, which would drain the ring.
//  * Query-major B: plane p's fragments are derived from plane 0's with DPP
//    shifts. Digit-major B: plane p uses digit tile s - p as loaded.
__device__ __forceinline__ void LdsBarrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int N>
__device__ __forceinline__ void WaitVmcnt() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

__device__ __forceinline__ void DmaLoad16(const v4i* src, v4i* lds_dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_dst, 16,
                                   0, 0);
}

// The same with the non-temporal cache policy, for data streamed once.
__device__ __forceinline__ void DmaLoadStream16(const v4i* src, v4i* lds_dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_dst, 16,
                                   0, 2);
}

template <int NP, int CT>
struct WideShape {
  static constexpr int RS = 8;                // strips per workgroup
  static constexpr int RW = 4;                // strips per wave
  static constexpr int AF = RS * NP / 8;      // A fragments DMA'd per wave per k-block
  static constexpr int BH0 = (CT + 1) / 2;    // B tiles DMA'd by the h = 0 wave of a group
  static constexpr int BH1 = CT - BH0;
  static constexpr int kA = RS * NP * 64;     // v4i of A per stage
  static constexpr int kB = 4 * CT * 64;      // v4i of B per stage
  static constexpr int kStage = kA + kB;
  static constexpr int kBudget = 150 * 1024 / 16;
  static constexpr int S = kBudget / kStage >= 4 ? 4 : (kBudget / kStage >= 3 ? 3 : 2);
  static constexpr int kRedWords = 8 * RW * 16 * kQMax;
  static constexpr size_t kLdsBytes =
      (static_cast<size_t>(S) * kStage * 16 > static_cast<size_t>(kRedWords) * 8)
          ? static_cast<size_t>(S) * kStage * 16
          : static_cast<size_t>(kRedWords) * 8;
};

// Workgroup order. Workgroups are dispatched to the 8 XCDs round robin
// (workgroup w to XCD w % 8), each XCD with its own L2. The nz column blocks
// of one (row block, k-split) tile read the same A tiles, so they are given
// consecutive slots on one XCD: w = (u * nz + j) * 8 + x, tile = u * 8 + x,
// column block j. They run side by side and share A through that XCD's L2
// (A is fetched from HBM about once instead of nz times). Tiles past the
// last (the grid is padded to whole groups of 8) exit at once.
struct WideTile {
  int strip_block, ks, zb;
};

__device__ __forceinline__ bool DecodeWideTile(int w, int nx, int ntiles, int nz, WideTile* t) {
  const int x = w & 7;
  const int rest = w >> 3;
  const int tile = (rest / nz) * 8 + x;
  if (tile >= ntiles) return false;
  t->strip_block = tile % nx;
  t->ks = tile / nx;
  t->zb = rest % nz;
  return true;
}

// w: the workgroup's position in the launch order (blockIdx.x of its own
// launch); lds: the launch's dynamic LDS (WideShape<NP, CT>::kLdsBytes).
template <int NP, int CT, bool DM>
__device__ __forceinline__ void WideScanBody(
    int w, const int8_t* __restrict__ tiles, int strips, int rows, int kblocks, int kbps,
    int ksplit, const int8_t* __restrict__ btiles, int ngp, int nq, int ns, int qpg,
    u64* __restrict__ partial, v4i* lds) {
  using W = WideShape<NP, CT>;
  constexpr int RS = W::RS, RW = W::RW, AF = W::AF, S = W::S;
  const int nx = (strips + RS - 1) / RS;
  WideTile wt;
  if (!DecodeWideTile(w, nx, nx * ksplit, ngp / 4, &wt)) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int h = wave >> 2;       // strip half
  const int g = wave & 3;        // group within the workgroup
  const int strip0 = wt.strip_block * RS;
  const int ks = wt.ks;
  const int grp = wt.zb * 4 + g;
  const int kb0 = ks * kbps;
  const int nkb = min(kblocks, kb0 + kbps) - kb0;
  const v4i* A = reinterpret_cast<const v4i*>(tiles);
  const v4i* Bg = reinterpret_cast<const v4i*>(btiles) + (long long)grp * CT * 64 + lane;
  // A fragment f = wave*AF + jf of a k-block is (strip f / NP, plane f % NP)
  // and lands at stage offset f*64: the stage holds A as [strip][plane].
  const v4i* Af[AF];
#pragma unroll
  for (int jf = 0; jf < AF; ++jf) {
    const int f = wave * AF + jf;
    Af[jf] = A + TileIndex(min(strip0 + f / NP, strips - 1), 0, f % NP, kblocks, NP) * 64 + lane;
  }
  const int c0 = h == 0 ? 0 : W::BH0;  // first B tile this wave issues
  const unsigned cmask = ColumnMask(CT, NP, ns, lane);
  v4i acc[RW][CT];
#pragma unroll
  for (int r = 0; r < RW; ++r)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[r][c] = v4i{0, 0, 0, 0};

  // Steps past the k-range read A's last k-block and B's all-zero block.
  auto issue = [&](int i) {
    v4i* st = lds + (i % S) * W::kStage;
    const long long kba = kb0 + min(i, nkb - 1);
    const long long kbb = i < nkb ? kb0 + i : kblocks;
#pragma unroll
    for (int jf = 0; jf < AF; ++jf) {
      DmaLoadStream16(Af[jf] + kba * kStripGroup * NP * 64, st + (wave * AF + jf) * 64);
    }
#pragma unroll
    for (int cc = 0; cc < W::BH0; ++cc) {
      if (h == 0 || cc < W::BH1) {
        const int c = c0 + cc;
        DmaLoad16(Bg + (kbb * ngp * CT + c) * 64, st + W::kA + (g * CT + c) * 64);
      }
    }
  };
  auto compute = [&](int i) {
    const v4i* st = lds + (i % S) * W::kStage;
    v4i b0[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) b0[c] = st[W::kA + (g * CT + c) * 64 + lane];
    if constexpr (DM) {
      // Output tile c takes plane p times digit c - p: no data movement.
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int r = 0; r < RW; ++r) {
          const v4i af = st[((h * RW + r) * NP + p) * 64 + lane];
#pragma unroll
          for (int c = p; c < CT; ++c)
            acc[r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, b0[c - p], acc[r][c], 0, 0, 0);
        }
      }
      return;
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      v4i bp[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const v4i prev = c > 0 ? b0[c > 0 ? c - 1 : 0] : v4i{0, 0, 0, 0};
        v4i x;
        if (p == 0) x = b0[c];
        else if (p == 1) x = ShiftColumns<1>(b0[c], prev);
        else if (p == 2) x = ShiftColumns<2>(b0[c], prev);
        else x = ShiftColumns<3>(b0[c], prev);
        bp[c] = p == 0 ? x : Masked(x, (cmask >> (c * 4 + p)) & 1u);
      }
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        const v4i af = st[((h * RW + r) * NP + p) * 64 + lane];
#pragma unroll
        for (int c = 0; c < CT; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bp[c], acc[r][c], 0, 0, 0);
      }
    }
  };
#pragma unroll
  for (int s = 0; s < S - 1; ++s) issue(s);
  for (int i = 0; i < nkb; ++i) {
    // This wave's DMAs of step i have landed (the later S-2 steps' may not).
    if (h == 0) {
      WaitVmcnt<(AF + W::BH0) * (S - 2)>();
    } else {
      WaitVmcnt<(AF + W::BH1) * (S - 2)>();
    }
    LdsBarrier();  // every wave's DMAs of step i have landed; step i-1 is consumed
    issue(i + S - 1);
    compute(i);
  }
  // Drain the ring (the last issues target clamped / zero blocks) before
  // its memory is reused by the epilogue.
  WaitVmcnt<0>();
  LdsBarrier();
  u64* red = reinterpret_cast<u64*>(lds) + wave * (RW * 16 * kQMax);
  if constexpr (DM) {
    WaveEpilogueDigitMajor<RW, CT>(acc, red, lane, grp, nq, rows,
                                   (long long)(strip0 + h * RW) * 16, ks, partial);
  } else {
    WaveEpilogue<RW, CT>(acc, red, lane, ns, qpg, grp, nq, rows,
                         (long long)(strip0 + h * RW) * 16, ks, partial);
  }
}

template <int NP, int CT, bool DM>
__global__ __launch_bounds__(512) void WideScanKernel(
    const int8_t* __restrict__ tiles, int strips, int rows, int kblocks, int kbps, int ksplit,
    const int8_t* __restrict__ btiles, int ngp, int nq, int ns, int qpg,
    u64* __restrict__ partial) {
  extern __shared__ v4i lds[];
  WideScanBody<NP, CT, DM>(blockIdx.x, tiles, strips, rows, kblocks, kbps, ksplit, btiles, ngp,
                           nq, ns, qpg, partial, lds);
}

// Digit-major WideScan with software-pipelined fragments (NP <= 2): the
// fragments of step i + 1 are read from LDS into a second register set while
// step i's MFMAs run, and the step's LDS-DMA issues and fragment reads are
// spread between its RW x NP MFMA blocks (fenced by sched_barrier) so the
// memory pipeline drains while the matrix cores work. Same LDS ring, tile
// order, vmcnt accounting and epilogue as WideScanKernel; the loop runs an
// even number of steps (a step past the k-range reads the all-zero B block).
template <int NP, int CT>
__global__ __launch_bounds__(512) void WideScanPipeKernel(
    const int8_t* __restrict__ tiles, int strips, int rows, int kblocks, int kbps, int ksplit,
    const int8_t* __restrict__ btiles, int ngp, int nq, u64* __restrict__ partial) {
  using W = WideShape<NP, CT>;
  constexpr int RS = W::RS, RW = W::RW, AF = W::AF, S = W::S;
  extern __shared__ v4i lds[];
  const int nx = (strips + RS - 1) / RS;
  WideTile wt;
  if (!DecodeWideTile(blockIdx.x, nx, nx * ksplit, ngp / 4, &wt)) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int h = wave >> 2;
  const int g = wave & 3;
  const int strip0 = wt.strip_block * RS;
  const int ks = wt.ks;
  const int grp = wt.zb * 4 + g;
  const int kb0 = ks * kbps;
  const int nkb = min(kblocks, kb0 + kbps) - kb0;
  const v4i* A = reinterpret_cast<const v4i*>(tiles);
  const v4i* Bg = reinterpret_cast<const v4i*>(btiles) + (long long)grp * CT * 64 + lane;
  const v4i* Af[AF];
#pragma unroll
  for (int jf = 0; jf < AF; ++jf) {
    const int f = wave * AF + jf;
    Af[jf] = A + TileIndex(min(strip0 + f / NP, strips - 1), 0, f % NP, kblocks, NP) * 64 + lane;
  }
  const int c0 = h == 0 ? 0 : W::BH0;
  v4i acc[RW][CT];
#pragma unroll
  for (int r = 0; r < RW; ++r)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[r][c] = v4i{0, 0, 0, 0};

  constexpr int kDma = AF + W::BH0;    // DMA slots of a step (an h = 1 wave skips the last)
  constexpr int kFrag = CT + RW * NP;  // fragments a wave reads per step
  constexpr int kBlk = RW * NP;        // MFMA blocks per step, one per (strip, plane)
  // DMA slot j of step i: A fragment j, then this wave's share of the B tiles.
  auto dma = [&](int i, int j) {
    v4i* st = lds + (i % S) * W::kStage;
    if (j < AF) {
      const long long kba = kb0 + min(i, nkb - 1);
      DmaLoadStream16(Af[j] + kba * kStripGroup * NP * 64, st + (wave * AF + j) * 64);
    } else {
      const int cc = j - AF;
      if (h == 0 || cc < W::BH1) {
        const long long kbb = i < nkb ? kb0 + i : kblocks;
        const int c = c0 + cc;
        DmaLoad16(Bg + (kbb * ngp * CT + c) * 64, st + W::kA + (g * CT + c) * 64);
      }
    }
  };
  auto frag = [&](int i, int j, v4i (&a)[RW][NP], v4i (&b)[CT]) {
    const v4i* st = lds + (i % S) * W::kStage;
    if (j < CT) {
      b[j] = st[W::kA + (g * CT + j) * 64 + lane];
    } else {
      const int rp = j - CT;
      a[rp / NP][rp % NP] = st[((h * RW + rp / NP) * NP + rp % NP) * 64 + lane];
    }
  };
  // Step i on (ca, cb); DMAs of step i + S; fragments of step i + 1 into (na, nb).
  auto step = [&](int i, v4i (&ca)[RW][NP], v4i (&cb)[CT], v4i (&na)[RW][NP],
                  v4i (&nb)[CT]) {
#pragma unroll
    for (int blk = 0; blk < kBlk; ++blk) {
      const int r = blk / NP, p = blk % NP;
#pragma unroll
      for (int c = p; c < CT; ++c)
        acc[r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ca[r][p], cb[c - p], acc[r][c], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = blk * kDma / kBlk; j < (blk + 1) * kDma / kBlk; ++j) dma(i + S, j);
#pragma unroll
      for (int j = blk * kFrag / kBlk; j < (blk + 1) * kFrag / kBlk; ++j) frag(i + 1, j, na, nb);
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // Own DMAs per step: all but the newest (S - 1 or S - 2) steps' must have landed.
  auto wait_all_but = [&](auto steps_c) {
    constexpr int n = decltype(steps_c)::value;
    if (h == 0) {
      WaitVmcnt<(AF + W::BH0) * n>();
    } else {
      WaitVmcnt<(AF + W::BH1) * n>();
    }
  };
  v4i fa0[RW][NP], fb0[CT], fa1[RW][NP], fb1[CT];
#pragma unroll
  for (int s = 0; s < S; ++s) {
#pragma unroll
    for (int j = 0; j < kDma; ++j) dma(s, j);
  }
  wait_all_but(std::integral_constant<int, S - 1>{});
  LdsBarrier();
#pragma unroll
  for (int j = 0; j < kFrag; ++j) frag(0, j, fa0, fb0);
  const int nkb2 = (nkb + 1) & ~1;
  for (int i = 0; i < nkb2; i += 2) {
    // Step i+1's DMAs have landed everywhere; step i's fragments are in
    // registers, so its stage is free for step i + S.
    wait_all_but(std::integral_constant<int, S - 2>{});
    LdsBarrier();
    step(i, fa0, fb0, fa1, fb1);
    wait_all_but(std::integral_constant<int, S - 2>{});
    LdsBarrier();
    step(i + 1, fa1, fb1, fa0, fb0);
  }
  WaitVmcnt<0>();
  LdsBarrier();
  u64* red = reinterpret_cast<u64*>(lds) + wave * (RW * 16 * kQMax);
  WaveEpilogueDigitMajor<RW, CT>(acc, red, lane, grp, nq, rows,
                                 (long long)(strip0 + h * RW) * 16, ks, partial);
}

// The streaming digit-major scans. The tall one (NP = 2, 7 digit tiles): each
// wave owns 6 strips (a workgroup 12 strips, 192 rows) by one 16-query group,
// so a workgroup moves 52 KiB of operands per k-block for 624 MFMAs where
// WideScanPipeKernel moves 44 KiB for 416, and each wave reads 19 fragments
// from LDS per 78 MFMAs instead of 15 per 52. The 168 accumulators leave no
// room for a second set of A fragments, so A is streamed one MFMA block ahead
// (one fragment per block) while B is read a whole step ahead, with one
// barrier per step, before its last block. Rings of SA and SB stages give
// leads of SA - 1 and SB - 1 steps. The same schedule serves four-plane
// matrices (NP = 4, 22 MFMAs per strip and group), whose 4 KiB A strips leave
// LDS for two or three A stages.
//
// GW query groups per wave: with 4 digit tiles (log q <= 32, 10 MFMAs per
// strip and group on four planes) one group's B tiles are few for the A
// bytes a strip brings in, so each wave takes two adjacent groups (a
// workgroup 8 groups, 128 queries): 4 strips x 2 groups moves 64 KiB per
// k-block for 640 MFMAs, where 4 strips x 1 group moves 48 KiB for 320.
template <int NP_, int CT, int RW_, int SA_, int SB_, int GW_ = 1>
struct StreamShape {
  static constexpr int NP = NP_, RW = RW_, RS = 2 * RW_, AF = RS * NP / 8;
  static_assert((RS * NP) % 8 == 0, "A fragments split evenly over 8 waves");
  static constexpr int GW = GW_;       // query groups per wave
  static constexpr int CW = GW_ * CT;  // B tiles per wave
  static constexpr int BH0 = (CW + 1) / 2, BH1 = CW - BH0;
  static constexpr int kA = RS * NP * 64;  // v4i of A per stage
  static constexpr int kB = 4 * CW * 64;   // v4i of B per stage
  static constexpr int SA = SA_, SB = SB_;
  // Within a step the ring with the shorter lead issues first, so that the
  // DMAs still allowed in flight at the step's barrier are the latest issued.
  static constexpr bool kBFirst = SB_ < SA_;
  static_assert((kBFirst && SB_ == 2) || (!kBFirst && SA_ == SB_), "in-order vmcnt waits");
  static constexpr size_t kRing = static_cast<size_t>(SA * kA + SB * kB) * 16;
  // One epilogue region per (wave, group): RW * 16 rows x kQMax u64.
  static constexpr int kRedWave = RW * 16 * kQMax;
  static constexpr size_t kRed = static_cast<size_t>(8) * GW * kRedWave * 8;
  static constexpr size_t kLdsBytes = kRing > kRed ? kRing : kRed;
};

template <int CT>
using TallShape = StreamShape<2, CT, 6, 3, 3>;

// vmcnt of the prologue's wait: the DMAs issued after the last of A(0), B(0)
// and B(1) (the operands step 0 reads) may stay in flight. Issue order: for
// s = 0, 1, ...: A(s) (s < SA - 1) and B(s) (s < SB), B first when kBFirst.
template <typename W>
constexpr int PrologueVmcnt(int bh) {
  int total = 0, needed_end = 0;
  const int steps = (W::SA - 1 > W::SB) ? W::SA - 1 : W::SB;
  for (int st = 0; st < steps; ++st) {
    for (int part = 0; part < 2; ++part) {
      const bool is_b = W::kBFirst ? part == 0 : part == 1;
      if (is_b && st < W::SB) {
        total += bh;
        if (st <= 1) needed_end = total;
      } else if (!is_b && st < W::SA - 1) {
        total += W::AF;
        if (st == 0) needed_end = total;
      }
    }
  }
  return total - needed_end;
}

template <int CT, typename W = TallShape<CT>>
__device__ __forceinline__ void WideScanTallBody(
    int w, const int8_t* __restrict__ tiles, int strips, int rows, int kblocks, int kbps,
    int ksplit, const int8_t* __restrict__ btiles, int ngp, int nq, u64* __restrict__ partial,
    v4i* lds) {
  constexpr int NP = W::NP, RS = W::RS, RW = W::RW, AF = W::AF, SA = W::SA, SB = W::SB;
  constexpr int GW = W::GW, CW = W::CW;
  const int nx = (strips + RS - 1) / RS;
  WideTile wt;
  if (!DecodeWideTile(w, nx, nx * ksplit, ngp / (4 * GW), &wt)) return;
  const int lane = threadIdx.x & 63;
  // Wave-uniform, so that addresses derived from it stay in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int h = wave >> 2;
  const int g = wave & 3;
  const int strip0 = wt.strip_block * RS;
  const int ks = wt.ks;
  // The wave's first group; its GW groups' B tiles are adjacent in btiles.
  const int grp = (wt.zb * 4 + g) * GW;
  const int kb0 = ks * kbps;
  const int nkb = min(kblocks, kb0 + kbps) - kb0;
  const v4i* A = reinterpret_cast<const v4i*>(tiles);
  const v4i* Bg = reinterpret_cast<const v4i*>(btiles) + (long long)grp * CT * 64 + lane;
  const v4i* Af[AF];
#pragma unroll
  for (int jf = 0; jf < AF; ++jf) {
    const int f = wave * AF + jf;
    Af[jf] = A + TileIndex(min(strip0 + f / NP, strips - 1), 0, f % NP, kblocks, NP) * 64 + lane;
  }
  const int c0 = h == 0 ? 0 : W::BH0;
  v4i* const ring_a = lds;
  v4i* const ring_b = lds + SA * W::kA;
  v4i acc[GW][RW][CT];
#pragma unroll
  for (int gi = 0; gi < GW; ++gi)
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[gi][r][c] = v4i{0, 0, 0, 0};

  // A fragment j of step s (steps past the k-range re-read the last k-block),
  // and this wave's B share cc of step s (past the range: the zero block).
  // A keeps the default cache policy: the four column-block workgroups of a
  // tile read it through one L2, and measured against the non-temporal hint
  // the 256-query scan ran 10.42 against 10.57 ms.
  auto dma_a = [&](int s, int j) {
    const long long kba = kb0 + min(s, nkb - 1);
    DmaLoad16(Af[j] + kba * kStripGroup * NP * 64,
              ring_a + (s % SA) * W::kA + (wave * AF + j) * 64);
  };
  auto dma_b = [&](int s, int cc) {
    if (h == 0 || cc < W::BH1) {
      const long long kbb = s < nkb ? kb0 + s : kblocks;
      const int c = c0 + cc;
      DmaLoad16(Bg + (kbb * ngp * CT + c) * 64, ring_b + (s % SB) * W::kB + (g * CW + c) * 64);
    }
  };
  auto read_a = [&](int s, int blk) -> v4i {
    const int r = blk / NP, p = blk % NP;
    return ring_a[(s % SA) * W::kA + ((h * RW + r) * NP + p) * 64 + lane];
  };
  auto read_b = [&](int s, int c) -> v4i {
    return ring_b[(s % SB) * W::kB + (g * CW + c) * 64 + lane];
  };
  constexpr int kBlk = RW * NP;      // MFMA blocks per step, one per (strip, plane)
  constexpr int kDma = AF + W::BH0;  // DMA slots per step (an h = 1 wave skips the last)
  static_assert(kDma <= kBlk && CW <= kBlk, "one DMA slot and one B read per block at most");
  // Own DMAs allowed in flight: na steps' A issues and nb steps' B issues.
  auto wait_own = [&](auto na_c, auto nb_c) {
    constexpr int na = decltype(na_c)::value, nb = decltype(nb_c)::value;
    if (h == 0) {
      WaitVmcnt<AF * na + W::BH0 * nb>();
    } else {
      WaitVmcnt<AF * na + W::BH1 * nb>();
    }
  };
  // Step i: B fragments of step i in cb, its first A fragment in a_in. Issues
  // A(i + SA - 1) and B(i + SB), reads step i + 1's B into nb and leaves its
  // first A fragment in a_in (the tall shape, SA = SB = 3: prologue A(0) B(0)
  // A(1) B(1) B(2), then per step A(i + 2) B(i + 3)). Each block's MFMAs are
  // issued after the read of the next block's A fragment, so that read lands
  // while they run. The step's one barrier, after the tail of its
  // next-to-last block (whose read of A(i)'s last fragment it waits for),
  // needs A(i + 1) and B(i + 2) landed everywhere (in flight only issues of
  // later steps), and orders every wave's reads of A(i) and B(i + 1) before
  // the next steps' DMAs that refill those slots. The read of A(i + 1)'s
  // first fragment after it then lands under the last block's MFMAs.
  // Measured at 256 queries: 2% faster than reading each fragment after the
  // previous block's MFMAs with the barrier after the last block, with
  // identical partial sums, and the query tiles' HBM traffic fell from about
  // 20 to 15 GB per dispatch.
  constexpr int kBar = kBlk - 2;
  static_assert(kBar >= kDma - 1 && kBar >= CW - 1,
                "the barrier follows the step's DMA issues and B reads");
  auto step = [&](int i, v4i (&cb)[CW], v4i (&nb)[CW], v4i& a_in) {
    v4i ab[2];
    ab[0] = a_in;
#pragma unroll
    for (int blk = 0; blk < kBlk; ++blk) {
      const int r = blk / NP, p = blk % NP;
      const v4i acur = ab[blk & 1];
      if (blk + 1 < kBlk) ab[(blk + 1) & 1] = read_a(i, blk + 1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int gi = 0; gi < GW; ++gi)
#pragma unroll
        for (int c = p; c < CT; ++c)
          acc[gi][r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(acur, cb[gi * CT + c - p],
                                                                acc[gi][r][c], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (blk < kDma) {
        if constexpr (W::kBFirst) {
          if (blk < W::BH0) {
            dma_b(i + SB, blk);
          } else {
            dma_a(i + SA - 1, blk - W::BH0);
          }
        } else {
          if (blk < AF) {
            dma_a(i + SA - 1, blk);
          } else {
            dma_b(i + SB, blk - AF);
          }
        }
      }
      if (blk < CW) nb[blk] = read_b(i + 1, blk);
      if (blk == kBar) {
        wait_own(std::integral_constant<int, SA - 2>{}, std::integral_constant<int, SB - 2>{});
        LdsBarrier();
        a_in = read_a(i + 1, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  v4i cb0[CW], cb1[CW], a_in;
  constexpr int kPro = (SA - 1 > SB) ? SA - 1 : SB;
#pragma unroll
  for (int st = 0; st < kPro; ++st) {
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const bool is_b = W::kBFirst ? part == 0 : part == 1;
      if (is_b && st < SB) {
#pragma unroll
        for (int cc = 0; cc < W::BH0; ++cc) dma_b(st, cc);
      } else if (!is_b && st < SA - 1) {
#pragma unroll
        for (int j = 0; j < AF; ++j) dma_a(st, j);
      }
    }
  }
  // A(0), B(0) and B(1) have landed everywhere.
  if (h == 0) {
    WaitVmcnt<PrologueVmcnt<W>(W::BH0)>();
  } else {
    WaitVmcnt<PrologueVmcnt<W>(W::BH1)>();
  }
  LdsBarrier();
#pragma unroll
  for (int c = 0; c < CW; ++c) cb0[c] = read_b(0, c);
  a_in = read_a(0, 0);
  // Every wave has read B(0) before step 0's B(SB) DMA refills its slot.
  LdsBarrier();
  const int nkb2 = (nkb + 1) & ~1;
  for (int i = 0; i < nkb2; i += 2) {
    step(i, cb0, cb1, a_in);
    step(i + 1, cb1, cb0, a_in);
  }
  WaitVmcnt<0>();
  LdsBarrier();
#pragma unroll
  for (int gi = 0; gi < GW; ++gi) {
    u64* red = reinterpret_cast<u64*>(lds) + (wave * GW + gi) * W::kRedWave;
    WaveEpilogueDigitMajor<RW, CT>(acc[gi], red, lane, grp + gi, nq, rows,
                                   (long long)(strip0 + h * RW) * 16, ks, partial);
  }
}

template <int CT>
__global__ __launch_bounds__(512) void WideScanTallKernel(
    const int8_t* __restrict__ tiles, int strips, int rows, int kblocks, int kbps, int ksplit,
    const int8_t* __restrict__ btiles, int ngp, int nq, u64* __restrict__ partial) {
  extern __shared__ v4i lds[];
  WideScanTallBody<CT>(blockIdx.x, tiles, strips, rows, kblocks, kbps, ksplit, btiles, ngp, nq,
                       partial, lds);
}

// The streaming schedule on four-plane matrices (shape W).
template <int NP, int CT, int RW, int SA, int SB, int GW = 1>
__global__ __launch_bounds__(512) void WideScanStreamKernel(
    const int8_t* __restrict__ tiles, int strips, int rows, int kblocks, int kbps, int ksplit,
    const int8_t* __restrict__ btiles, int ngp, int nq, u64* __restrict__ partial) {
  extern __shared__ v4i lds[];
  WideScanTallBody<CT, StreamShape<NP, CT, RW, SA, SB, GW>>(blockIdx.x, tiles, strips, rows,
                                                            kblocks, kbps, ksplit, btiles, ngp,
                                                            nq, partial, lds);
}

// One operand of a fused scan launch.
struct ScanOperand {
  const int8_t* tiles;
  int strips, rows, kblocks, kbps, ksplit;
  const int8_t* btiles;
  int ngp, nq, ns, qpg;
  u64* partial;
};

// The DB product's tall scan and the pack product's digit-major WideScan
// (four planes) as one launch: workgroups [0, db_blocks) are the DB scan's,
// in its own order, the rest the pack scan's. Workgroups start in launch
// order, so the pack product's take the CUs the DB product's last, partial
// round leaves idle instead of waiting for the whole DB product. db_blocks
// is a multiple of 8, so both keep their XCD placement (workgroup w on XCD
// w % 8).
__global__ __launch_bounds__(512) void DbPackScanKernel(ScanOperand db, ScanOperand pk,
                                                        int db_blocks) {
  extern __shared__ v4i lds[];
  const int w = static_cast<int>(blockIdx.x);
  if (w < db_blocks) {
    WideScanTallBody<7>(w, db.tiles, db.strips, db.rows, db.kblocks, db.kbps, db.ksplit,
                        db.btiles, db.ngp, db.nq, db.partial, lds);
  } else {
    WideScanBody<4, 7, true>(w - db_blocks, pk.tiles, pk.strips, pk.rows, pk.kblocks, pk.kbps,
                             pk.ksplit, pk.btiles, pk.ngp, pk.nq, pk.ns, pk.qpg, pk.partial, lds);
  }
}

constexpr size_t kFusedScanLds = TallShape<7>::kLdsBytes > WideShape<4, 7>::kLdsBytes
                                     ? TallShape<7>::kLdsBytes
                                     : WideShape<4, 7>::kLdsBytes;

int GridBlocks(long long total, int threads) {
  long long b = (total + threads - 1) / threads;
  if (b > 65535LL * 16) b = 65535LL * 16;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

constexpr int kNarrowR = 4;
constexpr int kNarrowD = 4;

template <int NP, int T>
hipError_t LaunchNarrowT(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                         const int8_t* btiles, u64* partial, hipStream_t s) {
  dim3 grid((L.strips + 4 * kNarrowR - 1) / (4 * kNarrowR), P.ksplit,
            P.group_rows > 0 ? 1 : P.ngp);
  hipLaunchKernelGGL((NarrowScanKernel<NP, kNarrowR, T, kNarrowD>), grid, dim3(256), 0, s,
                     tiles, L.strips, L.rows, L.kblocks, P.kb_per_split, btiles, P.ngp,
                     P.nq, P.ns, P.qpg, P.group_rows, partial);
  return hipGetLastError();
}

template <int NP, int CT, bool DM>
hipError_t LaunchWideT(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                       const int8_t* btiles, u64* partial, hipStream_t s) {
  using W = WideShape<NP, CT>;
  static_assert(W::kLdsBytes <= 160 * 1024, "WideScan LDS");
  const long long ntiles = static_cast<long long>((L.strips + W::RS - 1) / W::RS) * P.ksplit;
  const long long blocks = (ntiles + 7) / 8 * 8 * (P.ngp / 4);
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((WideScanKernel<NP, CT, DM>), dim3(static_cast<unsigned>(blocks)), dim3(512),
                     W::kLdsBytes, s, tiles, L.strips, L.rows, L.kblocks, P.kb_per_split,
                     P.ksplit, btiles, P.ngp, P.nq, P.ns, P.qpg, partial);
  return hipGetLastError();
}

template <int NP, int CT>
hipError_t LaunchWidePipeT(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                           const int8_t* btiles, u64* partial, hipStream_t s) {
  using W = WideShape<NP, CT>;
  const long long ntiles = static_cast<long long>((L.strips + W::RS - 1) / W::RS) * P.ksplit;
  const long long blocks = (ntiles + 7) / 8 * 8 * (P.ngp / 4);
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((WideScanPipeKernel<NP, CT>), dim3(static_cast<unsigned>(blocks)),
                     dim3(512), W::kLdsBytes, s, tiles, L.strips, L.rows, L.kblocks,
                     P.kb_per_split, P.ksplit, btiles, P.ngp, P.nq, partial);
  return hipGetLastError();
}

template <int CT>
hipError_t LaunchWideTallT(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                           const int8_t* btiles, u64* partial, hipStream_t s) {
  using W = TallShape<CT>;
  static_assert(W::kLdsBytes <= 160 * 1024, "WideScanTall LDS");
  const long long ntiles = static_cast<long long>((L.strips + W::RS - 1) / W::RS) * P.ksplit;
  const long long blocks = (ntiles + 7) / 8 * 8 * (P.ngp / 4);
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((WideScanTallKernel<CT>), dim3(static_cast<unsigned>(blocks)), dim3(512),
                     W::kLdsBytes, s, tiles, L.strips, L.rows, L.kblocks, P.kb_per_split,
                     P.ksplit, btiles, P.ngp, P.nq, partial);
  return hipGetLastError();
}

template <int NP, int CT, int RW, int SA, int SB, int GW = 1>
hipError_t LaunchWideStreamT(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                             const int8_t* btiles, u64* partial, hipStream_t s) {
  using W = StreamShape<NP, CT, RW, SA, SB, GW>;
  static_assert(W::kLdsBytes <= 160 * 1024, "WideScanStream LDS");
  if (P.ngp % (4 * GW) != 0) return hipErrorInvalidValue;
  const long long ntiles = static_cast<long long>((L.strips + W::RS - 1) / W::RS) * P.ksplit;
  const long long blocks = (ntiles + 7) / 8 * 8 * (P.ngp / (4 * GW));
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((WideScanStreamKernel<NP, CT, RW, SA, SB, GW>),
                     dim3(static_cast<unsigned>(blocks)), dim3(512), W::kLdsBytes, s, tiles,
                     L.strips, L.rows, L.kblocks, P.kb_per_split, P.ksplit, btiles, P.ngp, P.nq,
                     partial);
  return hipGetLastError();
}

template <int NP>
hipError_t LaunchScanNP(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                        const int8_t* btiles, u64* partial, hipStream_t s) {
  if (!P.wide) {
    switch (P.ctg) {
      case 1: return LaunchNarrowT<NP, 1>(tiles, L, P, btiles, partial, s);
      case 2: return LaunchNarrowT<NP, 2>(tiles, L, P, btiles, partial, s);
      case 4: return LaunchNarrowT<NP, 4>(tiles, L, P, btiles, partial, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (P.ctg) {
    case 2: return LaunchWideT<NP, 2, false>(tiles, L, P, btiles, partial, s);
    case 4:
      if (!P.digit_major) return LaunchWideT<NP, 4, false>(tiles, L, P, btiles, partial, s);
      if constexpr (NP == 4) {
        // Two 16-query groups per wave (128 queries per workgroup), 6 or 8
        // strips per workgroup.
        if (P.tall == 1) return LaunchWideStreamT<4, 4, 3, 3, 2, 2>(tiles, L, P, btiles, partial, s);
        if (P.tall == 2) return LaunchWideStreamT<4, 4, 4, 2, 2, 2>(tiles, L, P, btiles, partial, s);
      }
      if constexpr (NP <= 2) {
        return LaunchWidePipeT<NP, 4>(tiles, L, P, btiles, partial, s);
      } else {
        return LaunchWideT<NP, 4, true>(tiles, L, P, btiles, partial, s);
      }
    case 7:
      if (!P.digit_major) return LaunchWideT<NP, 7, false>(tiles, L, P, btiles, partial, s);
      if constexpr (NP == 2) {
        if (P.tall) return LaunchWideTallT<7>(tiles, L, P, btiles, partial, s);
      }
      if constexpr (NP == 4) {
        if (P.tall == 1) return LaunchWideStreamT<4, 7, 5, 2, 2>(tiles, L, P, btiles, partial, s);
        if (P.tall == 2) return LaunchWideStreamT<4, 7, 4, 3, 2>(tiles, L, P, btiles, partial, s);
      }
      // Two register sets of fragments fit beside the accumulators for NP <= 2.
      if constexpr (NP <= 2) {
        return LaunchWidePipeT<NP, 7>(tiles, L, P, btiles, partial, s);
      } else {
        return LaunchWideT<NP, 7, true>(tiles, L, P, btiles, partial, s);
      }
    default: return hipErrorInvalidValue;
  }
}

}  // namespace


hipError_t LaunchDbToTiles(const u64* cdata, int db_bytes, int rows, int cols,
                           int row0, int nrows, const LimbLayout& L,
                           int8_t* tiles, hipStream_t s) {
  const long long total =
      (long long)((row0 + nrows + 15) / 16 - row0 / 16) * L.kblocks * 64;
  hipLaunchKernelGGL(DbToTilesKernel, dim3(GridBlocks(total, 256)), dim3(256), 0,
                     s, cdata, db_bytes, rows, cols, row0, nrows, L.strips,
                     L.kblocks, L.planes, tiles);
  return hipGetLastError();
}

hipError_t LaunchPackToTiles(const void* pack, int mat_bytes, int rows,
                             int cols, int row0, int nrows,
                             const LimbLayout& L, int8_t* tiles,
                             hipStream_t s) {
  const long long total =
      (long long)((row0 + nrows + 15) / 16 - row0 / 16) * L.kblocks * 64;
  hipLaunchKernelGGL(PackToTilesKernel, dim3(GridBlocks(total, 256)), dim3(256),
                     0, s, pack, mat_bytes, rows, cols, row0, nrows, L.strips,
                     L.kblocks, L.planes, tiles);
  return hipGetLastError();
}

namespace {

// 16 bytes per lane and step; words an odd count leaves to lane 0.
__global__ __launch_bounds__(256) void CopyWordsKernel(const u64* __restrict__ src,
                                                       u64* __restrict__ dst, long long pairs,
                                                       long long words) {
  const long long stride = static_cast<long long>(gridDim.x) * 256;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < pairs; i += stride) {
    const ulonglong2 v = reinterpret_cast<const ulonglong2*>(src)[i];
    reinterpret_cast<ulonglong2*>(dst)[i] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (words & 1)) dst[words - 1] = src[words - 1];
}

}  // namespace

hipError_t LaunchCopyWords(const u64* src, u64* dst, size_t words, hipStream_t s) {
  if (words == 0) return hipSuccess;
  // Both ends 16-byte aligned (allocations and record strides are).
  if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) {
    return hipErrorInvalidValue;
  }
  const long long pairs = static_cast<long long>(words / 2);
  const int blocks = static_cast<int>(std::min<long long>(1024, std::max<long long>(1, (pairs + 255) / 256)));
  hipLaunchKernelGGL(CopyWordsKernel, dim3(blocks), dim3(256), 0, s, src, dst, pairs,
                     static_cast<long long>(words));
  return hipGetLastError();
}

hipError_t LaunchTilesToValues(const int8_t* tiles, const LimbLayout& L, int row0, int nrows,
                              bool signed_top, u64 mask, u64* out, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  if (row0 < 0 || row0 + nrows > L.rows) return hipErrorInvalidValue;
  const long long total = static_cast<long long>(nrows) * L.kblocks * 4;
  hipLaunchKernelGGL(TilesToValuesKernel, dim3(GridBlocks(total, 256)), dim3(256), 0, s,
                     tiles, L.cols, row0, nrows, L.kblocks, L.planes, signed_top ? 1 : 0, mask,
                     out);
  return hipGetLastError();
}

template <int NP, int CT, bool DM = false>
hipError_t ConfigureWide() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(WideScanKernel<NP, CT, DM>),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(WideShape<NP, CT>::kLdsBytes));
}

template <int CT>
hipError_t ConfigureWideTall() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(WideScanTallKernel<CT>),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(TallShape<CT>::kLdsBytes));
}

template <int NP, int CT, int RW, int SA, int SB, int GW = 1>
hipError_t ConfigureWideStream() {
  return hipFuncSetAttribute(
      reinterpret_cast<const void*>(WideScanStreamKernel<NP, CT, RW, SA, SB, GW>),
      hipFuncAttributeMaxDynamicSharedMemorySize,
      static_cast<int>(StreamShape<NP, CT, RW, SA, SB, GW>::kLdsBytes));
}

template <int NP, int CT>
hipError_t ConfigureWidePipe() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(WideScanPipeKernel<NP, CT>),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(WideShape<NP, CT>::kLdsBytes));
}

// The WideScan instantiations need more than the default 64 KiB of dynamic
// LDS; set once, outside any stream capture.
hipError_t ConfigureScanKernels() {
  const hipError_t es[] = {ConfigureWide<1, 2>(), ConfigureWide<1, 4>(), ConfigureWide<1, 7>(),
                           ConfigureWide<2, 2>(), ConfigureWide<2, 4>(), ConfigureWide<2, 7>(),
                           ConfigureWide<4, 2>(), ConfigureWide<4, 4>(), ConfigureWide<4, 7>(),
                           ConfigureWide<4, 7, true>(), ConfigureWidePipe<1, 7>(),
                           ConfigureWidePipe<2, 7>(), ConfigureWide<4, 4, true>(),
                           ConfigureWidePipe<1, 4>(), ConfigureWidePipe<2, 4>(),
                           ConfigureWideTall<7>(), ConfigureWideStream<4, 7, 5, 2, 2>(),
                           ConfigureWideStream<4, 7, 4, 3, 2>(),
                           ConfigureWideStream<4, 4, 4, 2, 2, 2>(),
                           ConfigureWideStream<4, 4, 3, 3, 2, 2>(),
                           hipFuncSetAttribute(reinterpret_cast<const void*>(DbPackScanKernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(kFusedScanLds))};
  for (hipError_t e : es) {
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int QueriesPerGroup(int tiles, int ns) {
  const int q = tiles * 16 / ns;
  return q < kQMax ? q : kQMax;
}

// Tile widths the kernels are instantiated for.
constexpr int kNarrowTiles[] = {1, 2, 4};
constexpr int kWideTiles[] = {2, 4, 7};

// Query groups per wide workgroup: 8 for the streaming scans of four-plane
// matrices with 4 digit tiles (two groups per wave), 4 otherwise.
int WideGroupsPerWorkgroup(const LimbLayout& L, const ScanPlan& P) {
  return P.wide && P.digit_major && P.ns == 4 && L.planes == 4 && P.tall != 0 ? 8 : 4;
}

// Finishes a plan whose kind and tile width are set: groups, K-split and
// B-tile size. ks_scale multiplies the workgroup target of the K-split.
void CompletePlan(const LimbLayout& L, int ks_scale, ScanPlan* P) {
  P->qpg = P->digit_major ? 16 : QueriesPerGroup(P->ctg, P->ns);
  const int groups = (P->nq + P->qpg - 1) / P->qpg;
  const int gpw = WideGroupsPerWorkgroup(L, *P);
  P->ngp = P->wide ? (groups + gpw - 1) / gpw * gpw : groups;
  // K-split: exact int32 sums need at most 65536 / planes columns per split;
  // beyond that, enough workgroups to occupy every CU a few times over.
  const int max_kb = kMaxColsPerSplitNP1 / L.planes / 64;
  // Strips per wide workgroup (the streaming shapes of LaunchScanNP).
  const int wide_rows = !P->digit_major || P->tall == 0 ? 8
                        : P->ns == 4                   ? (P->tall == 1 ? 6 : 8)
                        : P->tall == 1                 ? (L.planes == 4 ? 10 : 12)
                                                       : 8;
  const int row_wgs = P->wide ? (L.strips + wide_rows - 1) / wide_rows
                              : (L.strips + 4 * kNarrowR - 1) / (4 * kNarrowR);
  const int z = P->wide ? P->ngp / gpw : P->ngp;
  const int target = (P->wide ? 512 : 2048) * std::max(ks_scale, 1);
  int ks = (target + row_wgs * z - 1) / (row_wgs * z);
  const int min_split = (L.kblocks + max_kb - 1) / max_kb;
  if (ks < min_split) ks = min_split;
  if (ks > L.kblocks) ks = L.kblocks;
  if (ks < 1) ks = 1;
  P->kb_per_split = (L.kblocks + ks - 1) / ks;
  P->ksplit = (L.kblocks + P->kb_per_split - 1) / P->kb_per_split;
  P->btile_bytes = static_cast<size_t>(L.kblocks + 1) * P->ngp * P->ctg * kTileBytes;
}

// Digit-major WideScan is instantiated for 7 shifts per query (log q in
// (48, 56]) and for 4 (log q in (24, 32]).
inline bool DigitMajorShifts(int ns) { return ns == 7 || ns == 4; }

std::vector<ScanPlan> ScanPlanCandidates(const LimbLayout& L, int nq, int out_bits) {
  std::vector<ScanPlan> out;
  const int ns = (out_bits + 7) / 8;
  auto add = [&](int wide, int t, int dm, int tall) {
    for (int ks_scale : {1, 2}) {
      ScanPlan P{};
      P.wide = wide;
      P.nq = nq;
      P.ns = ns;
      P.ctg = t;
      P.digit_major = dm;
      P.tall = tall;
      CompletePlan(L, ks_scale, &P);
      bool dup = false;
      for (const ScanPlan& o : out) {
        dup = dup || (o.wide == P.wide && o.ctg == P.ctg && o.ksplit == P.ksplit &&
                      o.digit_major == P.digit_major && o.tall == P.tall);
      }
      if (!dup) out.push_back(P);
    }
  };
  for (int wide = 0; wide < 2; ++wide) {
    for (int t : (wide ? std::vector<int>{2, 4, 7} : std::vector<int>{1, 2, 4})) {
      // A narrow workgroup re-reads A for every group: keep those with few groups.
      if (!wide && (nq + QueriesPerGroup(t, ns) - 1) / QueriesPerGroup(t, ns) > 2) continue;
      add(wide, t, 0, 0);
    }
  }
  // Digit-major groups hold 16 queries: worth trying from two groups on.
  if (DigitMajorShifts(ns) && nq > 16) {
    add(1, ns, 1, 0);
    if (L.planes == 2 && ns == 7) add(1, ns, 1, 1);
    if (L.planes == 4) {
      add(1, ns, 1, 1);
      add(1, ns, 1, 2);
    }
  }
  return out;
}

ScanPlan MakeScanPlan(const LimbLayout& L, int nq, int out_bits, int wide_min_queries,
                      bool allow_digit_major, bool allow_tall, int four_plane_stream) {
  ScanPlan P{};
  P.nq = nq;
  P.ns = (out_bits + 7) / 8;
  const int ns = P.ns;
  // Narrow when the queries fit one group of at most 4 tiles and the batch is
  // below the wide threshold.
  int narrow_tiles = 0;
  for (int t : kNarrowTiles) {
    if (QueriesPerGroup(t, ns) >= nq) {
      narrow_tiles = t;
      break;
    }
  }
  if (narrow_tiles > 0 && nq < wide_min_queries) {
    P.wide = 0;
    P.ctg = narrow_tiles;
  } else if (allow_digit_major && DigitMajorShifts(ns) && nq > 32) {
    // Digit-major: no plane shifts and (2*ns - 1)/(2*ns) of the MFMAs, at a
    // granularity of 64 queries per column block.
    P.wide = 1;
    P.ctg = ns;
    P.digit_major = 1;
    // Four planes, by default: the 8-strip streaming scan with 7 digit
    // tiles; with 4 its two-groups-per-wave form, whose 128-query workgroups
    // pay from about 100 queries on (at 64 the plain WideScan took 3.2 ms
    // against 4.5 ms for the streaming form's half-empty workgroups, at 256
    // 10.7 against 8.3 ms).
    const int fps = four_plane_stream >= 0 ? four_plane_stream
                    : ns == 7              ? 2
                    : nq > 96              ? 2
                                           : 0;
    P.tall = L.planes == 4 ? fps : ns == 7 && L.planes == 2 ? (allow_tall ? 1 : 0) : 0;
  } else {
    P.wide = 1;
    P.ctg = kWideTiles[2];
    for (int t : kWideTiles) {
      if (4 * QueriesPerGroup(t, ns) >= nq) {
        P.ctg = t;
        break;
      }
    }
  }
  CompletePlan(L, 1, &P);
  return P;
}

ScanPlan MakeGroupedScanPlan(const LimbLayout& L, int ngroups, int group_rows, int out_bits) {
  ScanPlan P{};
  P.wide = 0;
  P.nq = ngroups;
  P.ns = (out_bits + 7) / 8;
  P.ctg = 1;
  P.qpg = 1;
  P.ngp = ngroups;
  P.group_rows = group_rows;
  const int max_kb = kMaxColsPerSplitNP1 / L.planes / 64;
  const int row_wgs = (L.strips + 4 * kNarrowR - 1) / (4 * kNarrowR);
  int ks = (2048 + row_wgs - 1) / row_wgs;
  const int min_split = (L.kblocks + max_kb - 1) / max_kb;
  if (ks < min_split) ks = min_split;
  if (ks > L.kblocks) ks = L.kblocks;
  if (ks < 1) ks = 1;
  P.kb_per_split = (L.kblocks + ks - 1) / ks;
  P.ksplit = (L.kblocks + P.kb_per_split - 1) / P.kb_per_split;
  P.btile_bytes = static_cast<size_t>(L.kblocks + 1) * P.ngp * P.ctg * kTileBytes;
  return P;
}

hipError_t LaunchQueryDigits(const u64* vec, int stride, int len, const LimbLayout& L,
                             const ScanPlan& P, int8_t* btiles, hipStream_t s) {
  // kblocks + 1 blocks: the last one is all zero (its k's are >= len).
  if (P.digit_major) {
    const long long total = (long long)(L.kblocks + 1) * P.ngp * 64;
    hipLaunchKernelGGL(QueryDigitsDigitMajorKernel, dim3(GridBlocks(total, 256)), dim3(256), 0,
                       s, vec, P.nq, stride, len, L.kblocks + 1, P.ns, P.ngp, btiles);
    return hipGetLastError();
  }
  const long long total = (long long)(L.kblocks + 1) * P.ngp * P.ctg * 64;
  hipLaunchKernelGGL(QueryDigitsKernel, dim3(GridBlocks(total, 256)), dim3(256), 0, s, vec,
                     P.nq, stride, len, L.kblocks + 1, P.ns, P.qpg, P.ngp, P.ctg,
                     P.digit_major, btiles);
  return hipGetLastError();
}

int VectorSumScratchWords(int nq) { return nq * kVectorSumChunks; }

int DigitSumChunks(const LimbLayout& L) {
  return (L.kblocks + 1 + kDigitSumWaves - 1) / kDigitSumWaves;
}

hipError_t LaunchQueryDigitsSums(const u64* vec, int stride, int len, const LimbLayout& L,
                                 const ScanPlan& P, int8_t* btiles, u64 c_off, u64* scratch,
                                 u64* corr, hipStream_t s) {
  if (!P.digit_major) {
    hipError_t e = LaunchQueryDigits(vec, stride, len, L, P, btiles, s);
    if (e != hipSuccess) return e;
    return LaunchVectorSums(vec, P.nq, stride, len, c_off, scratch, corr, s);
  }
  // kblocks + 1 blocks (the last one is all zero), one per wave: the
  // scratch holds nq x DigitSumChunks(L) sums.
  const int kbt = L.kblocks + 1;
  const int nchunks = DigitSumChunks(L);
  hipLaunchKernelGGL(QueryDigitsSumsKernel, dim3(nchunks, P.ngp), dim3(64 * kDigitSumWaves), 0, s,
                     vec, P.nq, stride, len, kbt, P.ns, P.ngp, kDigitSumWaves, nchunks, btiles,
                     scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ChunkSumFinalKernel, dim3(P.nq), dim3(64), 0, s, scratch, nchunks, c_off,
                     corr);
  return hipGetLastError();
}

hipError_t LaunchVectorSums(const u64* vec, int nq, int stride, int len, u64 c_off,
                            u64* scratch, u64* corr, hipStream_t s) {
  const int chunk = (len + kVectorSumChunks - 1) / kVectorSumChunks;
  const int nchunks = chunk > 0 ? (len + chunk - 1) / chunk : 1;
  hipLaunchKernelGGL(VectorPartialSumKernel, dim3(nchunks, nq), dim3(256), 0, s, vec,
                     stride, len, chunk > 0 ? chunk : 1, nchunks, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(VectorSumFinalKernel, dim3((nq + 63) / 64), dim3(64), 0, s, scratch,
                     nchunks, nq, c_off, corr);
  return hipGetLastError();
}

hipError_t LaunchScan(const int8_t* tiles, const LimbLayout& L, const ScanPlan& P,
                      const int8_t* btiles, u64* partial, hipStream_t s) {
  if (P.nq <= 0) return hipSuccess;
  if (P.qpg < 1 || P.ngp < 1) return hipErrorInvalidValue;
  if (P.digit_major && (!P.wide || !DigitMajorShifts(P.ctg) || P.ns != P.ctg ||
                        P.qpg != 16 || P.ngp % WideGroupsPerWorkgroup(L, P) != 0)) {
    return hipErrorInvalidValue;
  }
  if (P.group_rows > 0 && (P.wide || P.group_rows % (16 * kNarrowR) != 0 ||
                           static_cast<long long>(P.group_rows) * P.ngp < L.rows)) {
    return hipErrorInvalidValue;
  }
  switch (L.planes) {
    case 1: return LaunchScanNP<1>(tiles, L, P, btiles, partial, s);
    case 2: return LaunchScanNP<2>(tiles, L, P, btiles, partial, s);
    case 4: return LaunchScanNP<4>(tiles, L, P, btiles, partial, s);
    default: return hipErrorInvalidValue;
  }
}

bool ScanPairFusable(const LimbLayout& Ld, const ScanPlan& Pd, const LimbLayout& Lp,
                     const ScanPlan& Pp) {
  return Pd.nq > 0 && Pp.nq > 0 && Pd.wide && Pd.digit_major && Pd.tall == 1 && Pp.tall == 0 &&
         Pd.ctg == 7 &&
         Pd.ns == 7 && Pd.qpg == 16 && Pd.ngp % 4 == 0 && Pd.group_rows == 0 &&
         Ld.planes == 2 && Pp.wide && Pp.digit_major && Pp.ctg == 7 && Pp.ns == 7 &&
         Pp.qpg == 16 && Pp.ngp % 4 == 0 && Pp.group_rows == 0 && Lp.planes == 4;
}

long long ScanWorkgroups(const LimbLayout& L, const ScanPlan& P) {
  const long long tiles =
      static_cast<long long>((L.strips + TallShape<7>::RS - 1) / TallShape<7>::RS) * P.ksplit;
  return (tiles + 7) / 8 * 8 * (P.ngp / 4);
}

hipError_t LaunchScanPair(const int8_t* db_tiles, const LimbLayout& Ld, const ScanPlan& Pd,
                          const int8_t* db_btiles, u64* db_partial, const int8_t* pk_tiles,
                          const LimbLayout& Lp, const ScanPlan& Pp, const int8_t* pk_btiles,
                          u64* pk_partial, hipStream_t s) {
  if (!ScanPairFusable(Ld, Pd, Lp, Pp)) return hipErrorInvalidValue;
  const long long db_blocks = ScanWorkgroups(Ld, Pd);
  const long long pk_tiles_n =
      static_cast<long long>((Lp.strips + WideShape<4, 7>::RS - 1) / WideShape<4, 7>::RS) *
      Pp.ksplit;
  const long long pk_blocks = (pk_tiles_n + 7) / 8 * 8 * (Pp.ngp / 4);
  if (db_blocks + pk_blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const ScanOperand db{db_tiles, Ld.strips, Ld.rows, Ld.kblocks, Pd.kb_per_split, Pd.ksplit,
                       db_btiles, Pd.ngp, Pd.nq, Pd.ns, Pd.qpg, db_partial};
  const ScanOperand pk{pk_tiles, Lp.strips, Lp.rows, Lp.kblocks, Pp.kb_per_split, Pp.ksplit,
                       pk_btiles, Pp.ngp, Pp.nq, Pp.ns, Pp.qpg, pk_partial};
  hipLaunchKernelGGL(DbPackScanKernel, dim3(static_cast<unsigned>(db_blocks + pk_blocks)),
                     dim3(512), kFusedScanLds, s, db, pk, static_cast<int>(db_blocks));
  return hipGetLastError();
}

}  // namespace pirg
