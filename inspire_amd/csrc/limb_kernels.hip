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
#include "pirg_internal.h"

namespace pirg {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kWaves = 4;        // waves per workgroup in the scan kernel
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
      int8_t* dst = tiles + (((strip * kblocks + kb) * planes + p) * 64 + lane) * 16;
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
      int8_t* dst = tiles + (((strip * kblocks + kb) * planes + p) * 64 + lane) * 16;
      *reinterpret_cast<int4*>(dst) = *reinterpret_cast<const int4*>(out);
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

// B operand tiles. For k-block kb, 16-column tile `tile` and plane p, column
// j = (query ql, shift s) holds digit (s - p) of the query (zero when s < p),
// so that plane p's product lands in the accumulator of total shift 8s.
// Layout [kb][tile][plane][64 lanes][16 bytes]. One thread per (kb, tile, lane).
__global__ void VectorsToTilesKernel(const u64* __restrict__ vec, int nq,
                                     int stride, int len, int kblocks, int nl,
                                     int qpt, int nt, int planes,
                                     int8_t* __restrict__ btiles) {
  const long long total = (long long)kblocks * nt * 64;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int lane = static_cast<int>(g & 63);
    const long long t = g >> 6;
    const int tile = static_cast<int>(t % nt);
    const int kb = static_cast<int>(t / nt);
    const int j = lane & 15;
    const int ql = j / nl, sh = j % nl;
    const int q = tile * qpt + ql;
    const bool valid = ql < qpt && q < nq;
    const int kbase = kb * 64 + 16 * (lane >> 4);
    u64 ys[16];
    for (int b = 0; b < 16; ++b) {
      const int k = kbase + b;
      ys[b] = (valid && k < len) ? vec[(long long)q * stride + k] : 0ULL;
    }
    for (int p = 0; p < planes; ++p) {
      int8_t out[16];
      const int c = sh - p;
      for (int b = 0; b < 16; ++b) {
        out[b] = (c >= 0) ? static_cast<int8_t>(BalancedDigit(ys[b], c)) : int8_t{0};
      }
      int8_t* dst = btiles + ((((long long)kb * nt + tile) * planes + p) * 64 + lane) * 16;
      *reinterpret_cast<int4*>(dst) = *reinterpret_cast<const int4*>(out);
    }
  }
}

// corr[q] = c_off * sum_k vec[q][k] mod 2^64. One block per vector.
__global__ void VectorSumKernel(const u64* __restrict__ vec, int stride,
                                int len, u64 c_off, u64* __restrict__ corr) {
  __shared__ u64 part[256];
  const int q = blockIdx.x;
  u64 s = 0;
  for (int k = threadIdx.x; k < len; k += blockDim.x) s += vec[(long long)q * stride + k];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) corr[q] = c_off * part[0];
}

// The scan. A workgroup of 4 waves owns 4*R consecutive 16-row strips and a
// k-range; wave w owns strips [strip0 + w*R, +R). Per 64-column block:
//  * the T*NP B tiles of the block are staged in LDS (double buffered) by the
//    whole workgroup and shared by its 4 waves;
//  * each wave holds its R*NP A tiles in registers, loaded one block ahead;
//  * plane p of strip r times B tile (t, p) accumulates into acc[r][t]: the
//    B tile of plane p carries the query digits shifted by p, so one int32
//    accumulator per (row, query, total shift) collects every (plane, digit)
//    pair of that shift.
// |acc| <= NP * 2^14 * columns, so a k-range of 2^16 / NP columns stays exact.
template <int NP, int R, int T>
__global__ __launch_bounds__(256) void LimbScanKernel(
    const int8_t* __restrict__ tiles, int strips, int rows, int kblocks,
    int kb_per_split, const int8_t* __restrict__ btiles, int nt, int nq,
    int nl, int qpt, int out_bits, u64* __restrict__ partial) {
  __shared__ v4i bsh[2][T * NP * 64];
  __shared__ u64 red[kWaves][16][16];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int strip0 = (blockIdx.x * kWaves + wave) * R;
  const int ks = blockIdx.y;
  const int nt0 = blockIdx.z * T;
  const int kb0 = ks * kb_per_split;
  const int kb1 = min(kblocks, kb0 + kb_per_split);
  const v4i* A = reinterpret_cast<const v4i*>(tiles);
  const v4i* Bt = reinterpret_cast<const v4i*>(btiles);
  constexpr int kBVecs = T * NP * 64;             // v4i per staged block
  constexpr int kBPerThread = (kBVecs + 255) / 256;

  v4i acc[R][T];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int tt = 0; tt < T; ++tt) acc[r][tt] = v4i{0, 0, 0, 0};

  auto load_a = [&](int kb, v4i (&a)[R][NP]) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int strip = strip0 + r;
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        a[r][p] = (strip < strips)
                      ? A[(((long long)strip * kblocks + kb) * NP + p) * 64 + lane]
                      : v4i{0, 0, 0, 0};
      }
    }
  };
  // B tiles of block kb for tiles [nt0, nt0 + T): contiguous in btiles when
  // all T tiles exist; missing tiles read as zero.
  auto load_b = [&](int kb, v4i (&tmp)[kBPerThread]) {
#pragma unroll
    for (int i = 0; i < kBPerThread; ++i) {
      const int v = threadIdx.x + i * 256;
      const int tt = v / (NP * 64);
      v4i x = v4i{0, 0, 0, 0};
      if (v < kBVecs && nt0 + tt < nt) {
        x = Bt[((long long)kb * nt + nt0) * NP * 64 + v];
      }
      tmp[i] = x;
    }
  };
  auto store_b = [&](int buf, const v4i (&tmp)[kBPerThread]) {
#pragma unroll
    for (int i = 0; i < kBPerThread; ++i) {
      const int v = threadIdx.x + i * 256;
      if (v < kBVecs) bsh[buf][v] = tmp[i];
    }
  };

  if (kb0 < kb1) {
    v4i a_cur[R][NP], a_next[R][NP];
    v4i btmp[kBPerThread];
    load_a(kb0, a_cur);
    load_b(kb0, btmp);
    store_b(0, btmp);
    __syncthreads();
    for (int kb = kb0; kb < kb1; ++kb) {
      const int buf = (kb - kb0) & 1;
      const bool more = kb + 1 < kb1;
      if (more) {
        load_a(kb + 1, a_next);
        load_b(kb + 1, btmp);
      }
#pragma unroll
      for (int tt = 0; tt < T; ++tt) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const v4i b = bsh[buf][(tt * NP + p) * 64 + lane];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            acc[r][tt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a_cur[r][p], b, acc[r][tt],
                                                               0, 0, 0);
          }
        }
      }
      if (more) store_b(buf ^ 1, btmp);
      __syncthreads();
      if (more) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int p = 0; p < NP; ++p) a_cur[r][p] = a_next[r][p];
      }
    }
  }

  // Epilogue. MFMA 16x16 output layout: lane l, element v holds
  // C[row = 4*(l/16) + v][col = l % 16]; col = (query, shift).
  const int col = lane & 15;
  const int sh = 8 * (col % nl);
  const bool col_ok = col < qpt * nl && sh < out_bits;
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int tt = 0; tt < T; ++tt) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const u64 val = col_ok ? static_cast<u64>(static_cast<i64>(acc[r][tt][v])) << sh : 0ULL;
        red[wave][4 * (lane >> 4) + v][col] = val;
      }
      __syncthreads();
      if (lane < 16 * qpt) {
        const int row = lane & 15;
        const int ql = lane >> 4;
        u64 sum = 0;
        for (int cc = 0; cc < nl; ++cc) sum += red[wave][row][ql * nl + cc];
        const int q = (nt0 + tt) * qpt + ql;
        const long long grow = (long long)(strip0 + r) * 16 + row;
        if (q < nq && grow < rows) {
          partial[((long long)ks * nq + q) * rows + grow] = sum;
        }
      }
      __syncthreads();
    }
  }
}

int PickT(int nt) {
  if (nt <= 1) return 1;
  if (nt <= 2) return 2;
  if (nt <= 4) return 4;
  if (nt <= 8) return 8;
  return 16;
}

int GridBlocks(long long total, int threads) {
  long long b = (total + threads - 1) / threads;
  if (b > 65535LL * 16) b = 65535LL * 16;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

template <int NP, int R, int T>
hipError_t LaunchScanT(const int8_t* tiles, const LimbLayout& L,
                       const int8_t* btiles, int nt, int nq, int nl, int qpt,
                       int out_bits, int ksplit, u64* partial, hipStream_t s) {
  const int kbps = (L.kblocks + ksplit - 1) / ksplit;
  dim3 grid((L.strips + kWaves * R - 1) / (kWaves * R), ksplit, (nt + T - 1) / T);
  hipLaunchKernelGGL((LimbScanKernel<NP, R, T>), grid, dim3(256), 0, s, tiles,
                     L.strips, L.rows, L.kblocks, kbps, btiles, nt, nq, nl, qpt,
                     out_bits, partial);
  return hipGetLastError();
}

// R strips per wave; R * T * 4 accumulator registers stay <= 128 so the
// wave fits its registers without spilling.
template <int NP>
hipError_t LaunchScanNP(const int8_t* tiles, const LimbLayout& L,
                        const int8_t* btiles, int nt, int nq, int nl, int qpt,
                        int out_bits, int ksplit, u64* partial, hipStream_t s) {
  switch (PickT(nt)) {
    case 1:
      return LaunchScanT<NP, 4, 1>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit, partial, s);
    case 2:
      return LaunchScanT<NP, 4, 2>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit, partial, s);
    case 4:
      return LaunchScanT<NP, 4, 4>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit, partial, s);
    case 8:
      return LaunchScanT<NP, 4, 8>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit, partial, s);
    default:
      return LaunchScanT<NP, 2, 16>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit,
                                    partial, s);
  }
}

}  // namespace

// Plain streaming copy (one 16-byte load and store per element); the
// reference kernel for calibrating HBM byte counters.
__global__ void CopyProbeKernel(const int4* __restrict__ src, int4* __restrict__ dst,
                                long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    dst[i] = src[i];
  }
}

hipError_t LaunchCopyProbe(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const long long n = static_cast<long long>(bytes / 16);
  hipLaunchKernelGGL(CopyProbeKernel, dim3(4096), dim3(256), 0, s,
                     static_cast<const int4*>(src), static_cast<int4*>(dst), n);
  return hipGetLastError();
}

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

hipError_t LaunchVectorsToTiles(const u64* vec, int nq, int stride, int len,
                                int kblocks, int nl, int qpt, int planes, u64 c_off,
                                int8_t* btiles, u64* corr, hipStream_t s) {
  const int nt = (nq + qpt - 1) / qpt;
  const long long total = (long long)kblocks * nt * 64;
  hipLaunchKernelGGL(VectorsToTilesKernel, dim3(GridBlocks(total, 256)),
                     dim3(256), 0, s, vec, nq, stride, len, kblocks, nl, qpt,
                     nt, planes, btiles);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (corr != nullptr) {
    hipLaunchKernelGGL(VectorSumKernel, dim3(nq), dim3(256), 0, s, vec, stride,
                       len, c_off, corr);
    e = hipGetLastError();
  }
  return e;
}

int LimbScanKSplit(const LimbLayout& L, int nq) {
  (void)nq;
  const int R = 4;  // strips per wave of the common configurations
  const int wgs = (L.strips + kWaves * R - 1) / (kWaves * R);
  int ksplit = (1024 + wgs - 1) / wgs;
  const int max_kb = kMaxColsPerSplitNP1 / L.planes / 64;
  const int min_split = (L.kblocks + max_kb - 1) / max_kb;
  if (ksplit < min_split) ksplit = min_split;
  if (ksplit > L.kblocks) ksplit = L.kblocks;
  if (ksplit < 1) ksplit = 1;
  const int kbps = (L.kblocks + ksplit - 1) / ksplit;
  return (L.kblocks + kbps - 1) / kbps;
}

hipError_t LaunchLimbScan(const int8_t* tiles, const LimbLayout& L,
                          const int8_t* btiles, int nq, int nl, int qpt,
                          int out_bits, int ksplit, u64* partial,
                          hipStream_t s) {
  const int nt = (nq + qpt - 1) / qpt;
  switch (L.planes) {
    case 1:
      return LaunchScanNP<1>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit, partial, s);
    case 2:
      return LaunchScanNP<2>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit, partial, s);
    case 4:
      return LaunchScanNP<4>(tiles, L, btiles, nt, nq, nl, qpt, out_bits, ksplit, partial, s);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace pirg
