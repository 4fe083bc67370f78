// Integer kernels of the offline preprocessing, PreprocessingServer::Preprocess
// (crypto/pir/server/preprocessing_server.h:68-273): the DB interpolation and
// transpose, EvaluateNttInPlace over Z_(2^w)[X]/(X^d + 1), and the glue of
// PreprocessPack / BuildPreprocessMatrix (crypto/lwes_to_rlwe.cc:88-368). The
// FFT products live in fft_kernels.hip.
//
// All of these are streaming passes (HBM-bound): one thread per output word,
// grid-stride loops, reads and writes contiguous across a wave except where a
// transpose goes through LDS.
#include "preprocess_internal.h"

namespace pirg {
namespace {

unsigned GridBlocks(long long total) {
  long long b = (total + 255) / 256;
  if (b > (1LL << 20)) b = 1LL << 20;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

__device__ __forceinline__ int Brev(int x, int bits) {
  return bits == 0 ? 0 : static_cast<int>(__brev(static_cast<unsigned>(x)) >> (32 - bits));
}

// ---------------------------------------------------------------------------
// Interpolate (db_row_interpolation.cc:115-228). Thread e owns coefficient r
// of butterfly pair pi of row group g; the negacyclic rotation X^power
// (MultiplyByXPower, :43-78) becomes a shifted read of the upper polynomial.
template <typename E>
__global__ __launch_bounds__(256) void InterpStageKernel(InterpStage A) {
  const int d = A.d, t = A.t, half = A.len >> 1, two_d = 2 * A.d;
  const int log_pg = A.log_t - 1 + A.log_d;  // log2((t/2) * d)
  const long long total = A.groups << log_pg;
  const int step = (two_d - two_d / A.len) % two_d;
  const E* src = static_cast<const E*>(A.src);
  E* dst = static_cast<E*>(A.dst);
  const uint32_t p = A.p;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int r = static_cast<int>(e & (d - 1));
    const int pi = static_cast<int>((e >> A.log_d) & ((t >> 1) - 1));
    const long long g = e >> log_pg;
    const int i = (pi / half) * A.len;
    const int j = pi % half;
    const int lo = i + j, hi = lo + half;
    int power = static_cast<int>((static_cast<long long>(j) * step) % two_d);
    bool neg = false;
    if (power >= d) {
      neg = true;
      power -= d;
    }
    int sc;
    if (r < power) {
      sc = d - power + r;
      neg = !neg;
    } else {
      sc = r - power;
    }
    uint32_t u, v;
    if (A.first) {
      const long long rec_lo = g * t + Brev(lo, A.log_t);
      const long long rec_hi = g * t + Brev(hi, A.log_t);
      u = static_cast<uint32_t>(src[(rec_lo * A.ell + A.shard) * d + r]) % p;
      v = static_cast<uint32_t>(src[(rec_hi * A.ell + A.shard) * d + sc]) % p;
    } else {
      u = src[(g * t + lo) * d + r];
      v = src[(g * t + hi) * d + sc];
    }
    if (neg) v = v == 0 ? 0 : p - v;
    uint32_t nu = u + v;
    if (nu >= p) nu -= p;
    uint32_t nv = u + p - v;
    if (nv >= p) nv -= p;
    if (A.t_inv != 0) {
      nu = static_cast<uint32_t>(static_cast<uint64_t>(nu) * A.t_inv % p);
      nv = static_cast<uint32_t>(static_cast<uint64_t>(nv) * A.t_inv % p);
    }
    dst[(g * t + lo) * d + r] = static_cast<E>(nu);
    dst[(g * t + hi) * d + r] = static_cast<E>(nv);
  }
}

// All log2 t stages in registers for t <= 16. The stage-len twiddles are
// X^power with power a multiple of R0 = 2d/t, and X^power only moves a
// coefficient within its class {c0 + m R0}. Thread e owns class c0 of row
// group g: t polynomials x t/2 members, so every butterfly and rotation is
// a register operation fixed at compile time.
template <typename E, int T>
__global__ __launch_bounds__(256) void InterpFusedKernel(InterpStage A) {
  constexpr int HALF = T / 2;
  constexpr int LOG_T = T == 2 ? 1 : T == 4 ? 2 : T == 8 ? 3 : 4;
  const int d = A.d;
  const int r0 = 2 * d / T;
  const long long total = A.groups * r0;
  const E* src = static_cast<const E*>(A.src);
  E* dst = static_cast<E*>(A.dst);
  const uint32_t p = A.p;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const long long g = e / r0;
    const int c0 = static_cast<int>(e - g * r0);
    uint32_t v[T][HALF];
#pragma unroll
    for (int x = 0; x < T; ++x) {
      // Position x holds record bit_rev(x) of the group (:165-176).
      const long long rec = g * T + Brev(x, LOG_T);
      const E* row = src + (rec * A.ell + A.shard) * d + c0;
#pragma unroll
      for (int m = 0; m < HALF; ++m) v[x][m] = static_cast<uint32_t>(row[m * r0]) % p;
    }
#pragma unroll
    for (int len = 2; len <= T; len <<= 1) {
      const int half = len / 2;
      // power = j * (2d - 2d/len) mod 2d, in units of r0 (2d = T r0).
      const int stepu = (T - T / len) % T;
#pragma unroll
      for (int i = 0; i < T; i += len) {
#pragma unroll
        for (int j = 0; j < half; ++j) {
          int pu = (j * stepu) % T;
          bool neg_all = false;
          if (pu >= HALF) {
            neg_all = true;
            pu -= HALF;
          }
          uint32_t rot[HALF];
#pragma unroll
          for (int m = 0; m < HALF; ++m) {
            uint32_t val;
            bool neg = neg_all;
            if (m < pu) {
              val = v[i + j + half][m - pu + HALF];
              neg = !neg;
            } else {
              val = v[i + j + half][m - pu];
            }
            rot[m] = (neg && val != 0) ? p - val : val;
          }
#pragma unroll
          for (int m = 0; m < HALF; ++m) {
            const uint32_t u = v[i + j][m];
            uint32_t a = u + rot[m];
            if (a >= p) a -= p;
            uint32_t b = u + p - rot[m];
            if (b >= p) b -= p;
            v[i + j][m] = a;
            v[i + j + half][m] = b;
          }
        }
      }
    }
#pragma unroll
    for (int x = 0; x < T; ++x) {
      E* row = dst + (g * T + x) * d + c0;
#pragma unroll
      for (int m = 0; m < HALF; ++m) {
        row[m * r0] = static_cast<E>(static_cast<uint64_t>(v[x][m]) * A.t_inv % p);
      }
    }
  }
}

// 64 x 64 tiles through LDS; block b covers source rows (groups)
// [64 * (b / nx), +64) and source columns [64 * (b % nx), +64).
template <typename E>
__global__ __launch_bounds__(256) void TransposeKernel(const E* __restrict__ src,
                                                       long long src_stride, int rows, int cols,
                                                       E* __restrict__ dst, long long dst_cols,
                                                       long long nx) {
  __shared__ E tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const long long x0 = (blockIdx.x % nx) * 64;
  const long long g0 = (blockIdx.x / nx) * 64;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const long long g = g0 + ty + 4 * k, x = x0 + tx;
    E v = 0;
    if (g < rows && x < cols) v = src[g * src_stride + x];
    tile[ty + 4 * k][tx] = v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const long long x = x0 + ty + 4 * k, g = g0 + tx;
    if (x < cols && g < dst_cols) dst[x * dst_cols + g] = tile[tx][ty + 4 * k];
  }
}

// ---------------------------------------------------------------------------
// EvaluateNttInPlace (lwes_to_rlwe.cc:88-124), one stage.
template <typename T>
__global__ __launch_bounds__(256) void RingNttStageKernel(const T* __restrict__ in,
                                                          T* __restrict__ out, long long batches,
                                                          int log_d, int len) {
  const int d = 1 << log_d;
  const long long total = batches << (2 * log_d - 1);
  const int half = len >> 1;
  const int rot = 2 * (d / len);  // X^(2 j step), step = d / len
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int r = static_cast<int>(e & (d - 1));
    const int pi = static_cast<int>((e >> log_d) & ((d >> 1) - 1));
    const long long base = (e >> (2 * log_d - 1)) << (2 * log_d);
    const int i = (pi / half) * len;
    const int j = pi % half;
    const int P = j * rot;
    const long long plo = base + static_cast<long long>(i + j) * d;
    const long long phi = plo + static_cast<long long>(half) * d;
    const T u = in[plo + r];
    const T v = r >= P ? in[phi + r - P] : static_cast<T>(T(0) - in[phi + d - P + r]);
    out[plo + r] = static_cast<T>(u + v);
    out[phi + r] = static_cast<T>(u - v);
  }
}

// PreprocessPack (lwes_to_rlwe.cc:247-286) into the transform's input order.
__global__ __launch_bounds__(256) void PackNttInitKernel(const u64* __restrict__ hint,
                                                         int nblocks, int log_d, int shift,
                                                         u64 qmask, u64* __restrict__ A) {
  const int d = 1 << log_d;
  const long long total = static_cast<long long>(nblocks) << (2 * log_d);
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int r = static_cast<int>(e & (d - 1));
    const int k = static_cast<int>((e >> log_d) & (d - 1));
    const long long b = e >> (2 * log_d);
    const int j = (r - k) & (d - 1);
    const u64* a_j = hint + ((b << log_d) + j) * d;
    // a_tilde_j = sum_i a_j[i] X^-i: coefficient k is a_j[0] for k = 0 and
    // -a_j[d - k] otherwise; then LowBits, RightShift(log2 d), LowBits.
    const u64 coef = k == 0 ? a_j[0] : (0ULL - a_j[d - k]);
    const u64 at = shift >= 64 ? 0ULL : (((coef & qmask) >> shift) & qmask);
    // X^k * p_k: coefficient r is p_k[r - k], negated when r < k.
    A[((b << log_d) + Brev(k, log_d)) * d + r] = r < k ? (0ULL - at) : at;
  }
}

__global__ __launch_bounds__(256) void ChainCombineKernel(const u64* __restrict__ A, int log_d,
                                                          int aidx_g, int aidx_gh,
                                                          const u64* __restrict__ part,
                                                          int npairs, int nchains,
                                                          u64* __restrict__ st) {
  const int d = 1 << log_d;
  const long long total = static_cast<long long>(nchains) << log_d;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int pos = static_cast<int>(e & (d - 1));
    const int c = static_cast<int>(e >> log_d);
    const long long b = c >> 1;
    const int aidx = (c & 1) ? aidx_gh : aidx_g;
    u64 v = A[((b << log_d) + aidx) * d + pos];
    for (int p = 0; p < npairs; ++p) v += part[(static_cast<long long>(c) * npairs + p) * d + pos];
    st[e] = v;
  }
}

__global__ __launch_bounds__(256) void ChainFinalKernel(const u64* __restrict__ st,
                                                        const u64* __restrict__ part, int npairs,
                                                        int nblocks, int log_d, u64 qmask,
                                                        u64* __restrict__ a_tilde) {
  const int d = 1 << log_d;
  const long long total = static_cast<long long>(nblocks) << log_d;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int pos = static_cast<int>(e & (d - 1));
    const long long b = e >> log_d;
    u64 v = st[((2 * b) << log_d) + pos];
    for (int p = 0; p < npairs; ++p) v += part[(b * npairs + p) * d + pos];
    a_tilde[e] = v & qmask;
  }
}

// BuildPreprocessMatrix (lwes_to_rlwe.cc:195-226): tile of 64 columns c x 64
// rows r of one (block, l) through LDS.
__global__ __launch_bounds__(256) void MatrixOutKernel(const uint32_t* __restrict__ T, int k,
                                                       int d, int tiles1, int mat_bytes,
                                                       void* __restrict__ out) {
  __shared__ uint32_t tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const long long blk = blockIdx.x;
  const int tiles2 = tiles1 * tiles1;
  const long long bl = blk / tiles2;  // b * k + l
  const int tt = static_cast<int>(blk % tiles2);
  const int c0 = (tt / tiles1) * 64, r0 = (tt % tiles1) * 64;
  const long long b = bl / k;
  const int l = static_cast<int>(bl % k);
  const uint32_t* src = T + bl * d * d;
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const int c = c0 + ty + 4 * kk, r = r0 + tx;
    uint32_t v = 0;
    if (c < d && r < d) {
      const uint32_t x = src[static_cast<long long>(c) * d + ((r - c) & (d - 1))];
      v = r < c ? 0u - x : x;
    }
    tile[ty + 4 * kk][tx] = v;
  }
  __syncthreads();
  const long long kd = static_cast<long long>(k) * d;
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const int r = r0 + ty + 4 * kk, c = c0 + tx;
    if (r < d && c < d) {
      const long long o = (b * d + r) * kd + static_cast<long long>(l) * d + c;
      const uint32_t v = tile[tx][ty + 4 * kk];
      if (mat_bytes == 4) {
        static_cast<int32_t*>(out)[o] = static_cast<int32_t>(v);
      } else {
        static_cast<int16_t*>(out)[o] = static_cast<int16_t>(static_cast<uint16_t>(v));
      }
    }
  }
}

__global__ __launch_bounds__(256) void SumPartsKernel(const u64* __restrict__ part, int nparts,
                                                      int log_d, long long rows,
                                                      u64* __restrict__ out) {
  const int d = 1 << log_d;
  const long long total = rows << log_d;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int m = static_cast<int>(e & (d - 1));
    const long long x = e >> log_d;
    const u64* p = part + x * nparts * d + m;
    u64 v = 0;
    for (int q = 0; q < nparts; ++q) v += p[static_cast<long long>(q) * d];
    out[e] = v;
  }
}

int Log2(int x) {
  int l = 0;
  while ((1 << l) < x) ++l;
  return l;
}

}  // namespace

hipError_t LaunchInterpStage(const InterpStage& A, hipStream_t s) {
  if (A.t < 2 || (A.t & (A.t - 1)) != 0 || A.d < 1 || (1 << A.log_d) != A.d ||
      (1 << A.log_t) != A.t || A.p == 0) {
    return hipErrorInvalidValue;
  }
  const long long total = A.groups * (A.t / 2) * static_cast<long long>(A.d);
  if (A.elem_bytes == 2) {
    hipLaunchKernelGGL(InterpStageKernel<uint16_t>, dim3(GridBlocks(total)), dim3(256), 0, s, A);
  } else if (A.elem_bytes == 1) {
    hipLaunchKernelGGL(InterpStageKernel<uint8_t>, dim3(GridBlocks(total)), dim3(256), 0, s, A);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t LaunchInterpFused(const InterpStage& A, hipStream_t s) {
  if (A.d < 1 || (1 << A.log_d) != A.d || A.p == 0 || 2 * A.d < A.t) return hipErrorInvalidValue;
  const long long total = A.groups * (2LL * A.d / A.t);
  const dim3 grid(GridBlocks(total)), block(256);
#define PIRG_INTERP_CASE(TT)                                                           \
  case TT:                                                                             \
    if (A.elem_bytes == 2) {                                                           \
      hipLaunchKernelGGL((InterpFusedKernel<uint16_t, TT>), grid, block, 0, s, A);     \
    } else if (A.elem_bytes == 1) {                                                    \
      hipLaunchKernelGGL((InterpFusedKernel<uint8_t, TT>), grid, block, 0, s, A);      \
    } else {                                                                           \
      return hipErrorInvalidValue;                                                     \
    }                                                                                  \
    break;
  switch (A.t) {
    PIRG_INTERP_CASE(2)
    PIRG_INTERP_CASE(4)
    PIRG_INTERP_CASE(8)
    PIRG_INTERP_CASE(16)
    default: return hipErrorInvalidValue;
  }
#undef PIRG_INTERP_CASE
  return hipGetLastError();
}

hipError_t LaunchTranspose(const void* src, long long src_stride, int rows, int cols,
                           int elem_bytes, void* dst, long long dst_cols, hipStream_t s) {
  const long long nx = (cols + 63) / 64;
  const long long ny = (dst_cols + 63) / 64;
  const long long blocks = nx * ny;
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  if (elem_bytes == 2) {
    hipLaunchKernelGGL(TransposeKernel<uint16_t>, dim3(static_cast<unsigned>(blocks)), dim3(256),
                       0, s, static_cast<const uint16_t*>(src), src_stride, rows, cols,
                       static_cast<uint16_t*>(dst), dst_cols, nx);
  } else if (elem_bytes == 1) {
    hipLaunchKernelGGL(TransposeKernel<uint8_t>, dim3(static_cast<unsigned>(blocks)), dim3(256),
                       0, s, static_cast<const uint8_t*>(src), src_stride, rows, cols,
                       static_cast<uint8_t*>(dst), dst_cols, nx);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t LaunchRingNttStage(const void* in, void* out, int word_bytes, long long batches,
                              int log_d, int len, hipStream_t s) {
  if (log_d < 1 || len < 2 || len > (1 << log_d)) return hipErrorInvalidValue;
  const long long total = batches << (2 * log_d - 1);
  if (word_bytes == 8) {
    hipLaunchKernelGGL(RingNttStageKernel<u64>, dim3(GridBlocks(total)), dim3(256), 0, s,
                       static_cast<const u64*>(in), static_cast<u64*>(out), batches, log_d, len);
  } else if (word_bytes == 4) {
    hipLaunchKernelGGL(RingNttStageKernel<uint32_t>, dim3(GridBlocks(total)), dim3(256), 0, s,
                       static_cast<const uint32_t*>(in), static_cast<uint32_t*>(out), batches,
                       log_d, len);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t LaunchPackNttInit(const u64* hint, int nblocks, int d, int log_d, int shift,
                             u64 qmask, u64* A, hipStream_t s) {
  if ((1 << log_d) != d) return hipErrorInvalidValue;
  const long long total = static_cast<long long>(nblocks) << (2 * log_d);
  hipLaunchKernelGGL(PackNttInitKernel, dim3(GridBlocks(total)), dim3(256), 0, s, hint, nblocks,
                     log_d, shift, qmask, A);
  return hipGetLastError();
}

hipError_t LaunchChainCombine(const u64* A, int d, int aidx_g, int aidx_gh, const u64* part,
                              int npairs, int nchains, u64* st, hipStream_t s) {
  const int log_d = Log2(d);
  hipLaunchKernelGGL(ChainCombineKernel, dim3(GridBlocks(static_cast<long long>(nchains) * d)),
                     dim3(256), 0, s, A, log_d, aidx_g, aidx_gh, part, npairs, nchains, st);
  return hipGetLastError();
}

hipError_t LaunchChainFinal(const u64* st, const u64* part, int npairs, int nblocks, int d,
                            u64 qmask, u64* a_tilde, hipStream_t s) {
  const int log_d = Log2(d);
  hipLaunchKernelGGL(ChainFinalKernel, dim3(GridBlocks(static_cast<long long>(nblocks) * d)),
                     dim3(256), 0, s, st, part, npairs, nblocks, log_d, qmask, a_tilde);
  return hipGetLastError();
}

hipError_t LaunchMatrixOut(const uint32_t* T, int nblocks, int k, int d, int mat_bytes,
                           void* out, hipStream_t s) {
  const int tiles1 = (d + 63) / 64;
  const long long blocks = static_cast<long long>(nblocks) * k * tiles1 * tiles1;
  if (blocks <= 0) return hipSuccess;
  if (blocks > 0x7fffffffLL || (mat_bytes != 2 && mat_bytes != 4)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(MatrixOutKernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, T, k,
                     d, tiles1, mat_bytes, out);
  return hipGetLastError();
}

hipError_t LaunchSumParts(const u64* part, int nparts, int d, long long rows, u64* out,
                          hipStream_t s) {
  const int log_d = Log2(d);
  hipLaunchKernelGGL(SumPartsKernel, dim3(GridBlocks(rows * d)), dim3(256), 0, s, part, nparts,
                     log_d, rows, out);
  return hipGetLastError();
}

}  // namespace pirg
