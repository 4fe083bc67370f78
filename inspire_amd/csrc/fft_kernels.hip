// FFT polynomial products, FinalizeMatrixPack, ExternalProduct / EvalPoly and
// the modulus switch of PirServer::ProcessResponse.
//
// Polynomial::InnerProductFft (crypto/polynomial_fft.cc:68-197) rounds
// double-precision FFT results back to integers. With the reference's chunk
// sizes those doubles can sit within 0.5 of a rounding boundary (at d = 4096
// the ExternalProduct chunks reach 2^52.6 in magnitude), so its integer output
// depends on the exact sequence of IEEE operations. These kernels therefore
// perform, butterfly for butterfly, the same operations as crypto/fft.h:106-151
// -- radix-2 decimation in time on bit-reversed input, twiddles computed on the
// host by the same cos/sin expression, complex products as (ac - bd, ad + bc)
// with no fused multiply-add (this file is compiled with -ffp-contract=off) --
// and accumulate the chunk-pair spectra in the reference's order. Grouping two
// radix-2 stages per LDS pass changes the schedule, not the arithmetic.
//
// A workgroup owns one 2d-point transform held in LDS (16 * 2d bytes: 128 KiB
// at d = 4096, within the 160 KiB of a CDNA4 CU).
#include "pirg_internal.h"

namespace pirg {
namespace {

constexpr int kFftThreads = 512;

// crypto/fft.h:122-125 for one butterfly: v = y * w, x' = x + v, y' = x - v.
__device__ __forceinline__ void Butterfly(double2& x, double2& y, const double2 w) {
  const double vr = y.x * w.x - y.y * w.y;
  const double vi = y.x * w.y + y.y * w.x;
  const double ux = x.x, uy = x.y;
  x = make_double2(ux + vr, uy + vi);
  y = make_double2(ux - vr, uy - vi);
}

// Twiddles of the radix-4 groups a thread owns in the pass starting at stage
// s: w1 = tw[j * step1] (stage s), w2 = tw[j * step2], w3 = tw[(j + h) * step2]
// (stage s + 1).
constexpr int kMaxGroups = 4;  // n / 4 <= kFftThreads * kMaxGroups (n <= 8192)

__device__ __forceinline__ void LoadPassTwiddles(const double2* __restrict__ tw,
                                                 int log_n, int s, double2* w1,
                                                 double2* w2, double2* w3) {
  const int n = 1 << log_n;
  const int h = 1 << (s - 1);
  const int step1 = n >> s;
  const int step2 = n >> (s + 1);
#pragma unroll
  for (int m = 0; m < kMaxGroups; ++m) {
    const int g = threadIdx.x + m * blockDim.x;
    if (g < n / 4) {
      const int j = g & (h - 1);
      w1[m] = tw[j * step1];
      w2[m] = tw[j * step2];
      w3[m] = tw[(j + h) * step2];
    }
  }
}

// crypto/fft.h:117-128 on a bit-reversed LDS buffer, two stages per pass; the
// next pass's twiddles are loaded while the current pass runs.
__device__ void FftLds(double2* buf, const double2* __restrict__ tw, int log_n) {
  const int n = 1 << log_n;
  int s = 1;
  double2 w1[kMaxGroups], w2[kMaxGroups], w3[kMaxGroups];
  double2 n1[kMaxGroups], n2[kMaxGroups], n3[kMaxGroups];
  if (log_n >= 2) LoadPassTwiddles(tw, log_n, 1, w1, w2, w3);
  for (; s + 1 <= log_n; s += 2) {
    const int h = 1 << (s - 1);
    if (s + 3 <= log_n) LoadPassTwiddles(tw, log_n, s + 2, n1, n2, n3);
#pragma unroll
    for (int m = 0; m < kMaxGroups; ++m) {
      const int g = threadIdx.x + m * blockDim.x;
      if (g < n / 4) {
        const int j = g & (h - 1);
        const int e0 = ((g >> (s - 1)) << (s + 1)) + j;
        double2 x0 = buf[e0];
        double2 x1 = buf[e0 + h];
        double2 x2 = buf[e0 + 2 * h];
        double2 x3 = buf[e0 + 3 * h];
        Butterfly(x0, x1, w1[m]);
        Butterfly(x2, x3, w1[m]);
        Butterfly(x0, x2, w2[m]);
        Butterfly(x1, x3, w3[m]);
        buf[e0] = x0;
        buf[e0 + h] = x1;
        buf[e0 + 2 * h] = x2;
        buf[e0 + 3 * h] = x3;
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kMaxGroups; ++m) {
      w1[m] = n1[m];
      w2[m] = n2[m];
      w3[m] = n3[m];
    }
  }
  if (s == log_n) {
    const int h = 1 << (s - 1);
    const int step = n >> s;
    for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
      const int j = b & (h - 1);
      const int e0 = ((b >> (s - 1)) << s) + j;
      double2 x0 = buf[e0];
      double2 x1 = buf[e0 + h];
      Butterfly(x0, x1, tw[j * step]);
      buf[e0] = x0;
      buf[e0 + h] = x1;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int BitRev(int j, int log_n) {
  return static_cast<int>(__brev(static_cast<unsigned>(j)) >> (32 - log_n));
}

// Forward spectra (crypto/polynomial_fft.cc:119-152): one workgroup per
// (polynomial, chunk).
__global__ __launch_bounds__(kFftThreads) void SpectraKernel(
    const u64* __restrict__ src, int per_group, long long group_stride,
    ChunkRule rule, int d, int log_n, const double2* __restrict__ tw,
    double2* __restrict__ dst) {
  extern __shared__ double2 buf[];
  const int n = 1 << log_n;
  const int job = blockIdx.x;
  const int p = job / rule.num_chunks;
  const int c = job % rule.num_chunks;
  const u64* poly = src + (long long)(p / per_group) * group_stride +
                    (long long)(p % per_group) * d;
  constexpr int kElems = 8;
  for (int j0 = threadIdx.x; j0 < n; j0 += blockDim.x * kElems) {
    u64 x[kElems];
#pragma unroll
    for (int e = 0; e < kElems; ++e) {
      const int j = j0 + e * blockDim.x;
      x[e] = (j < d) ? poly[j] : 0ULL;
    }
#pragma unroll
    for (int e = 0; e < kElems; ++e) {
      const int j = j0 + e * blockDim.x;
      if (j < n) {
        const double v = (j < d) ? ChunkValue(rule, x[e], c) : 0.0;
        buf[BitRev(j, log_n)] = make_double2(v, 0.0);
      }
    }
  }
  __syncthreads();
  FftLds(buf, tw, log_n);
  double2* out = dst + (long long)job * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) out[j] = buf[j];
}

struct IpJob {
  const double2* const* uptr;  // [n_out] base of the u spectra (terms x nu)
  const double2* const* vptr;  // [n_out] base of the v spectra (terms x nv)
  u64* const* optr;            // [n_out] base of the pair partials (npairs x d)
};

// Accumulate-invert-round (crypto/polynomial_fft.cc:154-196) for one
// (output, chunk pair): acc[j] = sum_i U_i,a[j] * V_i,b[j] in term order, the
// inverse FFT, real / 2d, round half away from zero, shift by chunk_bits*(a+b)
// and fold X^d = -1. One workgroup per (output, pair).
__global__ __launch_bounds__(kFftThreads) void IpKernel(
    IpJob jobs, int nterms, int nu, int nv, PairList pairs, int d, int log_n,
    const double2* __restrict__ tw, unsigned long long* margin) {
  extern __shared__ double2 buf[];
  const int n = 1 << log_n;
  const int o = blockIdx.x / pairs.count;
  const int pi = blockIdx.x % pairs.count;
  const int a = pairs.a[pi], b = pairs.b[pi], sh = pairs.shift[pi];
  const double2* U = jobs.uptr[o];
  const double2* V = jobs.vptr[o];
  // kElems independent elements per thread keep that many load chains in
  // flight; the term order of each element's sum is the reference's.
  constexpr int kElems = 8;
  for (int j0 = threadIdx.x; j0 < n; j0 += blockDim.x * kElems) {
    double ax[kElems], ay[kElems];
#pragma unroll
    for (int e = 0; e < kElems; ++e) {
      ax[e] = 0.0;
      ay[e] = 0.0;
    }
    for (int i = 0; i < nterms; ++i) {
      const double2* Ui = U + ((long long)i * nu + a) * n;
      const double2* Vi = V + ((long long)i * nv + b) * n;
      double2 uu[kElems], vv[kElems];
#pragma unroll
      for (int e = 0; e < kElems; ++e) {
        const int j = j0 + e * blockDim.x;
        uu[e] = (j < n) ? Ui[j] : make_double2(0.0, 0.0);
        vv[e] = (j < n) ? Vi[j] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int e = 0; e < kElems; ++e) {
        const double pr = uu[e].x * vv[e].x - uu[e].y * vv[e].y;
        const double pim = uu[e].x * vv[e].y + uu[e].y * vv[e].x;
        ax[e] = ax[e] + pr;
        ay[e] = ay[e] + pim;
      }
    }
#pragma unroll
    for (int e = 0; e < kElems; ++e) {
      const int j = j0 + e * blockDim.x;
      if (j < n) buf[BitRev(j, log_n)] = make_double2(ax[e], ay[e]);
    }
  }
  __syncthreads();
  FftLds(buf, tw, log_n);
  // x / 2d of polynomial_fft.cc:184: 2d is a power of two, so the product by
  // its exact reciprocal is the same correctly rounded value.
  const double inv_n = ldexp(1.0, -log_n);
  u64* out = jobs.optr[o] + (long long)pi * d;
  double local_max = 0.0;
  for (int m = threadIdx.x; m < d; m += blockDim.x) {
    const double x0 = buf[m].x * inv_n;
    const double x1 = buf[m + d].x * inv_n;
    const double r0 = round(x0);
    const double r1 = round(x1);
    local_max = fmax(local_max, fmax(fabs(x0 - r0), fabs(x1 - r1)));
    const u64 v0 = static_cast<u64>(static_cast<i64>(r0)) << sh;
    const u64 v1 = static_cast<u64>(static_cast<i64>(r1)) << sh;
    out[m] = v0 - v1;
  }
  if (margin != nullptr) {
    atomicMax(margin, static_cast<unsigned long long>(__double_as_longlong(local_max)));
  }
}

__device__ __forceinline__ u64 RescaleCoeff(u64 c, int log_cur, int log_new) {
  // crypto/polynomial.cc:287-304 in 128-bit.
  const unsigned __int128 v = (static_cast<unsigned __int128>(c) << log_new) +
                              (static_cast<unsigned __int128>(1) << (log_cur - 1));
  return static_cast<u64>(v >> log_cur);
}

__device__ __forceinline__ u64 LowBitsD(u64 x, int amount) {
  return amount >= 64 ? x : (x & ((1ULL << amount) - 1));
}

}  // namespace

// ---------------------------------------------------------------------------
// Stage kernels of ProcessResponse, batched over nq requests.


// FinalizeMatrixPack's integer part (crypto/lwes_to_rlwe.cc:462-476):
// b_agg = LowBits(LowBits(b + b_part) + inner_h); then the Horner start
// (pir_server.cc:137) or, when t == 1, the modulus switch (:288-297).
__global__ void FinalizeCombineKernel(FinalizeArgs A) {
  const long long total = (long long)A.nq * A.M;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int q = static_cast<int>(g / A.M);
    const int row = static_cast<int>(g % A.M);  // (s*t + i)*d + m
    u64 bprime = A.corr_db[q];
    for (int k = 0; k < A.ksplit_db; ++k) {
      bprime += A.db_partial[((long long)k * A.nq + q) * A.M + row];
    }
    u64 bpart = A.corr_pk[q];
    for (int k = 0; k < A.ksplit_pk; ++k) {
      bpart += A.pk_partial[((long long)k * A.nq + q) * A.M + row];
    }
    const int oi = row / A.d, m = row % A.d;
    const u64* fp = A.fin_partial + (((long long)q * (A.M / A.d) + oi) * A.npairs) * A.d;
    u64 inner = 0;
    for (int pi = 0; pi < A.npairs; ++pi) inner += fp[(long long)pi * A.d + m];
    const u64 bagg = LowBitsD(LowBitsD(bprime + bpart, A.log_q) + inner, A.log_q);
    A.ct_b[(long long)q * A.M + row] = bagg;
    const int s = oi / A.t, i = oi % A.t;
    if (i == A.t - 1) {
      const u64 a = A.a_tilde[(long long)oi * A.d + m];
      if (A.t == 1) {
        u64* o = A.out + (((long long)q * A.ell + s) * 2) * A.d;
        o[m] = RescaleCoeff(a, A.log_q, A.log_q1);
        o[A.d + m] = RescaleCoeff(bagg, A.log_q, A.log_q2);
      } else {
        A.st_a[((long long)q * A.ell + s) * A.d + m] = a;
        A.st_b[((long long)q * A.ell + s) * A.d + m] = bagg;
      }
    }
  }
}


// ExternalProduct's decomposition (crypto/external_product.cc:60-72,
// polynomial.cc:340-374) fused with the forward FFT of each digit chunk.
__global__ __launch_bounds__(kFftThreads) void DigitsSpectraKernel(
    DigitsArgs A, int log_n, const double2* __restrict__ tw) {
  extern __shared__ double2 buf[];
  const int n = 1 << log_n;
  const int nu = A.rule.num_chunks;
  const int job = blockIdx.x;                 // ((q*ell + s)*2e + jj)*nu + a
  const int a = job % nu;
  const int jj = (job / nu) % (2 * A.e);
  const long long qs = job / (nu * 2 * A.e);  // q*ell + s
  const u64* st = (jj < A.e ? A.st_a : A.st_b) + qs * A.d;
  const int digit = jj % A.e;
  const int to_shift = A.log_q - A.e * A.log_digit;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double v = 0.0;
    if (j < A.d) {
      u64 x = LowBitsD(st[j], A.log_q);
      if (to_shift > 0) {
        x = static_cast<u64>((static_cast<unsigned __int128>(x) +
                              (1ULL << (to_shift - 1))) >> to_shift);
      }
      u64 cur = x;
      for (int t = 0; t < digit; ++t) cur = (A.log_digit >= 64) ? 0 : (cur >> A.log_digit);
      const u64 dg = LowBitsD(cur, A.log_digit);
      v = ChunkValue(A.rule, dg, a);
    }
    buf[BitRev(j, log_n)] = make_double2(v, 0.0);
  }
  __syncthreads();
  FftLds(buf, tw, log_n);
  double2* out = A.dspec + (long long)job * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) out[j] = buf[j];
}


// EvalPoly's Horner update (pir_server.cc:141-146): result = product + c_i,
// not reduced; after the last step the modulus switch (pir_server.cc:288-297).
__global__ void ExtCombineKernel(ExtCombineArgs A) {
  const long long total = (long long)A.nq * A.ell * 2 * A.d;
  for (long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x; g < total;
       g += (long long)gridDim.x * blockDim.x) {
    const int m = static_cast<int>(g % A.d);
    const int which = static_cast<int>((g / A.d) % 2);
    const long long qs = g / (2LL * A.d);
    const int s = static_cast<int>(qs % A.ell);
    const long long q = qs / A.ell;
    const u64* pp = A.ext_partial + ((qs * 2 + which) * A.npairs) * (long long)A.d;
    u64 v = 0;
    for (int pi = 0; pi < A.npairs; ++pi) v += pp[(long long)pi * A.d + m];
    const long long oi = (long long)s * A.t + A.step_i;
    const u64 c = (which == 0) ? A.a_tilde[oi * A.d + m]
                               : A.ct_b[(q * A.ell * A.t + oi) * A.d + m];
    v += c;
    if (A.step_i == 0) {
      A.out[(qs * 2 + which) * A.d + m] =
          RescaleCoeff(v, A.log_q, which == 0 ? A.log_q1 : A.log_q2);
    } else {
      (which == 0 ? A.st_a : A.st_b)[qs * A.d + m] = v;
    }
  }
}

// Sum of pair partials for the standalone InnerProductFft entry point.
__global__ void SumPairsKernel(const u64* __restrict__ partial, int npairs,
                               int d, u64* __restrict__ out) {
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < d;
       m += gridDim.x * blockDim.x) {
    u64 v = 0;
    for (int pi = 0; pi < npairs; ++pi) v += partial[(long long)pi * d + m];
    out[m] = v;
  }
}

// ---------------------------------------------------------------------------
// Launchers.

static int GridFor(long long total) {
  long long b = (total + 255) / 256;
  if (b > 65535LL * 8) b = 65535LL * 8;
  return static_cast<int>(b < 1 ? 1 : b);
}

hipError_t ConfigureFftKernels(int n) {
  const int bytes = n * static_cast<int>(sizeof(double2));
  hipError_t e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(SpectraKernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(IpKernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(DigitsSpectraKernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  return e;
}

hipError_t LaunchSpectra(const u64* src, int npolys, int per_group,
                         long long group_stride, const ChunkRule& rule,
                         const FftDims& F, const double2* tw_fwd,
                         double2* dst, hipStream_t s) {
  if (npolys <= 0) return hipSuccess;
  hipLaunchKernelGGL(SpectraKernel, dim3(npolys * rule.num_chunks),
                     dim3(kFftThreads), F.N * sizeof(double2), s, src, per_group,
                     group_stride, rule, F.d, F.log_N, tw_fwd, dst);
  return hipGetLastError();
}

hipError_t LaunchIp(const double2* const* uptr, const double2* const* vptr,
                    u64* const* optr, int n_out, int nterms, int nu, int nv,
                    const PairList& pairs, const FftDims& F,
                    const double2* tw_bwd, unsigned long long* margin,
                    hipStream_t s) {
  if (n_out <= 0) return hipSuccess;
  IpJob jobs{uptr, vptr, optr};
  hipLaunchKernelGGL(IpKernel, dim3(n_out * pairs.count), dim3(kFftThreads),
                     F.N * sizeof(double2), s, jobs, nterms, nu, nv, pairs, F.d,
                     F.log_N, tw_bwd, margin);
  return hipGetLastError();
}

hipError_t LaunchFinalizeCombine(const FinalizeArgs& A, hipStream_t s) {
  hipLaunchKernelGGL(FinalizeCombineKernel, dim3(GridFor((long long)A.nq * A.M)),
                     dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t LaunchDigitsSpectra(const DigitsArgs& A, const FftDims& F,
                               const double2* tw_fwd, hipStream_t s) {
  const int jobs = A.nq * A.ell * 2 * A.e * A.rule.num_chunks;
  hipLaunchKernelGGL(DigitsSpectraKernel, dim3(jobs), dim3(kFftThreads),
                     F.N * sizeof(double2), s, A, F.log_N, tw_fwd);
  return hipGetLastError();
}

hipError_t LaunchExtCombine(const ExtCombineArgs& A, hipStream_t s) {
  hipLaunchKernelGGL(ExtCombineKernel,
                     dim3(GridFor((long long)A.nq * A.ell * 2 * A.d)), dim3(256),
                     0, s, A);
  return hipGetLastError();
}

hipError_t LaunchSumPairs(const u64* partial, int npairs, int d, u64* out,
                          hipStream_t s) {
  hipLaunchKernelGGL(SumPairsKernel, dim3(GridFor(d)), dim3(256), 0, s, partial,
                     npairs, d, out);
  return hipGetLastError();
}

}  // namespace pirg
