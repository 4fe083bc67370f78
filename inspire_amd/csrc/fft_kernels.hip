// FFT polynomial products, FinalizeMatrixPack, ExternalProduct / EvalPoly and
// the modulus switch of PirServer::ProcessResponse.
//
// Polynomial::InnerProductFft (crypto/polynomial_fft.cc:68-197) rounds
// double-precision FFT results back to integers. With the reference's chunk
// sizes those doubles can sit within 0.5 of a rounding boundary (at d = 4096
// the ExternalProduct chunks reach 2^52.6 in magnitude), so its integer output
// depends on the exact sequence of IEEE operations. These kernels therefore
// perform, butterfly for butterfly, the same operations as crypto/fft.h:106-151
// (the register schedule of fft_regs.h), with twiddles computed on the host by
// the same cos/sin expression, complex products as (ac - bd, ad + bc) with no
// fused multiply-add (this file is compiled with -ffp-contract=off), and
// accumulate the chunk-pair spectra in the reference's term order.
//
// A workgroup of N/16 threads owns one N-point transform (N = 2d); the kernels
// are instantiated per log2 N and dispatched at launch.
#include <algorithm>
#include "fft_regs.h"
#include "preprocess_internal.h"

namespace pirg {
namespace {

using fftr::Geo;
using fftr::InSlot;
using fftr::OutSlot;
using fftr::Transform;

// Forward spectra (crypto/polynomial_fft.cc:119-152): one workgroup per
// (polynomial, chunk). Input j = t + NT*k is coefficient j for j < d, zero
// padding above.
template <int L, bool H>
__global__ __launch_bounds__(Geo<L>::NT) __attribute__((amdgpu_waves_per_eu(H ? 4 : 1))) void SpectraKernel(
    const u64* __restrict__ src, int per_group, long long group_stride,
    ChunkRule rule, const double2* __restrict__ st, double2* __restrict__ dst) {
  extern __shared__ __align__(16) double lds[];
  constexpr int N = Geo<L>::N, E = Geo<L>::E, NT = Geo<L>::NT, d = N / 2;
  const int t = threadIdx.x;
  const int job = blockIdx.x;
  const int p = job / rule.num_chunks;
  const int c = job % rule.num_chunks;
  const u64* poly = src + (long long)(p / per_group) * group_stride +
                    (long long)(p % per_group) * d + t;
  u64 x[E / 2];
#pragma unroll
  for (int k = 0; k < E / 2; ++k) x[k] = poly[NT * k];
  double2 v[16];
#pragma unroll
  for (int k = 0; k < E / 2; ++k) v[InSlot<L>(k)] = make_double2(ChunkValue(rule, x[k], c), 0.0);
#pragma unroll
  for (int k = E / 2; k < E; ++k) v[InSlot<L>(k)] = make_double2(0.0, 0.0);
  Transform<L, H>(v, lds, st);
  double2* out = dst + (long long)job * N + t;
#pragma unroll
  for (int e = 0; e < E; ++e) out[NT * e] = v[OutSlot<L>(e)];
}

struct IpJob {
  const double2* const* uptr;  // [n_out] base of the u spectra (terms x nu)
  const double2* const* vptr;  // [n_out] base of the v spectra (terms x nv)
  u64* const* optr;            // [n_out] base of the pair partials (npairs x d)
};

// Job placement of IpKernel: outs outputs per request; s > 0 places tiles of
// r requests x s outputs on one XCD.
struct IpTile {
  int outs, r, s;
};

// Raises *margin to bits unless it already holds as much. The margin settles
// within the first workgroups of a launch, so a plain load turns nearly every
// later atomic into none: tens of thousands of same-address atomics per
// 256-query batch took 0.6 ms of its 15.5.
__device__ __forceinline__ void RaiseMargin(unsigned long long* margin, unsigned long long bits) {
  const unsigned long long cur =
      __hip_atomic_load(margin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (bits > cur) atomicMax(margin, bits);
}

// Maximum of non-negative doubles (their bit patterns order the same way as
// the values): a wave-level reduction and one update per wave when the
// workgroup is made of full waves, one per thread otherwise.
template <int NT>
__device__ __forceinline__ void MarginMax(double local, unsigned long long* margin) {
  unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(local));
  if constexpr (NT % 64 == 0) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(bits, off);
      bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0) RaiseMargin(margin, bits);
  } else {
    RaiseMargin(margin, bits);
  }
}

// Accumulate-invert-round (crypto/polynomial_fft.cc:154-196) for one
// (output, chunk pair): acc[j] = sum_i U_i,a[j] * V_i,b[j] in term order, the
// inverse FFT, real / 2d, round half away from zero, shift by chunk_bits*(a+b)
// and fold X^d = -1. One workgroup per (output, pair).
//
// Job order: the G = outs_per_group * npairs jobs of one group (outputs that
// read the same u spectra) take consecutive slots on one XCD (workgroup w runs
// on XCD w % 8), so that group's spectra are fetched into one L2 rather than
// into all eight; the grid is padded to whole rounds of 8 groups.
//
// With many outputs per request (tile.s > 0) the jobs are placed in 2-D
// tiles instead: tile.r requests x tile.s of each request's tile.outs outputs
// (outputs q * tile.outs + oi), all pairs, on one XCD. A tile's u spectra
// (tile.s outputs) and v spectra (tile.r requests) then fit that XCD's L2
// together, where one request's group would stream every output's u spectra
// through it (and a single request's jobs would all run on one XCD).
//
// FA: the accumulated spectra come ready from MacKernel (acc, one N-point
// array per job) and the workgroup only inverts and rounds.
template <int L, bool H, bool FA>
__global__ __launch_bounds__(Geo<L>::NT) __attribute__((amdgpu_waves_per_eu(H ? 4 : 1))) void IpKernel(
    IpJob jobs, int n_jobs, int group_jobs, IpTile tile, int nterms, int nu, int nv,
    PairList pairs, const double2* __restrict__ st, unsigned long long* margin,
    const double2* __restrict__ acc) {
  extern __shared__ __align__(16) double lds[];
  constexpr int N = Geo<L>::N, E = Geo<L>::E, NT = Geo<L>::NT, d = N / 2;
  const int t = threadIdx.x;
  const int w = blockIdx.x;
  int job;
  if (tile.s <= 0) {
    job = ((w >> 3) / group_jobs * 8 + (w & 7)) * group_jobs + (w >> 3) % group_jobs;
  } else {
    const int slot = w >> 3;
    const int tix = slot / group_jobs * 8 + (w & 7);
    const int jt = slot % group_jobs;
    const int tiles_s = (tile.outs + tile.s - 1) / tile.s;
    const int rest = jt / pairs.count;
    const int q = tix / tiles_s * tile.r + rest / tile.s;
    const int oi = tix % tiles_s * tile.s + rest % tile.s;
    if (oi >= tile.outs) return;
    job = (q * tile.outs + oi) * pairs.count + jt % pairs.count;
  }
  if (job >= n_jobs) return;
  const int o = job / pairs.count;
  const int pi = job % pairs.count;
  const int a = pairs.a[pi], b = pairs.b[pi], sh = pairs.shift[pi];
  const double2* U = jobs.uptr[o] + t;
  const double2* V = jobs.vptr[o] + t;
  double2 v[16];
  if constexpr (FA) {
    const double2* A = acc + (long long)job * N + t;
#pragma unroll
    for (int k = 0; k < E; ++k) v[InSlot<L>(k)] = A[NT * k];
  } else {
#pragma unroll
  for (int k = 0; k < E; ++k) v[k] = make_double2(0.0, 0.0);
  for (int i = 0; i < nterms; ++i) {
    const double2* Ui = U + ((long long)i * nu + a) * N;
    const double2* Vi = V + ((long long)i * nv + b) * N;
    // H: positions in runs of 4, since two workgroups share a CU and each
    // keeps its loads in flight in a quarter of the registers; otherwise all
    // E positions' loads at once.
    constexpr int kRun = H ? 4 : E;
#pragma unroll
    for (int k0 = 0; k0 < E; k0 += kRun) {
      double2 uu[kRun], vv[kRun];
#pragma unroll
      for (int k = 0; k < kRun && k0 + k < E; ++k) {
        uu[k] = Ui[NT * (k0 + k)];
        vv[k] = Vi[NT * (k0 + k)];
      }
#pragma unroll
      for (int k = 0; k < kRun && k0 + k < E; ++k) {
        const double pr = uu[k].x * vv[k].x - uu[k].y * vv[k].y;
        const double pim = uu[k].x * vv[k].y + uu[k].y * vv[k].x;
        double2& ac = v[InSlot<L>(k0 + k)];
        ac.x = ac.x + pr;
        ac.y = ac.y + pim;
      }
    }
  }
  }
  Transform<L, H>(v, lds, st);
  // x / 2d of polynomial_fft.cc:184: 2d is a power of two, so the product by
  // its exact reciprocal is the same correctly rounded value.
  const double inv_n = 1.0 / static_cast<double>(N);
  u64* out = jobs.optr[o] + (long long)pi * d + t;
  double local_max = 0.0;
#pragma unroll
  for (int e = 0; e < E / 2; ++e) {
    const double x0 = v[OutSlot<L>(e)].x * inv_n;           // position t + NT*e
    const double x1 = v[OutSlot<L>(e + E / 2)].x * inv_n;   // position + d
    const double r0 = round(x0);
    const double r1 = round(x1);
    local_max = fmax(local_max, fmax(fabs(x0 - r0), fabs(x1 - r1)));
    const u64 v0 = static_cast<u64>(static_cast<i64>(r0)) << sh;
    const u64 v1 = static_cast<u64>(static_cast<i64>(r1)) << sh;
    out[NT * e] = v0 - v1;
  }
  if (margin != nullptr) MarginMax<NT>(local_max, margin);
}

// The accumulation of IpKernel spread over the GPU (one thread per job and
// position, same operations in the same term order), for launches with few
// jobs, where one workgroup per job would be bound by one CU's bandwidth.
__global__ __launch_bounds__(256) void MacKernel(IpJob jobs, int n_jobs, int nterms, int nu,
                                                 int nv, PairList pairs, int log_n,
                                                 double2* __restrict__ acc) {
  const int N = 1 << log_n;
  const long long total = static_cast<long long>(n_jobs) << log_n;
  for (long long e = blockIdx.x * 256LL + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int job = static_cast<int>(e >> log_n);
    const int pos = static_cast<int>(e & (N - 1));
    const int o = job / pairs.count;
    const int pi = job % pairs.count;
    const int a = pairs.a[pi], b = pairs.b[pi];
    const double2* U = jobs.uptr[o] + pos;
    const double2* V = jobs.vptr[o] + pos;
    double2 ac = make_double2(0.0, 0.0);
    for (int i = 0; i < nterms; ++i) {
      const double2 u = U[((long long)i * nu + a) * N];
      const double2 v = V[((long long)i * nv + b) * N];
      const double pr = u.x * v.x - u.y * v.y;
      const double pim = u.x * v.y + u.y * v.x;
      ac.x = ac.x + pr;
      ac.y = ac.y + pim;
    }
    acc[e] = ac;
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

// Partial sums of the combine kernels: p[0] + p[stride] + ... +
// p[(n - 1) * stride] mod 2^64 over short runs (K-splits, chunk pairs). A loop
// with a runtime bound waits for each load before it issues the next, so a
// run is loaded in two steps: Issue<W> starts the loads of its first W terms
// (callers issue several runs before adding any), Finish adds them and any
// terms past W.
template <int W>
struct StridedRun {
  u64 v[W];
  __device__ __forceinline__ void Issue(const u64* __restrict__ p, long long stride, int n) {
#pragma unroll
    for (int j = 0; j < W; ++j) v[j] = j < n ? p[(long long)j * stride] : 0ULL;
  }
  __device__ __forceinline__ u64 Finish(const u64* __restrict__ p, long long stride,
                                        int n) const {
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) s += v[j];
    for (int k = W; k < n; ++k) s += p[(long long)k * stride];
    return s;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Stage kernels of ProcessResponse, batched over nq requests.


// FinalizeMatrixPack's integer part (crypto/lwes_to_rlwe.cc:462-476):
// b_agg = LowBits(LowBits(b + b_part) + inner_h); then the Horner start
// (pir_server.cc:137) or, when t == 1, the modulus switch (:288-297).
// Requests along blockIdx.y (strided by gridDim.y); d is a power of two.
__global__ void FinalizeCombineKernel(FinalizeArgs A) {
  const int log_d = __builtin_ctz(static_cast<unsigned>(A.d));
  const long long qm = (long long)A.nq * A.M;
  for (int q = blockIdx.y; q < A.nq; q += gridDim.y)
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < A.M;  // (s*t + i)*d + m
       row += gridDim.x * blockDim.x) {
    const u64* dp = A.db_partial + (long long)q * A.M + row;
    const u64* kp = A.pk_partial + (long long)q * A.M + row;
    const int oi = row >> log_d, m = row & (A.d - 1);
    const u64* fp = A.fin_partial + (((long long)q * (A.M / A.d) + oi) * A.npairs) * A.d + m;
    StridedRun<8> rd;
    StridedRun<4> rk;
    StridedRun<8> rf;
    rd.Issue(dp, qm, A.ksplit_db);
    rk.Issue(kp, qm, A.ksplit_pk);
    rf.Issue(fp, A.d, A.npairs);
    const u64 bprime = A.corr_db[q] + rd.Finish(dp, qm, A.ksplit_db);
    const u64 bpart = A.corr_pk[A.corr_rows > 0 ? row / A.corr_rows : q] +
                      rk.Finish(kp, qm, A.ksplit_pk);
    const u64 inner = rf.Finish(fp, A.d, A.npairs);
    const u64 bagg = LowBitsD(LowBitsD(bprime + bpart, A.log_q) + inner, A.log_q);
    // The Horner steps read the packed ciphertexts' b; with t == 1 there are
    // none (only out is written).
    if (A.t > 1) A.ct_b[(long long)q * A.M + row] = bagg;
    const int s = oi / A.t, i = oi % A.t;
    if (i == A.t - 1) {
      const u64 a = A.a_tilde[(long long)oi * A.d + m] & A.cmask;
      if (A.t == 1) {
        // Rescale's result is cast to CoeffType (polynomial.cc:297-301).
        u64* o = A.out + (((long long)q * A.ell + s) * 2) * A.d;
        o[m] = RescaleCoeff(a, A.log_q, A.log_q1) & A.cmask;
        o[A.d + m] = RescaleCoeff(bagg, A.log_q, A.log_q2) & A.cmask;
      } else {
        A.st_a[((long long)q * A.ell + s) * A.d + m] = a;
        A.st_b[((long long)q * A.ell + s) * A.d + m] = bagg;
      }
    }
  }
}


// ExternalProduct's decomposition (crypto/external_product.cc:60-72,
// polynomial.cc:340-374) fused with the forward FFT of each digit chunk.
template <int L, bool H>
__global__ __launch_bounds__(Geo<L>::NT) __attribute__((amdgpu_waves_per_eu(H ? 4 : 1))) void DigitsSpectraKernel(
    DigitsArgs A, const double2* __restrict__ st) {
  extern __shared__ __align__(16) double lds[];
  constexpr int N = Geo<L>::N, E = Geo<L>::E, NT = Geo<L>::NT;
  const int t = threadIdx.x;
  const int nu = A.rule.num_chunks;
  const int job = blockIdx.x;                 // ((q*ell + s)*2e + jj)*nu + a
  const int a = job % nu;
  const int jj = (job / nu) % (2 * A.e);
  const long long qs = job / (nu * 2 * A.e);  // q*ell + s
  const int digit = jj % A.e;
  const int to_shift = A.log_q - A.e * A.log_digit;
  u64 x[E / 2];
  if (A.prev_step < 0) {
    const u64* src = (jj < A.e ? A.st_a : A.st_b) + qs * (N / 2) + t;
#pragma unroll
    for (int k = 0; k < E / 2; ++k) x[k] = src[NT * k];
  } else {
    // The previous Horner step's result (pir_server.cc:141-146): the sum of
    // its chunk-pair partials plus c_i, unreduced.
    constexpr int d = N / 2;
    const int which = jj < A.e ? 0 : 1;
    const long long q = qs / A.ell;
    const int s = static_cast<int>(qs % A.ell);
    const long long oi = (long long)s * A.t + A.prev_step;
    const u64* pp = A.ext_partial + ((qs * 2 + which) * A.npairs) * (long long)d + t;
    const u64* cp = (which == 0 ? A.a_tilde + oi * d : A.ct_b + (q * A.ell * A.t + oi) * d) + t;
#pragma unroll
    for (int k = 0; k < E / 2; ++k) x[k] = cp[NT * k];
    for (int pi = 0; pi < A.npairs; ++pi) {
#pragma unroll
      for (int k = 0; k < E / 2; ++k) x[k] += pp[(long long)pi * d + NT * k];
    }
  }
  double2 v[16];
#pragma unroll
  for (int k = 0; k < E / 2; ++k) {
    u64 y = LowBitsD(x[k], A.log_q);
    if (to_shift > 0) {
      y = static_cast<u64>((static_cast<unsigned __int128>(y) + (1ULL << (to_shift - 1))) >>
                           to_shift);
    }
    u64 cur = y;
    for (int r = 0; r < digit; ++r) cur = (A.log_digit >= 64) ? 0 : (cur >> A.log_digit);
    v[InSlot<L>(k)] = make_double2(ChunkValue(A.rule, LowBitsD(cur, A.log_digit), a), 0.0);
  }
#pragma unroll
  for (int k = E / 2; k < E; ++k) v[InSlot<L>(k)] = make_double2(0.0, 0.0);
  Transform<L, H>(v, lds, st);
  double2* out = A.dspec + (long long)job * N + t;
#pragma unroll
  for (int e2 = 0; e2 < E; ++e2) out[NT * e2] = v[OutSlot<L>(e2)];
}


// Forward spectra of DB row polynomials (ComputeHintMatrix,
// hint_computation.cc:120-152): job = (row * polys + k) * nu + a.
template <int L, typename Elem, bool H>
__global__ __launch_bounds__(Geo<L>::NT) __attribute__((amdgpu_waves_per_eu(H ? 4 : 1))) void DbSpectraKernel(
    const Elem* __restrict__ db, long long stride, long long valid, int polys, ChunkRule rule,
    const double2* __restrict__ st, double2* __restrict__ dst) {
  extern __shared__ __align__(16) double lds[];
  constexpr int N = Geo<L>::N, E = Geo<L>::E, NT = Geo<L>::NT, d = N / 2;
  const int t = threadIdx.x;
  const long long job = blockIdx.x;
  const long long pidx = job / rule.num_chunks;
  const int c = static_cast<int>(job % rule.num_chunks);
  const long long x = pidx / polys;
  const long long col0 = (pidx % polys) * d;
  const Elem* row = db + x * stride;
  double2 v[16];
#pragma unroll
  for (int k = 0; k < E / 2; ++k) {
    const long long col = col0 + t + NT * k;
    const u64 val = col < valid ? static_cast<u64>(row[col]) : 0ULL;
    v[InSlot<L>(k)] = make_double2(ChunkValue(rule, val, c), 0.0);
  }
#pragma unroll
  for (int k = E / 2; k < E; ++k) v[InSlot<L>(k)] = make_double2(0.0, 0.0);
  Transform<L, H>(v, lds, st);
  double2* out = dst + job * N + t;
#pragma unroll
  for (int e = 0; e < E; ++e) out[NT * e] = v[OutSlot<L>(e)];
}

// One step of PreprocessPack's reduction loop (lwes_to_rlwe.cc:313-346):
// SignedGadgetInv digits of the chain states (polynomial.cc:376-427) and the
// automorphed packing polynomials (polynomial.cc:429-456), each chunked and
// transformed. Digit jobs come first, then automorphism jobs.
template <int L>
__global__ __launch_bounds__(Geo<L>::NT) void ChainSpectraKernel(
    ChainSpectraArgs A, const double2* __restrict__ st) {
  extern __shared__ __align__(16) double lds[];
  constexpr int N = Geo<L>::N, E = Geo<L>::E, NT = Geo<L>::NT, d = N / 2;
  const int t = threadIdx.x;
  const int nu = A.ur.num_chunks, nv = A.vr.num_chunks;
  const int dig_jobs = A.n_dig * A.k * nu;
  const int job = blockIdx.x;
  double2 v[16];
  double2* out;
  if (job < dig_jobs) {
    const int ch = job / (A.k * nu);
    const int l = (job / nu) % A.k;
    const int a = job % nu;
    const int chain = A.dig_first + ch * A.dig_step;
    const u64* state = A.state + static_cast<long long>(chain) * d;
    const int ld = A.log_digit, ts = A.to_shift;
    const u64 mask = (1ULL << ld) - 1;
    const u64 halfm = 1ULL << (ld - 1);
    uint32_t* d32 = A.digits32 == nullptr
                        ? nullptr
                        : A.digits32 + ((static_cast<long long>(chain >> 1) * A.k + l) * d +
                                        ((chain & 1) ? A.slot_gh : A.slot_g)) * d;
    u64* d64 = A.digits64 == nullptr
                   ? nullptr
                   : A.digits64 + (static_cast<long long>(ch) * A.k + l) * d;
#pragma unroll
    for (int k = 0; k < E / 2; ++k) {
      const int pos = t + NT * k;
      const u64 x = state[pos] & A.qmask;
      // (x + 2^(ts-1)) >> ts without the 65-bit intermediate.
      u64 cur = ts > 0 ? (x >> ts) + ((x >> (ts - 1)) & 1ULL) : x;
      u64 dig = 0;
      for (int jj = 0; jj <= l; ++jj) {
        const u64 lower = cur & mask;
        if (lower >= halfm) {
          dig = static_cast<u64>(static_cast<i64>(lower) - static_cast<i64>(1ULL << ld));
          cur = (cur >> ld) + 1;
        } else {
          dig = lower;
          cur = cur >> ld;
        }
      }
      v[InSlot<L>(k)] = make_double2(ChunkValue(A.ur, dig, a), 0.0);
      if (a == 0) {
        if (d32 != nullptr) d32[pos] = static_cast<uint32_t>(dig);
        if (d64 != nullptr) d64[pos] = dig;
      }
    }
    out = A.uspec + static_cast<long long>(job) * N;
  } else {
    const int jw = job - dig_jobs;
    const int ty = jw / (A.k * nv);
    const int m = (jw / nv) % A.k;
    const int b = jw % nv;
    const u64* w = A.w + static_cast<long long>(m) * d;
    const int pinv = A.pinv[ty];
#pragma unroll
    for (int k = 0; k < E / 2; ++k) {
      const int pos = t + NT * k;
      // Automorph(w, power)[pos] = w[i] for i * power = pos (mod 2d), or
      // -w[i - d] when that i is >= d.
      const int i1 = static_cast<int>((static_cast<long long>(pos) * pinv) & (2 * d - 1));
      const u64 val = i1 < d ? w[i1] : (0ULL - w[i1 - d]);
      v[InSlot<L>(k)] = make_double2(ChunkValue(A.vr, val, b), 0.0);
    }
    out = A.vspec + static_cast<long long>(jw) * N;
  }
#pragma unroll
  for (int k = E / 2; k < E; ++k) v[InSlot<L>(k)] = make_double2(0.0, 0.0);
  Transform<L, false>(v, lds, st);
  out += t;
#pragma unroll
  for (int e = 0; e < E; ++e) out[NT * e] = v[OutSlot<L>(e)];
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
    StridedRun<8> rp;
    rp.Issue(pp + m, A.d, A.npairs);
    u64 v = rp.Finish(pp + m, A.d, A.npairs);
    const long long oi = (long long)s * A.t + A.step_i;
    const u64 c = (which == 0) ? A.a_tilde[oi * A.d + m]
                               : A.ct_b[(q * A.ell * A.t + oi) * A.d + m];
    // Horner's unreduced Add wraps at CoeffType's width (pir_server.cc:141-146);
    // Rescale reads that value and its result is cast to CoeffType.
    v = (v + c) & A.cmask;
    if (A.step_i == 0) {
      A.out[(qs * 2 + which) * A.d + m] =
          RescaleCoeff(v, A.log_q, which == 0 ? A.log_q1 : A.log_q2) & A.cmask;
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

namespace {

int GridFor(long long total) {
  long long b = (total + 255) / 256;
  if (b > 65535LL * 8) b = 65535LL * 8;
  return static_cast<int>(b < 1 ? 1 : b);
}

// Instantiations per log2 N: Fn<L>::Run(args...) for the runtime L.
template <template <int> class Fn, typename... Args>
hipError_t DispatchLogN(int L, Args&&... args) {
  switch (L) {
    case 1: return Fn<1>::Run(args...);
    case 2: return Fn<2>::Run(args...);
    case 3: return Fn<3>::Run(args...);
    case 4: return Fn<4>::Run(args...);
    case 5: return Fn<5>::Run(args...);
    case 6: return Fn<6>::Run(args...);
    case 7: return Fn<7>::Run(args...);
    case 8: return Fn<8>::Run(args...);
    case 9: return Fn<9>::Run(args...);
    case 10: return Fn<10>::Run(args...);
    case 11: return Fn<11>::Run(args...);
    case 12: return Fn<12>::Run(args...);
    case 13: return Fn<13>::Run(args...);
    default: return hipErrorInvalidValue;
  }
}

// Launches of at least this many workgroups use the two-per-CU form of a
// transform kernel (throughput); smaller ones the one-per-CU form, whose
// transform has half the barriers and preloaded twiddles (latency).
constexpr long long kHalfMinBlocks = 256;

template <int L>
struct Configure {
  template <typename K>
  static hipError_t Set(K* kernel, bool half) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               Geo<L>::LdsBytes(half));
  }
  static hipError_t Run() {
    hipError_t e;
    if ((e = Set(SpectraKernel<L, false>, false)) != hipSuccess) return e;
    if ((e = Set(SpectraKernel<L, true>, true)) != hipSuccess) return e;
    if ((e = Set(IpKernel<L, false, false>, false)) != hipSuccess) return e;
    if ((e = Set(IpKernel<L, true, false>, true)) != hipSuccess) return e;
    if ((e = Set(IpKernel<L, false, true>, false)) != hipSuccess) return e;
    if ((e = Set(DbSpectraKernel<L, uint16_t, false>, false)) != hipSuccess) return e;
    if ((e = Set(DbSpectraKernel<L, uint16_t, true>, true)) != hipSuccess) return e;
    if ((e = Set(DbSpectraKernel<L, uint8_t, false>, false)) != hipSuccess) return e;
    if ((e = Set(DbSpectraKernel<L, uint8_t, true>, true)) != hipSuccess) return e;
    if ((e = Set(ChainSpectraKernel<L>, false)) != hipSuccess) return e;
    if ((e = Set(DigitsSpectraKernel<L, false>, false)) != hipSuccess) return e;
    return Set(DigitsSpectraKernel<L, true>, true);
  }
};

template <int L>
struct SpectraL {
  static hipError_t Run(const u64* src, int npolys, int per_group, long long group_stride,
                        const ChunkRule& rule, const double2* st, double2* dst,
                        hipStream_t s) {
    const long long blocks = static_cast<long long>(npolys) * rule.num_chunks;
    if (blocks >= kHalfMinBlocks) {
      hipLaunchKernelGGL((SpectraKernel<L, true>), dim3(blocks), dim3(Geo<L>::NT),
                         Geo<L>::LdsBytes(true), s, src, per_group, group_stride, rule, st,
                         dst);
    } else {
      hipLaunchKernelGGL((SpectraKernel<L, false>), dim3(blocks), dim3(Geo<L>::NT),
                         Geo<L>::LdsBytes(false), s, src, per_group, group_stride, rule, st,
                         dst);
    }
    return hipGetLastError();
  }
};

template <int L>
struct IpL {
  static hipError_t Run(const IpJob& jobs, int n_out, int outs_per_group, const IpTile& tile,
                        int nterms, int nu, int nv, const PairList& pairs, const double2* st,
                        unsigned long long* margin, double2* acc, hipStream_t s) {
    const long long n_jobs = static_cast<long long>(n_out) * pairs.count;
    long long group_jobs = static_cast<long long>(outs_per_group) * pairs.count;
    long long groups = (n_out + outs_per_group - 1) / outs_per_group;
    if (tile.s > 0) {
      const long long nreq = n_out / tile.outs;
      group_jobs = static_cast<long long>(tile.r) * tile.s * pairs.count;
      groups = (nreq + tile.r - 1) / tile.r * ((tile.outs + tile.s - 1) / tile.s);
    }
    const long long blocks = (groups + 7) / 8 * 8 * group_jobs;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (n_jobs >= kHalfMinBlocks) {
      hipLaunchKernelGGL((IpKernel<L, true, false>), dim3(static_cast<unsigned>(blocks)),
                         dim3(Geo<L>::NT), Geo<L>::LdsBytes(true), s, jobs,
                         static_cast<int>(n_jobs), static_cast<int>(group_jobs), tile, nterms,
                         nu, nv,
                         pairs, st, margin, nullptr);
    } else if (acc != nullptr && n_jobs <= kIpSplitMaxJobs) {
      const long long total = n_jobs << L;
      hipLaunchKernelGGL(MacKernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256),
                         0, s, jobs, static_cast<int>(n_jobs), nterms, nu, nv, pairs, L, acc);
      hipLaunchKernelGGL((IpKernel<L, false, true>), dim3(static_cast<unsigned>(blocks)),
                         dim3(Geo<L>::NT), Geo<L>::LdsBytes(false), s, jobs,
                         static_cast<int>(n_jobs), static_cast<int>(group_jobs), tile, nterms,
                         nu, nv,
                         pairs, st, margin, acc);
    } else {
      hipLaunchKernelGGL((IpKernel<L, false, false>), dim3(static_cast<unsigned>(blocks)),
                         dim3(Geo<L>::NT), Geo<L>::LdsBytes(false), s, jobs,
                         static_cast<int>(n_jobs), static_cast<int>(group_jobs), tile, nterms,
                         nu, nv,
                         pairs, st, margin, nullptr);
    }
    return hipGetLastError();
  }
};

template <int L>
struct DbSpectraL {
  template <typename Elem, bool H>
  static void Go(const void* db, long long stride, long long valid, long long jobs, int polys,
                 const ChunkRule& rule, const double2* st, double2* dst, hipStream_t s) {
    hipLaunchKernelGGL((DbSpectraKernel<L, Elem, H>), dim3(static_cast<unsigned>(jobs)),
                       dim3(Geo<L>::NT), Geo<L>::LdsBytes(H), s, static_cast<const Elem*>(db),
                       stride, valid, polys, rule, st, dst);
  }
  static hipError_t Run(const void* db, int elem_bytes, long long stride, long long valid,
                        int rows, int polys, const ChunkRule& rule, const double2* st,
                        double2* dst, hipStream_t s) {
    const long long jobs = static_cast<long long>(rows) * polys * rule.num_chunks;
    if (jobs > 0x7fffffffLL) return hipErrorInvalidValue;
    const bool half = jobs >= kHalfMinBlocks;
    if (elem_bytes == 2) {
      if (half) Go<uint16_t, true>(db, stride, valid, jobs, polys, rule, st, dst, s);
      else Go<uint16_t, false>(db, stride, valid, jobs, polys, rule, st, dst, s);
    } else if (elem_bytes == 1) {
      if (half) Go<uint8_t, true>(db, stride, valid, jobs, polys, rule, st, dst, s);
      else Go<uint8_t, false>(db, stride, valid, jobs, polys, rule, st, dst, s);
    } else {
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
};

template <int L>
struct ChainSpectraL {
  static hipError_t Run(const ChainSpectraArgs& A, const double2* st, hipStream_t s) {
    const int jobs = A.n_dig * A.k * A.ur.num_chunks + A.n_wtypes * A.k * A.vr.num_chunks;
    if (jobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(ChainSpectraKernel<L>, dim3(jobs), dim3(Geo<L>::NT),
                       Geo<L>::LdsBytes(false), s, A, st);
    return hipGetLastError();
  }
};

template <int L>
struct DigitsL {
  static hipError_t Run(const DigitsArgs& A, const double2* st, hipStream_t s) {
    const int jobs = A.nq * A.ell * 2 * A.e * A.rule.num_chunks;
    if (jobs >= kHalfMinBlocks) {
      hipLaunchKernelGGL((DigitsSpectraKernel<L, true>), dim3(jobs), dim3(Geo<L>::NT),
                         Geo<L>::LdsBytes(true), s, A, st);
    } else {
      hipLaunchKernelGGL((DigitsSpectraKernel<L, false>), dim3(jobs), dim3(Geo<L>::NT),
                         Geo<L>::LdsBytes(false), s, A, st);
    }
    return hipGetLastError();
  }
};

bool ValidDims(const FftDims& F) {
  return F.log_N >= 1 && F.log_N <= fftr::kMaxLogN && F.N == (1 << F.log_N) &&
         F.d * 2 == F.N;
}

}  // namespace

hipError_t ConfigureFftKernels(int N) {
  int L = 0;
  while ((1 << L) < N) ++L;
  if ((1 << L) != N) return hipErrorInvalidValue;
  return DispatchLogN<Configure>(L);
}

hipError_t LaunchSpectra(const u64* src, int npolys, int per_group,
                         long long group_stride, const ChunkRule& rule,
                         const FftDims& F, const double2* tw_fwd,
                         double2* dst, hipStream_t s) {
  if (npolys <= 0) return hipSuccess;
  if (!ValidDims(F)) return hipErrorInvalidValue;
  return DispatchLogN<SpectraL>(F.log_N, src, npolys, per_group, group_stride, rule, tw_fwd,
                                dst, s);
}

hipError_t LaunchIp(const double2* const* uptr, const double2* const* vptr,
                    u64* const* optr, int n_out, int outs_per_group, int nterms, int nu,
                    int nv, const PairList& pairs, const FftDims& F,
                    const double2* tw_bwd, unsigned long long* margin,
                    hipStream_t s, double2* acc, int tile_requests, int tile_outs) {
  if (n_out <= 0 || pairs.count <= 0) return hipSuccess;
  if (!ValidDims(F) || outs_per_group <= 0) return hipErrorInvalidValue;
  IpTile tile{outs_per_group, 1, 0};
  if (tile_outs > 0) {
    if (tile_requests < 1 || n_out % outs_per_group != 0) return hipErrorInvalidValue;
    tile = IpTile{outs_per_group, tile_requests, tile_outs};
  }
  IpJob jobs{uptr, vptr, optr};
  return DispatchLogN<IpL>(F.log_N, jobs, n_out, outs_per_group, tile, nterms, nu, nv, pairs,
                           tw_bwd, margin, acc, s);
}

hipError_t LaunchDbSpectra(const void* db, int elem_bytes, long long stride, long long valid_cols,
                           int rows, int polys, const ChunkRule& rule, const FftDims& F,
                           const double2* tw_fwd, double2* dst, hipStream_t s) {
  if (rows <= 0 || polys <= 0) return hipSuccess;
  if (!ValidDims(F)) return hipErrorInvalidValue;
  return DispatchLogN<DbSpectraL>(F.log_N, db, elem_bytes, stride, valid_cols, rows, polys, rule,
                                  tw_fwd, dst, s);
}

hipError_t LaunchChainSpectra(const ChainSpectraArgs& A, const FftDims& F,
                              const double2* tw_fwd, hipStream_t s) {
  if (!ValidDims(F) || A.log_digit < 1 || A.log_digit > 63) return hipErrorInvalidValue;
  return DispatchLogN<ChainSpectraL>(F.log_N, A, tw_fwd, s);
}

hipError_t LaunchDigitsSpectra(const DigitsArgs& A, const FftDims& F,
                               const double2* tw_fwd, hipStream_t s) {
  if (!ValidDims(F)) return hipErrorInvalidValue;
  return DispatchLogN<DigitsL>(F.log_N, A, tw_fwd, s);
}

hipError_t LaunchFinalizeCombine(const FinalizeArgs& A, hipStream_t s) {
  if (A.nq <= 0 || A.M <= 0) return hipSuccess;
  if (A.d <= 0 || (A.d & (A.d - 1)) != 0) return hipErrorInvalidValue;
  const int bx = static_cast<int>(std::min<long long>((A.M + 255) / 256, 8192));
  hipLaunchKernelGGL(FinalizeCombineKernel, dim3(bx, std::min(A.nq, 65535)), dim3(256), 0, s, A);
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
