// Register-resident schedule of crypto/fft.h's radix-2 DIT FFT.
//
// The reference transform (crypto/fft.h:106-129) places input i at position
// bit_rev(i) and runs stages s = 1..L (L = log2 N). Stage s combines
// positions k and k + 2^(s-1) of each 2^s block with twiddle
// tw[(k mod 2^(s-1)) * (N >> s)]: v = y * w, x' = x + v, y' = x - v.
//
// Here one workgroup of NT = N / E threads runs one transform; each thread
// holds E = min(16, N) positions in registers and applies up to four
// consecutive stages to them without leaving registers ("super-stage" K runs
// stages 4K+1 .. 4K+P on groups of 2^P positions spaced 2^(4K) apart). The
// positions are exchanged through LDS between super-stages: for N = 8192 that
// is 4 super-stages and 3 exchanges, instead of 13 passes over LDS. Every
// butterfly is the reference's (same operands, same twiddle value, same
// operation order, no fused multiply-add), so the output is bit-identical;
// only the schedule differs.
//
// Super-stage 0 gives thread t the group bit_rev(t), whose inputs are
// t + NT * k (k < E): the loads of the caller are coalesced and need no
// bit-reversal pass. After the last super-stage, thread t holds positions
// t + NT * e (e < E), again coalesced for the caller's stores.
//
// Twiddles are read from a stage-ordered table: stage s occupies entries
// [2^(s-1) - 1, 2^s - 1) and entry 2^(s-1) - 1 + k equals the reference's
// tw[k * (N >> s)] (a reordered copy of the same doubles), so the loads of
// every super-stage after the first are contiguous across a wave.
#ifndef INSPIRE_AMD_FFT_REGS_H_
#define INSPIRE_AMD_FFT_REGS_H_

#include "pirg_internal.h"

namespace pirg {
namespace fftr {

constexpr int kMaxLogN = 13;  // N = 2d, d <= 4096

constexpr int Brev(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

template <int L>
struct Geo {
  static_assert(L >= 1 && L <= kMaxLogN, "transform size out of range");
  static constexpr int N = 1 << L;
  static constexpr int LE = L < 4 ? L : 4;   // log2 of the positions per thread
  static constexpr int E = 1 << LE;
  static constexpr int NT = N / E;           // threads per transform
  static constexpr int NS = (L + 3) / 4;     // super-stages
  // LDS bytes of the exchanges, one pad slot per N/16 positions: a double2
  // per position (one transform per CU), or a double per position with the
  // real and imaginary parts exchanged in turn (two transforms per CU).
  static constexpr int kLdsFull = NS > 1 ? (N + 16) * 16 : 16;
  static constexpr int kLdsHalf = NS > 1 ? (N + 16) * 8 : 16;
  static constexpr int LdsBytes(bool half) { return half ? kLdsHalf : kLdsFull; }
};

template <int L, int K>
struct Super {
  static constexpr int S0 = 4 * K;
  static constexpr int P = (L - S0) < 4 ? (L - S0) : 4;
  static constexpr int G = Geo<L>::E >> P;   // groups per thread
};

// Padded LDS slot of position x. Positions 16 * bit_rev(t) + m (the stores
// after super-stage 0) and runs of consecutive positions (every other
// exchange) both land in distinct banks across 16 lanes.
template <int L>
__device__ __forceinline__ int LdsSlot(int x) {
  if constexpr (L >= 4) {
    return x + (x >> (L - 4));
  } else {
    return x;
  }
}

// Position held by register idx = i * 2^P + m during super-stage K.
template <int L, int K>
__device__ __forceinline__ int Position(int t, int idx) {
  using S = Super<L, K>;
  constexpr int P = S::P;
  const int i = idx >> P;
  const int m = idx & ((1 << P) - 1);
  if constexpr (K == 0) {
    constexpr int gb = L - Geo<L>::LE;
    int gid = 0;
    if constexpr (gb > 0) gid = static_cast<int>(__brev(static_cast<unsigned>(t)) >> (32 - gb));
    (void)i;  // G == 1 in super-stage 0
    return (gid << P) + m;
  } else {
    const int g = t + Geo<L>::NT * i;
    const int glo = g & ((1 << S::S0) - 1);
    return glo + ((g >> S::S0) << (S::S0 + P)) + (m << S::S0);
  }
}

// Register that receives input index t + NT * k.
template <int L>
constexpr int InSlot(int k) {
  return Brev(k, Geo<L>::LE);
}

// Register that holds output position t + NT * e after the transform.
template <int L>
constexpr int OutSlot(int e) {
  using S = Super<L, Geo<L>::NS - 1>;
  return (e % S::G) * (1 << S::P) + e / S::G;
}

// crypto/fft.h:122-125: v = y * w; x' = x + v; y' = x - v.
__device__ __forceinline__ void Butterfly(double2& x, double2& y, const double2 w) {
  const double vr = y.x * w.x - y.y * w.y;
  const double vi = y.x * w.y + y.y * w.x;
  const double ux = x.x, uy = x.y;
  x = make_double2(ux + vr, uy + vi);
  y = make_double2(ux - vr, uy - vi);
}

// The butterfly at offset 0 of a stage, whose twiddle is tw[0] =
// (cos 0, sin ±0) = (1, ±0): y * tw[0] computes (y.x - y.y * ±0,
// y.x * ±0 + y.y), which equals y except possibly in the sign of a zero
// component. No later operation turns two zeros of opposite sign into
// different nonzero values (sums, differences and products by twiddles of
// the same real numbers round to the same real numbers), and the final
// round-to-integer maps both to 0, so every integer result is unchanged.
__device__ __forceinline__ void ButterflyUnit(double2& x, double2& y) {
  const double ux = x.x, uy = x.y;
  x = make_double2(ux + y.x, uy + y.y);
  y = make_double2(ux - y.x, uy - y.y);
}

// Stages 4K+1 .. 4K+P on every group the thread holds. H (two transforms
// per CU): a stage's twiddles are loaded just before its butterflies, to
// keep few registers live; otherwise all of them up front. In super-stage 0
// every thread's offset-0 butterflies (twiddle tw[0]) are known at compile
// time and skip the complex product: 15 of each thread's 104 butterflies at
// N = 8192.
template <int L, int K, bool H>
__device__ __forceinline__ void Compute(double2 (&v)[16], int t,
                                        const double2* __restrict__ st) {
  using S = Super<L, K>;
  constexpr int P = S::P, S0 = S::S0;
#pragma unroll
  for (int i = 0; i < S::G; ++i) {
    int glo = 0;
    if constexpr (K > 0) glo = (t + Geo<L>::NT * i) & ((1 << S0) - 1);
    if constexpr (H) {
#pragma unroll
      for (int r = 0; r < P; ++r) {
        double2 w[8];
#pragma unroll
        for (int q = (K == 0 ? 1 : 0); q < (1 << r); ++q) {
          w[q] = st[(1 << (S0 + r)) - 1 + glo + (q << S0)];
        }
#pragma unroll
        for (int m = 0; m < (1 << P); ++m) {
          if (m & (1 << r)) continue;
          const int q = m & ((1 << r) - 1);
          if (K == 0 && q == 0) {
            ButterflyUnit(v[i * (1 << P) + m], v[i * (1 << P) + m + (1 << r)]);
          } else {
            Butterfly(v[i * (1 << P) + m], v[i * (1 << P) + m + (1 << r)], w[q]);
          }
        }
      }
    } else {
      double2 w[(1 << P) - 1];
#pragma unroll
      for (int r = 0; r < P; ++r) {
#pragma unroll
        for (int q = (K == 0 ? 1 : 0); q < (1 << r); ++q) {
          w[(1 << r) - 1 + q] = st[(1 << (S0 + r)) - 1 + glo + (q << S0)];
        }
      }
#pragma unroll
      for (int r = 0; r < P; ++r) {
#pragma unroll
        for (int m = 0; m < (1 << P); ++m) {
          if (m & (1 << r)) continue;
          const int q = m & ((1 << r) - 1);
          if (K == 0 && q == 0) {
            ButterflyUnit(v[i * (1 << P) + m], v[i * (1 << P) + m + (1 << r)]);
          } else {
            Butterfly(v[i * (1 << P) + m], v[i * (1 << P) + m + (1 << r)],
                      w[(1 << r) - 1 + q]);
          }
        }
      }
    }
  }
}

// From the register layout of super-stage K to that of K + 1 through LDS:
// whole double2 values, or (H) the real parts first, then the imaginary
// parts, in half the LDS.
template <int L, int K, bool H>
__device__ __forceinline__ void Exchange(double2 (&v)[16], void* lds_raw, int t) {
  constexpr int E = Geo<L>::E;
  if constexpr (K > 0) __syncthreads();  // the previous exchange's reads are done
  if constexpr (H) {
    double* lds = static_cast<double*>(lds_raw);
#pragma unroll
    for (int idx = 0; idx < E; ++idx) lds[LdsSlot<L>(Position<L, K>(t, idx))] = v[idx].x;
    __syncthreads();
#pragma unroll
    for (int idx = 0; idx < E; ++idx) v[idx].x = lds[LdsSlot<L>(Position<L, K + 1>(t, idx))];
    __syncthreads();
#pragma unroll
    for (int idx = 0; idx < E; ++idx) lds[LdsSlot<L>(Position<L, K>(t, idx))] = v[idx].y;
    __syncthreads();
#pragma unroll
    for (int idx = 0; idx < E; ++idx) v[idx].y = lds[LdsSlot<L>(Position<L, K + 1>(t, idx))];
  } else {
    double2* lds = static_cast<double2*>(lds_raw);
#pragma unroll
    for (int idx = 0; idx < E; ++idx) lds[LdsSlot<L>(Position<L, K>(t, idx))] = v[idx];
    __syncthreads();
#pragma unroll
    for (int idx = 0; idx < E; ++idx) v[idx] = lds[LdsSlot<L>(Position<L, K + 1>(t, idx))];
  }
}

template <int L, int K, bool H>
__device__ __forceinline__ void Run(double2 (&v)[16], void* lds,
                                    const double2* __restrict__ st, int t) {
  Compute<L, K, H>(v, t, st);
  if constexpr (K + 1 < Geo<L>::NS) {
    Exchange<L, K, H>(v, lds, t);
    Run<L, K + 1, H>(v, lds, st, t);
  }
}

// The whole transform. On entry v[InSlot(k)] holds input t + NT * k and the
// LDS buffer (Geo<L>::LdsBytes(H) bytes) is free; on return v[OutSlot(e)]
// holds output position t + NT * e. Every thread of the workgroup
// (blockDim.x == NT) must call it.
template <int L, bool H>
__device__ __forceinline__ void Transform(double2 (&v)[16], void* lds,
                                          const double2* __restrict__ st) {
  Run<L, 0, H>(v, lds, st, static_cast<int>(threadIdx.x));
}

}  // namespace fftr
}  // namespace pirg

#endif  // INSPIRE_AMD_FFT_REGS_H_
