// CPU oracle for the ReinsPIRe online answer path — TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of google/InsPIRe's PIR server answer path and of
// the offline preprocessing / client code that produce and consume its
// inputs. It is the checker for the HIP product path (inspire_amd), never part
// of it: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
// load it. Each function cites the reference file:line it restates (paths are
// relative to the reference repository root).
//
// Pinning (what fixes each part independently of this restatement):
//   * matrix products: closed form sum mod 2^64 and the KATs of
//     crypto/matrix_test.cc:82-183 (tests/test_oracle_kats.py);
//   * FFT products: the exact 128-bit schoolbook checker below
//     (orc_inner_product_exact) on the mt19937_64 streams of
//     crypto/polynomial_test.cc:442-612 and the KAT at :297-309;
//   * gadget / shift / rescale ops: KATs of crypto/polynomial_test.cc:140-242;
//   * packing: Pack == MatrixPack (crypto/lwes_to_rlwe_test.cc:364-468) and
//     decrypt(pack) == message;
//   * external product / EvalPoly / ProcessResponse: decrypt(response) == DB
//     record (crypto/pir/server/pir_server_test.cc:246-412);
//   * Interpolate: Interpolate(Evaluate(P)) == P
//     (crypto/db_row_interpolation_test.cc:91-142).
// Seeded streams come from SplitMix64 (a published generator), not from the
// reference's ChaCha PRNG, so parity is defined on identical inputs.
//
// The FFT is compiled with -ffp-contract=off, like the reference's portable
// crypto/fft.h (no fused multiply-add in the butterflies).

#include "pir_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

using u64 = uint64_t;
using i64 = int64_t;
using u128 = unsigned __int128;
using Poly = std::vector<u64>;
using cd = std::complex<double>;
using CVec = std::vector<cd>;

struct OracleError {
  int code;
  std::string msg;
};
[[noreturn]] void Fail(int code, const std::string& m) {
  throw OracleError{code, m};
}
[[noreturn]] void InvalidArg(const std::string& m) { Fail(3, m); }

int Report(const OracleError& e, char* err, size_t errlen) {
  if (err != nullptr && errlen > 0) {
    std::snprintf(err, errlen, "%s", e.msg.c_str());
  }
  return e.code;
}

#define ORC_GUARD_BEGIN try {
#define ORC_GUARD_END(err, errlen)            \
  }                                           \
  catch (const OracleError& e) {              \
    return Report(e, err, errlen);            \
  }                                           \
  catch (const std::exception& e) {           \
    return Report(OracleError{13, e.what()}, err, errlen); \
  }                                           \
  return 0;

bool IsPow2(u64 x) { return x != 0 && (x & (x - 1)) == 0; }
int BitWidth(u64 x) { return x == 0 ? 0 : 64 - __builtin_clzll(x); }
u64 QMask(int log_q) { return log_q >= 64 ? ~0ULL : ((1ULL << log_q) - 1); }

// ---------------------------------------------------------------------------
// SplitMix64 (Steele, Lea, Flood 2014): the seeded stream used for every
// synthetic DB, key and query in this repository.
struct SplitMix64 {
  u64 s;
  explicit SplitMix64(u64 seed) : s(seed) {}
  u64 Next() {
    u64 z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  uint8_t Next8() { return static_cast<uint8_t>(Next() & 0xFF); }
};

// ---------------------------------------------------------------------------
// Polynomial coefficient ops, crypto/polynomial.cc.

// polynomial.cc:254-265 (wraparound add mod 2^64).
Poly Add(const Poly& a, const Poly& b) {
  if (a.size() != b.size()) InvalidArg("Polynomial lengths must match.");
  Poly r(a.size());
  for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] + b[i];
  return r;
}

// polynomial.cc:278-285.
Poly Negate(const Poly& a) {
  Poly r(a.size());
  for (size_t i = 0; i < a.size(); ++i) r[i] = -a[i];
  return r;
}

// polynomial.cc:323-338.
Poly LowBits(const Poly& a, int amount) {
  if (amount < 0) InvalidArg("Amount cannot be negative.");
  if (amount >= 64) return a;
  const u64 mask = (1ULL << amount) - 1;
  Poly r(a.size());
  for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] & mask;
  return r;
}

// polynomial.cc:306-321.
Poly RightShift(const Poly& a, int amount) {
  if (amount < 0) InvalidArg("Shift amount cannot be negative.");
  if (amount >= 64) return Poly(a.size(), 0);
  Poly r(a.size());
  for (size_t i = 0; i < a.size(); ++i) r[i] = a[i] >> amount;
  return r;
}

// polynomial.cc:287-304: round(c * 2^new / 2^cur) in 128-bit, truncated.
Poly Rescale(const Poly& a, int log_cur, int log_new) {
  const u128 cur = static_cast<u128>(1) << log_cur;
  const u128 nw = static_cast<u128>(1) << log_new;
  const u128 half = cur / 2;
  Poly r(a.size());
  for (size_t i = 0; i < a.size(); ++i) {
    r[i] = static_cast<u64>((static_cast<u128>(a[i]) * nw + half) / cur);
  }
  return r;
}

// polynomial.cc:340-374.
std::vector<Poly> GadgetInv(const Poly& a, int log_modulus, int log_digit,
                            int num_digits) {
  if (log_digit <= 0 || num_digits <= 0) {
    InvalidArg("log_digit and num_digits must be positive.");
  }
  if (log_modulus == 0) log_modulus = 64;
  const int to_shift = log_modulus - num_digits * log_digit;
  Poly cur;
  if (to_shift > 0) {
    const u64 round_val = 1ULL << (to_shift - 1);
    cur.resize(a.size());
    for (size_t i = 0; i < a.size(); ++i) {
      cur[i] = static_cast<u64>((static_cast<u128>(a[i]) + round_val) >>
                                to_shift);
    }
  } else {
    cur = a;
  }
  std::vector<Poly> out;
  out.reserve(num_digits);
  for (int i = 0; i < num_digits; ++i) {
    out.push_back(LowBits(cur, log_digit));
    cur = RightShift(cur, log_digit);
  }
  return out;
}

// polynomial.cc:376-427.
std::vector<Poly> SignedGadgetInv(const Poly& a, int log_modulus,
                                  int log_digit, int num_digits) {
  if (log_digit <= 0 || num_digits <= 0) {
    InvalidArg("log_digit and num_digits must be positive.");
  }
  if (log_modulus == 0) log_modulus = 64;
  const int to_shift = log_modulus - num_digits * log_digit;
  Poly cur = a;
  if (to_shift > 0) {
    const u64 round_val = 1ULL << (to_shift - 1);
    for (size_t i = 0; i < a.size(); ++i) {
      cur[i] = static_cast<u64>((static_cast<u128>(cur[i]) + round_val) >>
                                to_shift);
    }
  }
  std::vector<Poly> out;
  out.reserve(num_digits);
  const u64 mask = (1ULL << log_digit) - 1;
  const u64 half_mask = 1ULL << (log_digit - 1);
  for (int i = 0; i < num_digits; ++i) {
    Poly digit(a.size());
    for (size_t j = 0; j < a.size(); ++j) {
      const u64 lower = cur[j] & mask;
      if (lower >= half_mask) {
        digit[j] = static_cast<u64>(static_cast<i64>(lower) -
                                    static_cast<i64>(1ULL << log_digit));
        cur[j] = (cur[j] >> log_digit) + 1;
      } else {
        digit[j] = lower;
        cur[j] = cur[j] >> log_digit;
      }
    }
    out.push_back(std::move(digit));
  }
  return out;
}

// polynomial.cc:429-456.
Poly Automorph(const Poly& a, int power) {
  const int len = static_cast<int>(a.size());
  if (power < 0 || power % 2 == 0 || power >= 2 * len) {
    InvalidArg(
        "Automorphism power must be a non-negative odd integer less than 2*N.");
  }
  Poly r(len, 0);
  for (int i = 0; i < len; ++i) {
    int index = static_cast<int>((static_cast<i64>(i) * power) % (2 * len));
    if (index >= len) {
      r[index - len] -= a[i];
    } else {
      r[index] += a[i];
    }
  }
  return r;
}

// lwes_to_rlwe.cc:41-58: negacyclic shift by X^power, 0 <= power < n.
Poly MultiplyByXPower(const Poly& a, int power) {
  const int n = static_cast<int>(a.size());
  if (power < 0 || power >= n) InvalidArg("Invalid power for cyclic shift.");
  Poly r(n);
  for (int i = 0; i < n; ++i) {
    r[i] = (i < power) ? -a[n - power + i] : a[i - power];
  }
  return r;
}

// Exact negacyclic product of a (any u64) by a ternary s (coefficients 0, 1
// or any value > 1 meaning -1), mod 2^64. This is what Polynomial::Mult with a
// ternary NTT operand returns (polynomial.cc:458-560, ToNtt is_ternary at
// :489-491, signed CRT lift at :544-557) whenever the integer product fits the
// CRT modulus.
Poly MulTernary(const Poly& a, const Poly& s) {
  const size_t d = a.size();
  Poly r(d, 0);
  for (size_t j = 0; j < d; ++j) {
    if (s[j] == 0) continue;
    const bool neg = s[j] > 1;
    for (size_t i = 0; i < d; ++i) {
      size_t idx = i + j;
      bool wrap = idx >= d;
      if (wrap) idx -= d;
      const bool minus = wrap ^ neg;
      if (minus) {
        r[idx] -= a[i];
      } else {
        r[idx] += a[i];
      }
    }
  }
  return r;
}

// Exact negacyclic product mod 2^64 (u64 wraparound schoolbook).
Poly MulExact(const Poly& a, const Poly& b) {
  const size_t d = a.size();
  Poly r(d, 0);
  for (size_t i = 0; i < d; ++i) {
    if (a[i] == 0) continue;
    for (size_t j = 0; j < d; ++j) {
      const u64 p = a[i] * b[j];
      size_t idx = i + j;
      if (idx >= d) {
        r[idx - d] -= p;
      } else {
        r[idx] += p;
      }
    }
  }
  return r;
}

// ---------------------------------------------------------------------------
// crypto/fft.h:50-151: iterative radix-2 DIT complex FFT with bit-reversed
// input, twiddles cos/sin(-2*pi*i/N) (forward) and their conjugate angles
// (backward), no normalisation.
struct FftPlan {
  size_t size;
  std::vector<size_t> bit_rev;
  CVec tw_fwd, tw_bwd;
  explicit FftPlan(size_t n) : size(n) {
    bit_rev.resize(n);
    size_t log_n = 0;
    while ((static_cast<size_t>(1) << log_n) < n) ++log_n;
    for (size_t i = 0; i < n; ++i) {
      size_t rev = 0;
      for (size_t b = 0; b < log_n; ++b) {
        if (i & (static_cast<size_t>(1) << b)) {
          rev |= (static_cast<size_t>(1) << (log_n - 1 - b));
        }
      }
      bit_rev[i] = rev;
    }
    tw_fwd.resize(n / 2);
    tw_bwd.resize(n / 2);
    const double pi = std::acos(-1.0);
    for (size_t i = 0; i < n / 2; ++i) {
      double angle = -2.0 * pi * i / n;
      tw_fwd[i] = cd(std::cos(angle), std::sin(angle));
      tw_bwd[i] = cd(std::cos(-angle), std::sin(-angle));
    }
  }
};

// std::complex<double> multiplication as a build with floating-point
// contraction evaluates it where the target has FMA instructions (GCC's
// default -ffp-contract=fast for C++, or clang's =on, with e.g.
// -march=skylake-avx512): each part's first product is fused with the
// subtraction / addition, (fma(xr, wr, -(xi*wi)), fma(xr, wi, xi*wr)). One
// plausible contracted build of crypto/fft.h:117 and polynomial_fft.cc:162,
// used only to measure which roundings depend on the build.
inline cd MulFused(const cd& x, const cd& w) {
  return cd(std::fma(x.real(), w.real(), -(x.imag() * w.imag())),
            std::fma(x.real(), w.imag(), x.imag() * w.real()));
}

// Set by orc_set_fft_contraction: InnerProductFft (and everything built on
// it: ExternalProduct, FinalizeMatrixPack, ProcessResponse) evaluates its
// complex products as MulFused does. Off by default; on only to measure what
// a contracted build of the reference would answer.
bool g_fft_fused = false;

// fft.h:106-129 (kFused: the products as MulFused evaluates them).
template <bool kFused>
void ComputeFftT(const cd* in, cd* out, const FftPlan& plan, const CVec& tw) {
  const size_t n = plan.size;
  for (size_t i = 0; i < n; ++i) out[plan.bit_rev[i]] = in[i];
  for (size_t len = 2; len <= n; len <<= 1) {
    const size_t half = len >> 1;
    const size_t step = n / len;
    for (size_t i = 0; i < n; i += len) {
      for (size_t j = 0; j < half; ++j) {
        cd u = out[i + j];
        cd v = kFused ? MulFused(out[i + j + half], tw[j * step])
                      : out[i + j + half] * tw[j * step];
        out[i + j] = u + v;
        out[i + j + half] = u - v;
      }
    }
  }
}

void ComputeFft(const cd* in, cd* out, const FftPlan& plan, const CVec& tw) {
  ComputeFftT<false>(in, out, plan, tw);
}

// The chunk-extraction rule of InnerProductFft (polynomial_fft.cc:118-152):
// chunk c of a coefficient is (x >> (chunk_bits*c)) & mask; the top chunk of
// a signed operand is sign-extended from its remaining bits.
struct ChunkRule {
  int chunk_bits, num_chunks, bits;
  bool is_signed;
  double Value(u64 x, int c) const {
    const u64 mask = (1ULL << chunk_bits) - 1;
    const u64 chunk = (x >> (chunk_bits * c)) & mask;
    if (is_signed && c == num_chunks - 1) {
      const int bits_in_chunk = bits - chunk_bits * c;
      const int shift = 64 - bits_in_chunk;
      const i64 sc = static_cast<i64>(chunk << shift) >> shift;
      return static_cast<double>(sc);
    }
    return static_cast<double>(chunk);
  }
  // The integer the chunks represent, mod 2^64 (used by the exact checker).
  u64 Represented(u64 x) const {
    u64 acc = 0;
    const u64 mask = (1ULL << chunk_bits) - 1;
    for (int c = 0; c < num_chunks; ++c) {
      u64 chunk = (x >> (chunk_bits * c)) & mask;
      u64 val = chunk;
      if (is_signed && c == num_chunks - 1) {
        const int bits_in_chunk = bits - chunk_bits * c;
        const int shift = 64 - bits_in_chunk;
        val = static_cast<u64>(static_cast<i64>(chunk << shift) >> shift);
      }
      if (chunk_bits * c < 64) acc += val << (chunk_bits * c);
    }
    return acc;
  }
};

// Forward spectra of every chunk of one polynomial, zero-padded to 2d
// (polynomial_fft.cc:119-152).
template <bool kFused = false>
std::vector<CVec> ChunkSpectra(const u64* coeffs, int d, const ChunkRule& rule,
                               const FftPlan& plan) {
  const int len2 = 2 * d;
  std::vector<CVec> out(rule.num_chunks, CVec(len2));
  CVec in(len2);
  for (int c = 0; c < rule.num_chunks; ++c) {
    for (int j = 0; j < len2; ++j) {
      double val = 0.0;
      if (j < d) val = rule.Value(coeffs[j], c);
      in[j] = cd(val, 0.0);
    }
    ComputeFftT<kFused>(in.data(), out[c].data(), plan, plan.tw_fwd);
  }
  return out;
}

struct IpSetup {
  ChunkRule ur, vr;
};

// Validation and chunk counts of polynomial_fft.cc:72-107.
IpSetup InnerProductSetup(int n, int d, int u_bits, int v_bits, int chunk_bits,
                          bool u_is_signed) {
  if (n <= 0) InvalidArg("Polynomial vectors cannot be empty.");
  if (u_bits <= 0 || v_bits <= 0 || chunk_bits <= 0 || chunk_bits > 63) {
    InvalidArg("Invalid bits or chunk_bits.");
  }
  if (std::log2(static_cast<double>(n)) + std::log2(static_cast<double>(d)) +
          2.0 * chunk_bits >
      53.0) {
    InvalidArg("Potential precision overflow in accumulation.");
  }
  const int max_chunks = (64 + chunk_bits - 1) / chunk_bits;
  int u_num = (u_bits + chunk_bits - 1) / chunk_bits;
  int v_num = (v_bits + chunk_bits - 1) / chunk_bits;
  u_num = std::min(u_num, max_chunks);
  v_num = std::min(v_num, max_chunks);
  IpSetup s;
  s.ur = ChunkRule{chunk_bits, u_num, u_bits, u_is_signed};
  s.vr = ChunkRule{chunk_bits, v_num, v_bits, false};
  return s;
}


// Accumulators of InnerProductFft, one per chunk pair (a, b), indexed
// a * v_num + b (polynomial_fft.cc:109-111).
struct IpAccum {
  IpSetup s;
  int d;
  std::vector<CVec> acc;
  IpAccum(const IpSetup& setup, int d_)
      : s(setup), d(d_),
        acc(setup.ur.num_chunks * setup.vr.num_chunks,
            CVec(2 * d_, cd(0.0, 0.0))) {}
  // polynomial_fft.cc:154-164.
  template <bool kFused = false>
  void Mac(const std::vector<CVec>& us, const std::vector<CVec>& vs) {
    const int len2 = 2 * d;
    for (int a = 0; a < s.ur.num_chunks; ++a) {
      for (int b = 0; b < s.vr.num_chunks; ++b) {
        if (s.ur.chunk_bits * (a + b) >= 64) continue;
        CVec& ac = acc[a * s.vr.num_chunks + b];
        const CVec& ua = us[a];
        const CVec& vb = vs[b];
        for (int j = 0; j < len2; ++j) ac[j] += kFused ? MulFused(ua[j], vb[j]) : ua[j] * vb[j];
      }
    }
  }
  // polynomial_fft.cc:167-196: inverse FFT, divide by 2d, round to int64,
  // shift by chunk_bits*(a+b) and fold X^d = -1.
  template <bool kFused = false>
  Poly Finish(const FftPlan& plan, double* max_err) const {
    const int len2 = 2 * d;
    const int cb = s.ur.chunk_bits;
    Poly result(d, 0);
    CVec out(len2);
    for (int a = 0; a < s.ur.num_chunks; ++a) {
      for (int b = 0; b < s.vr.num_chunks; ++b) {
        if (cb * (a + b) >= 64) continue;
        ComputeFftT<kFused>(acc[a * s.vr.num_chunks + b].data(), out.data(), plan,
                            plan.tw_bwd);
        for (int j = 0; j < len2; ++j) {
          const double vd = out[j].real() / len2;
          const double rd = std::round(vd);
          if (max_err != nullptr) {
            *max_err = std::max(*max_err, std::fabs(vd - rd));
          }
          const u64 val = static_cast<u64>(static_cast<i64>(rd));
          if (j < d) {
            result[j] += (val << (cb * (a + b)));
          } else {
            result[j - d] -= (val << (cb * (a + b)));
          }
        }
      }
    }
    return result;
  }
};

// polynomial_fft.cc:68-197.
Poly InnerProductFft(const std::vector<Poly>& u, const std::vector<Poly>& v,
                     int u_bits, int v_bits, int chunk_bits, bool u_is_signed,
                     double* max_err) {
  if (u.size() != v.size()) {
    InvalidArg("Polynomial vectors must have the same size.");
  }
  if (u.empty()) InvalidArg("Polynomial vectors cannot be empty.");
  const int d = static_cast<int>(u[0].size());
  for (size_t i = 0; i < u.size(); ++i) {
    if (static_cast<int>(u[i].size()) != d ||
        static_cast<int>(v[i].size()) != d) {
      InvalidArg("Polynomial lengths or context mismatch.");
    }
  }
  IpSetup s = InnerProductSetup(static_cast<int>(u.size()), d, u_bits, v_bits,
                                chunk_bits, u_is_signed);
  FftPlan plan(2 * d);
  IpAccum acc(s, d);
  if (g_fft_fused) {
    for (size_t i = 0; i < u.size(); ++i) {
      acc.Mac<true>(ChunkSpectra<true>(u[i].data(), d, s.ur, plan),
                    ChunkSpectra<true>(v[i].data(), d, s.vr, plan));
    }
    return acc.Finish<true>(plan, max_err);
  }
  for (size_t i = 0; i < u.size(); ++i) {
    auto us = ChunkSpectra(u[i].data(), d, s.ur, plan);
    auto vs = ChunkSpectra(v[i].data(), d, s.vr, plan);
    acc.Mac(us, vs);
  }
  return acc.Finish(plan, max_err);
}

// The values InnerProductFft rounds, per accumulated chunk pair: pre[p][j] =
// backward_out[j].real() / len2 (polynomial_fft.cc:176-184) for the p-th pair
// (a, b) with chunk_bits*(a+b) < 64, in the reference's a-major order, before
// round() and before the shift and fold. kFused evaluates every complex
// product (butterflies and the spectra products of :162) with MulFused.
template <bool kFused>
int InnerProductFftPairs(const std::vector<Poly>& u, const std::vector<Poly>& v, int u_bits,
                         int v_bits, int chunk_bits, bool u_is_signed, double* pre,
                         int* pair_a, int* pair_b) {
  const int d = static_cast<int>(u[0].size());
  IpSetup s = InnerProductSetup(static_cast<int>(u.size()), d, u_bits, v_bits, chunk_bits,
                                u_is_signed);
  const int len2 = 2 * d;
  FftPlan plan(len2);
  const int nu = s.ur.num_chunks, nv = s.vr.num_chunks;
  std::vector<CVec> acc(nu * nv, CVec(len2, cd(0.0, 0.0)));
  for (size_t i = 0; i < u.size(); ++i) {
    auto us = ChunkSpectra<kFused>(u[i].data(), d, s.ur, plan);
    auto vs = ChunkSpectra<kFused>(v[i].data(), d, s.vr, plan);
    for (int a = 0; a < nu; ++a) {
      for (int b = 0; b < nv; ++b) {
        if (chunk_bits * (a + b) >= 64) continue;
        CVec& ac = acc[a * nv + b];
        for (int j = 0; j < len2; ++j) {
          ac[j] += kFused ? MulFused(us[a][j], vs[b][j]) : us[a][j] * vs[b][j];
        }
      }
    }
  }
  int np = 0;
  CVec out(len2);
  for (int a = 0; a < nu; ++a) {
    for (int b = 0; b < nv; ++b) {
      if (chunk_bits * (a + b) >= 64) continue;
      ComputeFftT<kFused>(acc[a * nv + b].data(), out.data(), plan, plan.tw_bwd);
      for (int j = 0; j < len2; ++j) pre[static_cast<size_t>(np) * len2 + j] = out[j].real() / len2;
      pair_a[np] = a;
      pair_b[np] = b;
      ++np;
    }
  }
  return np;
}

// Closed-form checker: sum_i (U_i * V_i) mod (X^d + 1, 2^64) where U_i, V_i are
// the integers the chunk rule makes of u_i, v_i (skipped chunk pairs only drop
// multiples of 2^64).
Poly InnerProductExact(const std::vector<Poly>& u, const std::vector<Poly>& v,
                       int u_bits, int v_bits, int chunk_bits,
                       bool u_is_signed) {
  const int d = static_cast<int>(u[0].size());
  IpSetup s = InnerProductSetup(static_cast<int>(u.size()), d, u_bits, v_bits,
                                chunk_bits, u_is_signed);
  Poly result(d, 0);
  for (size_t i = 0; i < u.size(); ++i) {
    Poly uu(d), vv(d);
    for (int j = 0; j < d; ++j) {
      uu[j] = s.ur.Represented(u[i][j]);
      vv[j] = s.vr.Represented(v[i][j]);
    }
    result = Add(result, MulExact(uu, vv));
  }
  return result;
}


// Width-generic InnerProductFft for Polynomial<uint32_t> / <uint64_t>
// (polynomial_fft.cc:68-197 with CoeffType of `width` bits): max_chunks and
// the pair-skip rule use 8*sizeof(CoeffType), results wrap mod 2^width.
u64 WidthMask(int width) { return width >= 64 ? ~0ULL : ((1ULL << width) - 1); }

Poly InnerProductFftW(const std::vector<Poly>& u, const std::vector<Poly>& v,
                      int u_bits, int v_bits, int chunk_bits, bool u_is_signed,
                      int width, double* max_err) {
  if (width == 64) {
    return InnerProductFft(u, v, u_bits, v_bits, chunk_bits, u_is_signed, max_err);
  }
  if (u.size() != v.size()) InvalidArg("Polynomial vectors must have the same size.");
  if (u.empty()) InvalidArg("Polynomial vectors cannot be empty.");
  const int d = static_cast<int>(u[0].size());
  if (u_bits <= 0 || v_bits <= 0 || chunk_bits <= 0 || chunk_bits > 63) {
    InvalidArg("Invalid bits or chunk_bits.");
  }
  if (std::log2(static_cast<double>(u.size())) + std::log2(static_cast<double>(d)) +
          2.0 * chunk_bits > 53.0) {
    InvalidArg("Potential precision overflow in accumulation.");
  }
  const int max_chunks = (width + chunk_bits - 1) / chunk_bits;
  IpSetup s;
  s.ur = ChunkRule{chunk_bits, std::min((u_bits + chunk_bits - 1) / chunk_bits, max_chunks),
                   u_bits, u_is_signed};
  s.vr = ChunkRule{chunk_bits, std::min((v_bits + chunk_bits - 1) / chunk_bits, max_chunks),
                   v_bits, false};
  FftPlan plan(2 * d);
  const int len2 = 2 * d;
  std::vector<CVec> acc(s.ur.num_chunks * s.vr.num_chunks, CVec(len2, cd(0.0, 0.0)));
  const u64 wm = WidthMask(width);
  for (size_t i = 0; i < u.size(); ++i) {
    Poly uu(d), vv(d);
    for (int j = 0; j < d; ++j) {
      uu[j] = u[i][j] & wm;
      vv[j] = v[i][j] & wm;
    }
    auto us = ChunkSpectra(uu.data(), d, s.ur, plan);
    auto vs = ChunkSpectra(vv.data(), d, s.vr, plan);
    for (int a = 0; a < s.ur.num_chunks; ++a) {
      for (int b = 0; b < s.vr.num_chunks; ++b) {
        if (chunk_bits * (a + b) >= width) continue;
        CVec& ac = acc[a * s.vr.num_chunks + b];
        for (int j = 0; j < len2; ++j) ac[j] += us[a][j] * vs[b][j];
      }
    }
  }
  Poly result(d, 0);
  CVec out(len2);
  for (int a = 0; a < s.ur.num_chunks; ++a) {
    for (int b = 0; b < s.vr.num_chunks; ++b) {
      if (chunk_bits * (a + b) >= width) continue;
      ComputeFft(acc[a * s.vr.num_chunks + b].data(), out.data(), plan, plan.tw_bwd);
      for (int j = 0; j < len2; ++j) {
        const double vd = out[j].real() / len2;
        const double rd = std::round(vd);
        if (max_err != nullptr) *max_err = std::max(*max_err, std::fabs(vd - rd));
        const u64 val = static_cast<u64>(static_cast<i64>(rd)) & wm;
        const u64 sh = (val << (chunk_bits * (a + b))) & wm;
        if (j < d) {
          result[j] = (result[j] + sh) & wm;
        } else {
          result[j - d] = (result[j - d] - sh) & wm;
        }
      }
    }
  }
  return result;
}

Poly InnerProductExactW(const std::vector<Poly>& u, const std::vector<Poly>& v,
                        int u_bits, int v_bits, int chunk_bits, bool u_is_signed,
                        int width) {
  const int d = static_cast<int>(u[0].size());
  const int max_chunks = (width + chunk_bits - 1) / chunk_bits;
  ChunkRule ur{chunk_bits, std::min((u_bits + chunk_bits - 1) / chunk_bits, max_chunks),
               u_bits, u_is_signed};
  ChunkRule vr{chunk_bits, std::min((v_bits + chunk_bits - 1) / chunk_bits, max_chunks),
               v_bits, false};
  const u64 wm = WidthMask(width);
  Poly result(d, 0);
  for (size_t i = 0; i < u.size(); ++i) {
    Poly uu(d), vv(d);
    for (int j = 0; j < d; ++j) {
      uu[j] = ur.Represented(u[i][j] & wm);
      vv[j] = vr.Represented(v[i][j] & wm);
    }
    result = Add(result, MulExact(uu, vv));
  }
  for (auto& x : result) x &= wm;
  return result;
}

// lwes_to_rlwe.cc:60-74: chunk = min(20, floor((53 - log2 n - log2 d) / 2)).
Poly InnerProductHelper(const std::vector<Poly>& a, const std::vector<Poly>& b,
                        int a_bits, int b_bits, bool a_is_signed,
                        double* max_err, int width = 64) {
  if (a.size() != b.size() || a.empty()) {
    InvalidArg("InnerProduct: size mismatch or empty");
  }
  const int max_allowed = static_cast<int>(
      (53.0 - std::log2(static_cast<double>(a.size())) -
       std::log2(static_cast<double>(a[0].size()))) /
      2.0);
  const int chunk_bits = std::min(20, max_allowed);
  return InnerProductFftW(a, b, a_bits, b_bits, chunk_bits, a_is_signed, width,
                          max_err);
}

// ---------------------------------------------------------------------------
// Matrix products, crypto/matrix.cc and crypto/matrix.h.

// matrix.cc:261-318. Word w of a row holds columns 4w..4w+3 in 16-bit lanes,
// least significant first; the product is the wraparound sum.
__attribute__((target_clones("avx512f", "avx2", "default")))
void CondensedMulU16Rows(const u64* cdata, int cols, int row0, int nrows,
                         const u64* vec, u64* out) {
  const size_t wpr = static_cast<size_t>(cols) / 4;
  for (int i = row0; i < row0 + nrows; ++i) {
    const u64* row = cdata + static_cast<size_t>(i) * wpr;
    u64 acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    for (size_t w = 0; w < wpr; ++w) {
      const u64 m = row[w];
      const u64* v = vec + 4 * w;
      acc0 += (m & 0xFFFF) * v[0];
      acc1 += ((m >> 16) & 0xFFFF) * v[1];
      acc2 += ((m >> 32) & 0xFFFF) * v[2];
      acc3 += (m >> 48) * v[3];
    }
    out[i - row0] = acc0 + acc1 + acc2 + acc3;
  }
}

// matrix.cc:376-462 (8 unsigned bytes per word).
__attribute__((target_clones("avx512f", "avx2", "default")))
void CondensedMulU8Rows(const u64* cdata, int cols, int row0, int nrows,
                        const u64* vec, u64* out) {
  const size_t wpr = static_cast<size_t>(cols) / 8;
  for (int i = row0; i < row0 + nrows; ++i) {
    const u64* row = cdata + static_cast<size_t>(i) * wpr;
    u64 acc = 0;
    for (size_t w = 0; w < wpr; ++w) {
      const u64 m = row[w];
      for (int k = 0; k < 8; ++k) acc += ((m >> (8 * k)) & 0xFF) * vec[8 * w + k];
    }
    out[i - row0] = acc;
  }
}

// matrix.cc:320-374: two int32 lanes per word, each sign-extended to 64 bits.
__attribute__((target_clones("avx512f", "avx2", "default")))
void CondensedMulI32Rows(const u64* cdata, int cols, int row0, int nrows,
                         const u64* vec, u64* out) {
  const size_t wpr = static_cast<size_t>(cols) / 2;
  for (int i = row0; i < row0 + nrows; ++i) {
    const u64* row = cdata + static_cast<size_t>(i) * wpr;
    u64 acc0 = 0, acc1 = 0;
    for (size_t w = 0; w < wpr; ++w) {
      const i64 ms = static_cast<i64>(row[w]);
      const u64 m1 = static_cast<u64>(ms >> 32);
      const u64 m0 = static_cast<u64>((ms << 32) >> 32);
      acc0 += m0 * vec[2 * w];
      acc1 += m1 * vec[2 * w + 1];
    }
    out[i - row0] = acc0 + acc1;
  }
}

// matrix.h:183-194: plain product with operands cast to uint64_t (int16 is
// sign-extended by the cast).
__attribute__((target_clones("avx512f", "avx2", "default")))
void MulI16(const int16_t* data, int rows, int cols, const u64* vec, u64* out) {
  for (int i = 0; i < rows; ++i) {
    u64 acc = 0;
    const int16_t* row = data + static_cast<size_t>(i) * cols;
    for (int j = 0; j < cols; ++j) acc += static_cast<u64>(row[j]) * vec[j];
    out[i] = acc;
  }
}

// Threads for the DB and pack products of ProcessResponse (1: the
// reference's single-threaded loop). Rows are independent, so splitting them
// changes no result.
int g_matrix_threads = 1;

// Runs f(row0, nrows) over [0, rows) in g_matrix_threads contiguous slices.
template <class F>
void ParallelRows(int rows, F f) {
  const int n = std::max(1, std::min(g_matrix_threads, rows));
  if (n == 1) {
    f(0, rows);
    return;
  }
  std::vector<std::thread> pool;
  const int per = (rows + n - 1) / n;
  for (int r0 = 0; r0 < rows; r0 += per) {
    pool.emplace_back(f, r0, std::min(per, rows - r0));
  }
  for (auto& th : pool) th.join();
}

// The pack matrix of the server in either representation.
// Matrix<uint64_t>::Multiply on a plain matrix (matrix.h:182-192): the
// wraparound sum of entry times vector element.
void MulU64(const u64* data, int rows, int cols, const u64* vec, u64* out) {
  for (int i = 0; i < rows; ++i) {
    const u64* row = data + static_cast<size_t>(i) * cols;
    u64 acc = 0;
    for (int j = 0; j < cols; ++j) acc += row[j] * vec[j];
    out[i] = acc;
  }
}

struct PackView {
  const u64* i32_cdata = nullptr;   // condensed int32 (mat_bytes == 4)
  const int16_t* i16 = nullptr;     // plain int16 (mat_bytes == 2)
  const u64* u64p = nullptr;        // plain uint64 (mat_bytes == 8)
  void Multiply(int row0, int nrows, int cols, const u64* vec, u64* out) const {
    if (i32_cdata != nullptr) {
      CondensedMulI32Rows(i32_cdata, cols, row0, nrows, vec, out);
    } else if (u64p != nullptr) {
      MulU64(u64p + static_cast<size_t>(row0) * cols, nrows, cols, vec, out);
    } else {
      MulI16(i16 + static_cast<size_t>(row0) * cols, nrows, cols, vec, out);
    }
  }
};

// The pack matrix of a PirServer<Db, Coeff, Mat>: int32_t (condensed words,
// mat_bytes 4; a plain uint32_t matrix has the same bytes), int16_t (2) or
// uint64_t (8).
PackView MakePackView(int mat_bytes, const void* pack) {
  PackView pv;
  if (mat_bytes == 4) {
    pv.i32_cdata = static_cast<const u64*>(pack);
  } else if (mat_bytes == 8) {
    pv.u64p = static_cast<const u64*>(pack);
  } else {
    pv.i16 = static_cast<const int16_t*>(pack);
  }
  return pv;
}

// ---------------------------------------------------------------------------
// Parameters.
struct Derived {
  int d, log_d, t, ell, k, e, r, S, K, M, n;
  int log_q, log_q1, log_q2;
  u64 qmask;
  int width;  // bits of CoeffType
  u64 wmask;  // 2^width - 1
};

Derived Derive(const orc_params* P) {
  if (P->log_degree < 1 || P->log_degree > 16) InvalidArg("Invalid log_degree.");
  if (P->log_q < 1 || P->log_q > 64) InvalidArg("Invalid log_q.");
  if (P->coeff_bits != 0 && P->coeff_bits != 32 && P->coeff_bits != 64) {
    InvalidArg("coeff_bits must be 32 or 64.");
  }
  Derived D;
  D.width = P->coeff_bits == 32 ? 32 : 64;
  if (P->log_q > D.width) InvalidArg("log_q exceeds the coefficient width.");
  D.wmask = D.width == 64 ? ~0ULL : ((1ULL << D.width) - 1);
  D.d = 1 << P->log_degree;
  D.log_d = P->log_degree;
  D.t = P->interp_degree;
  D.ell = P->entry_size_multiple;
  D.k = P->pack_num_digits;
  D.e = P->eval_num_digits;
  D.n = P->num_entries;
  // pir_params.cc:35-62.
  if (D.n <= 0) InvalidArg("num_entries must be strictly positive.");
  if (D.ell <= 0) InvalidArg("entry_size_multiple must be strictly positive.");
  if (D.t <= 0 || !IsPow2(static_cast<u64>(D.t))) {
    InvalidArg("interpolation_degree must be a power of two.");
  }
  if (D.n % D.t != 0) {
    InvalidArg("interpolation_degree (" + std::to_string(D.t) +
               ") must divide num_entries (" + std::to_string(D.n) + ").");
  }
  if (D.t > 2 * D.d) {
    InvalidArg("interpolation_degree (" + std::to_string(D.t) +
               ") must be less than or equal to 2 * ring degree (" +
               std::to_string(2 * D.d) + ").");
  }
  D.r = D.n / D.t;
  D.S = (D.r + D.d - 1) / D.d;
  const int simd = (P->db_bytes == 1) ? 64 : 32;
  D.K = ((D.r + simd - 1) / simd) * simd;
  D.M = D.ell * D.t * D.d;
  D.log_q = P->log_q;
  D.log_q1 = P->log_q1;
  D.log_q2 = P->log_q2;
  D.qmask = QMask(P->log_q);
  return D;
}

// ---------------------------------------------------------------------------
// Packing, crypto/lwes_to_rlwe.cc.

// lwes_to_rlwe.cc:448-479.
void FinalizeMatrixPack(const orc_params* P, int d, const u64* b,
                        const u64* b_part, const std::vector<Poly>& y_vec_h,
                        const std::vector<Poly>& t_vec_h, Poly* out_b,
                        double* max_err) {
  // CoeffType uint32_t: the sums wrap at 2^32 and the inner product is
  // Polynomial<uint32_t>'s; LowBits to log_q <= 32 then reads the same bits.
  const int width = P->coeff_bits == 32 ? 32 : 64;
  const u64 wm = WidthMask(width);
  Poly b_agg(d);
  for (int m = 0; m < d; ++m) b_agg[m] = (b[m] + b_part[m]) & wm;
  b_agg = LowBits(b_agg, P->log_q);
  Poly inner_h = InnerProductHelper(t_vec_h, y_vec_h, P->pack_log_digit,
                                    P->log_q, true, max_err, width);
  b_agg = Add(b_agg, inner_h);
  for (auto& x : b_agg) x &= wm;
  *out_b = LowBits(b_agg, P->log_q);
}

// lwes_to_rlwe.cc:76-83.
int ComputePower(int k, int two_n) {
  i64 res = 1;
  for (int i = 0; i < k; ++i) res = (res * 5) % two_n;
  return static_cast<int>(res);
}

// lwes_to_rlwe.cc:88-124: in-place "FFT" over Z_q[X]/(X^d+1) whose roots of
// unity are powers of X^2 (evaluates sum c[i] Z^i at Z = X^{2k}).
void EvaluateNttInPlace(std::vector<Poly>& c) {
  const int d = static_cast<int>(c.size());
  const int log2_d = BitWidth(static_cast<u64>(d - 1));
  for (int i = 0; i < d; ++i) {
    int rev = 0;
    for (int b = 0; b < log2_d; ++b) {
      if ((i >> b) & 1) rev |= (1 << (log2_d - 1 - b));
    }
    if (i < rev) std::swap(c[i], c[rev]);
  }
  for (int len = 2; len <= d; len *= 2) {
    const int half = len / 2;
    const int step = d / len;
    for (int i = 0; i < d; i += len) {
      for (int j = 0; j < half; ++j) {
        Poly u = c[i + j];
        Poly v = MultiplyByXPower(c[i + j + half], 2 * j * step);
        c[i + j] = Add(u, v);
        c[i + j + half] = Add(u, Negate(v));
      }
    }
  }
}

struct PreprocessPackOut {
  Poly a_tilde_agg;
  std::vector<std::vector<Poly>> t_vec_g, t_vec_gh;
  std::vector<Poly> t_vec_h;
};

// lwes_to_rlwe.cc:237-368. a_is: d LWE a-vectors of length d.
PreprocessPackOut PreprocessPack(const orc_params* P,
                                 const std::vector<Poly>& a_is,
                                 const std::vector<Poly>& w_vec_g,
                                 const std::vector<Poly>& w_vec_h) {
  const int d = static_cast<int>(a_is.size());
  const int log_q = P->log_q;
  const int log2_d = BitWidth(static_cast<u64>(d - 1));
  const int ld = P->pack_log_digit, nd = P->pack_num_digits;

  std::vector<Poly> a_tildes(d);
  for (int i = 0; i < d; ++i) {
    Poly coeffs(d, 0);
    coeffs[0] = a_is[i][0];
    for (int j = 1; j < d; ++j) coeffs[d - j] = -a_is[i][j];
    Poly p = LowBits(coeffs, log_q);
    p = RightShift(p, log2_d);
    a_tildes[i] = LowBits(p, log_q);
  }

  std::vector<Poly> a_hat_agg_prime(d);
  for (int k = 0; k < d; ++k) {
    Poly pk(d, 0);
    for (int j = 0; j < d; ++j) pk[j] = a_tildes[j][k];
    a_hat_agg_prime[k] = MultiplyByXPower(pk, k);
  }
  EvaluateNttInPlace(a_hat_agg_prime);

  std::vector<Poly> a_hat_agg(d);
  for (int k = 0; k < d / 2; ++k) {
    const int power = ComputePower(k, 2 * d);
    a_hat_agg[k] = a_hat_agg_prime[(power - 1) / 2];
  }
  for (int k = d / 2; k < d; ++k) {
    int power = ComputePower(k - d / 2, 2 * d);
    power = static_cast<int>((static_cast<i64>(power) * (2 * d - 1)) % (2 * d));
    a_hat_agg[k] = a_hat_agg_prime[(power - 1) / 2];
  }

  Poly t_g = a_hat_agg[d / 2 - 1];
  Poly t_gh = a_hat_agg[d - 1];
  PreprocessPackOut out;
  out.t_vec_g.resize(std::max(0, d / 2 - 1));
  out.t_vec_gh.resize(std::max(0, d / 2 - 1));
  for (int i = d / 2 - 2; i >= 0; --i) {
    Poly t_g_mod = LowBits(t_g, log_q);
    std::vector<Poly> t_vec_g_i = SignedGadgetInv(t_g_mod, log_q, ld, nd);
    out.t_vec_g[i] = t_vec_g_i;
    const int power_g = ComputePower(i, 2 * d);
    std::vector<Poly> w_i_g(w_vec_g.size());
    for (size_t k = 0; k < w_vec_g.size(); ++k) {
      w_i_g[k] = Automorph(w_vec_g[k], power_g);
    }
    Poly inner_g = InnerProductHelper(t_vec_g_i, w_i_g, ld, log_q, true, nullptr);
    t_g = Add(a_hat_agg[i], inner_g);

    const int power_gh =
        static_cast<int>((static_cast<i64>(power_g) * (2 * d - 1)) % (2 * d));
    std::vector<Poly> w_i_gh(w_vec_g.size());
    for (size_t k = 0; k < w_vec_g.size(); ++k) {
      w_i_gh[k] = Automorph(w_vec_g[k], power_gh);
    }
    Poly t_gh_mod = LowBits(t_gh, log_q);
    std::vector<Poly> t_vec_gh_i = SignedGadgetInv(t_gh_mod, log_q, ld, nd);
    out.t_vec_gh[i] = t_vec_gh_i;
    Poly inner_gh =
        InnerProductHelper(t_vec_gh_i, w_i_gh, ld, log_q, true, nullptr);
    t_gh = Add(a_hat_agg[d / 2 + i], inner_gh);
  }
  Poly t_gh_mod = LowBits(t_gh, log_q);
  out.t_vec_h = SignedGadgetInv(t_gh_mod, log_q, ld, nd);
  Poly inner_h =
      InnerProductHelper(out.t_vec_h, w_vec_h, ld, log_q, true, nullptr);
  out.a_tilde_agg = LowBits(Add(t_g, inner_h), log_q);
  return out;
}

// lwes_to_rlwe.cc:163-235: the d x (k*d) matrix of the map y_vec_g ->
// sum_j t_vec_g[j] Aut(y, 5^j) + t_vec_gh[j] Aut(y, (2d-1)5^j), entries
// centred mod 2^log_q and truncated to the matrix coefficient type.
// Returns row-major int64 values (already truncated to mat_bytes).
std::vector<i64> BuildPreprocessMatrix(const PreprocessPackOut& pp, int d,
                                       int k, int log_q, int mat_bytes) {
  std::vector<i64> m(static_cast<size_t>(d) * k * d, 0);
  for (int l = 0; l < k; ++l) {
    std::vector<Poly> t(d, Poly(d, 0));
    for (int j = 0; j <= d / 2 - 2; ++j) {
      const int power_g = ComputePower(j, 2 * d);
      t[(power_g - 1) / 2] = pp.t_vec_g[j][l];
      const int power_gh =
          static_cast<int>((static_cast<i64>(power_g) * (2 * d - 1)) % (2 * d));
      t[(power_gh - 1) / 2] = pp.t_vec_gh[j][l];
    }
    EvaluateNttInPlace(t);
    for (int c = 0; c < d; ++c) {
      Poly fp = MultiplyByXPower(t[c], c);
      for (int r = 0; r < d; ++r) {
        const u64 threshold =
            (log_q >= 64) ? (1ULL << 63) : (1ULL << (log_q - 1));
        u64 val = fp[r];
        if (val >= threshold && log_q < 64) val = val - (1ULL << log_q);
        i64 v;
        if (mat_bytes == 8) {
          v = static_cast<i64>(val);
        } else if (mat_bytes == 4) {
          v = static_cast<int32_t>(static_cast<uint32_t>(val));
        } else {
          v = static_cast<int16_t>(static_cast<uint16_t>(val));
        }
        m[static_cast<size_t>(r) * k * d + static_cast<size_t>(l) * d + c] = v;
      }
    }
  }
  return m;
}

// matrix.cc:124-140 (int32 condense: two values per word, low first).
void CondenseI32(const i64* vals, int rows, int cols, u64* out) {
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols / 2; ++j) {
      u64 w = 0;
      for (int kk = 0; kk < 2; ++kk) {
        const u64 v = static_cast<u64>(static_cast<uint32_t>(
            static_cast<int32_t>(vals[static_cast<size_t>(i) * cols + 2 * j + kk])));
        w |= (v & 0xFFFFFFFFULL) << (kk * 32);
      }
      out[static_cast<size_t>(i) * (cols / 2) + j] = w;
    }
  }
}

void StorePack(const orc_params* P, const std::vector<i64>& m, int rows,
               int cols, void* dst, size_t row_offset) {
  if (P->mat_bytes == 4) {
    CondenseI32(m.data(), rows, cols,
                static_cast<u64*>(dst) + row_offset * (cols / 2));
  } else if (P->mat_bytes == 8) {
    u64* o = static_cast<u64*>(dst) + row_offset * cols;
    for (size_t i = 0; i < m.size(); ++i) o[i] = static_cast<u64>(m[i]);
  } else {
    int16_t* o = static_cast<int16_t*>(dst) + row_offset * cols;
    for (size_t i = 0; i < m.size(); ++i) o[i] = static_cast<int16_t>(m[i]);
  }
}

// lwes_to_rlwe.cc:370-419.
void PackReference(const orc_params* P, const PreprocessPackOut& pp,
                   const Poly& b, const std::vector<Poly>& y_vec_g,
                   const std::vector<Poly>& y_vec_h, Poly* out_b) {
  const int log_q = P->log_q;
  const int d = static_cast<int>(b.size());
  const int ld = P->pack_log_digit;
  Poly b_agg = LowBits(b, log_q);
  for (int j = 0; j <= d / 2 - 2; ++j) {
    const int power_g = ComputePower(j, 2 * d);
    std::vector<Poly> yj(y_vec_g.size());
    for (size_t k = 0; k < y_vec_g.size(); ++k) yj[k] = Automorph(y_vec_g[k], power_g);
    b_agg = LowBits(Add(b_agg, InnerProductHelper(pp.t_vec_g[j], yj, ld, log_q,
                                                  true, nullptr)),
                    log_q);
    const int power_gh =
        static_cast<int>((static_cast<i64>(power_g) * (2 * d - 1)) % (2 * d));
    std::vector<Poly> yjh(y_vec_g.size());
    for (size_t k = 0; k < y_vec_g.size(); ++k) {
      yjh[k] = Automorph(y_vec_g[k], power_gh);
    }
    b_agg = LowBits(Add(b_agg, InnerProductHelper(pp.t_vec_gh[j], yjh, ld,
                                                  log_q, true, nullptr)),
                    log_q);
  }
  b_agg = LowBits(Add(b_agg, InnerProductHelper(pp.t_vec_h, y_vec_h, ld, log_q,
                                                true, nullptr)),
                  log_q);
  *out_b = b_agg;
}

// ---------------------------------------------------------------------------
// External product / EvalPoly.

// external_product.cc:35-114. rgsw_a / rgsw_b: 2e polys each, ordered
// (u_0..u_{e-1}, v_0..v_{e-1}).
void ExternalProduct(const orc_params* P, int d, const Poly& ct_a,
                     const Poly& ct_b, const std::vector<Poly>& rgsw_a,
                     const std::vector<Poly>& rgsw_b, Poly* out_a, Poly* out_b,
                     double* max_err) {
  const int e = P->eval_num_digits;
  if (static_cast<int>(ct_a.size()) != d || static_cast<int>(ct_b.size()) != d) {
    InvalidArg("ct lengths must match d.");
  }
  if (static_cast<int>(rgsw_a.size()) != 2 * e ||
      static_cast<int>(rgsw_b.size()) != 2 * e) {
    InvalidArg("rgsw_ct sizes must match num_digits.");
  }
  const int log_q = P->log_q;
  Poly a_mod = LowBits(ct_a, log_q);
  std::vector<Poly> a_inv = GadgetInv(a_mod, log_q, P->eval_log_digit, e);
  Poly b_mod = LowBits(ct_b, log_q);
  std::vector<Poly> b_inv = GadgetInv(b_mod, log_q, P->eval_log_digit, e);
  std::vector<Poly> u_joined, va, vb;
  for (int i = 0; i < e; ++i) {
    u_joined.push_back(a_inv[i]);
    va.push_back(rgsw_a[i]);
    vb.push_back(rgsw_b[i]);
  }
  for (int i = 0; i < e; ++i) {
    u_joined.push_back(b_inv[i]);
    va.push_back(rgsw_a[e + i]);
    vb.push_back(rgsw_b[e + i]);
  }
  const int max_allowed = static_cast<int>(
      (53.0 - std::log2(static_cast<double>(2 * e)) -
       std::log2(static_cast<double>(d))) /
      2.0);
  const int chunk_bits = std::min(20, max_allowed);
  const int width = P->coeff_bits == 32 ? 32 : 64;
  *out_a = InnerProductFftW(u_joined, va, P->eval_log_digit, log_q, chunk_bits,
                            false, width, max_err);
  *out_b = InnerProductFftW(u_joined, vb, P->eval_log_digit, log_q, chunk_bits,
                            false, width, max_err);
}

double NowMs() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---------------------------------------------------------------------------
// Offline preprocessing.

// db_row_interpolation.cc:43-78 (mod p, odd p).
std::vector<u64> MulXPowerModP(const std::vector<u64>& poly, int power, u64 p) {
  const int d = static_cast<int>(poly.size());
  power %= (2 * d);
  if (power < 0) power += 2 * d;
  bool negate_all = false;
  if (power >= d) {
    negate_all = true;
    power -= d;
  }
  std::vector<u64> r(d);
  for (int i = 0; i < d; ++i) {
    bool neg = negate_all;
    u64 val;
    if (i < power) {
      val = poly[d - power + i];
      neg = !neg;
    } else {
      val = poly[i - power];
    }
    r[i] = neg ? (val == 0 ? 0 : p - val) : val;
  }
  return r;
}

// db_row_interpolation.cc:85-111.
u64 ModInverse(u64 a, u64 m) {
  i64 t = 0, new_t = 1;
  i64 r = static_cast<i64>(m), new_r = static_cast<i64>(a);
  while (new_r != 0) {
    const i64 q = r / new_r;
    i64 tmp = t - q * new_t;
    t = new_t;
    new_t = tmp;
    tmp = r - q * new_r;
    r = new_r;
    new_r = tmp;
  }
  if (r > 1) InvalidArg("a is not invertible modulo m (not coprime).");
  if (t < 0) t += static_cast<i64>(m);
  return static_cast<u64>(t);
}

// db_row_interpolation.cc:115-228: inverse FFT over Z_p[X]/(X^d+1) with
// omega = X^{2d/t}.
std::vector<u64> Interpolate(const std::vector<u64>& row, int d, int t, u64 p) {
  if (static_cast<int>(row.size()) != d * t) {
    InvalidArg("Length of evals must be ring_degree * interpolation_degree.");
  }
  if (!IsPow2(static_cast<u64>(d))) InvalidArg("ring_degree must be a power of two.");
  if (!IsPow2(static_cast<u64>(t))) {
    InvalidArg("interpolation_degree must be a power of two.");
  }
  if (t > 2 * d) {
    InvalidArg("interpolation_degree must be less than or equal to 2 * ring_degree.");
  }
  if (t == 1) return row;
  if (p % 2 == 0) InvalidArg("plaintext_modulus must be odd.");
  const int log_t = __builtin_ctz(static_cast<unsigned>(t));
  auto bit_reverse = [](int n, int bits) {
    int res = 0;
    for (int b = 0; b < bits; ++b) res = (res << 1) | ((n >> b) & 1);
    return res;
  };
  std::vector<std::vector<u64>> polys(t);
  for (int i = 0; i < t; ++i) {
    std::vector<u64> poly(d);
    for (int j = 0; j < d; ++j) poly[j] = row[static_cast<size_t>(i) * d + j] % p;
    polys[bit_reverse(i, log_t)] = std::move(poly);
  }
  for (int len = 2; len <= t; len <<= 1) {
    const int half = len / 2;
    const int step = (2 * d - 2 * d / len) % (2 * d);
    for (int i = 0; i < t; i += len) {
      int cur = 0;
      for (int j = 0; j < half; ++j) {
        const std::vector<u64>& u = polys[i + j];
        std::vector<u64> v = MulXPowerModP(polys[i + j + half], cur, p);
        std::vector<u64> nu(d), nv(d);
        for (int c = 0; c < d; ++c) {
          nu[c] = (u[c] + v[c]) % p;
          nv[c] = (u[c] + p - v[c]) % p;
        }
        polys[i + j] = std::move(nu);
        polys[i + j + half] = std::move(nv);
        cur = (cur + step) % (2 * d);
      }
    }
  }
  const u64 t_inv = ModInverse(static_cast<u64>(t) % p, p);
  std::vector<u64> out;
  out.reserve(static_cast<size_t>(d) * t);
  for (int i = 0; i < t; ++i) {
    for (int j = 0; j < d; ++j) {
      out.push_back(static_cast<u64>((static_cast<u128>(polys[i][j]) * t_inv) % p));
    }
  }
  return out;
}

// encryption.cc:127-149.
std::vector<Poly> SampleAComponents(int d, u64 qmask, int num, SplitMix64& prng) {
  if (num <= 0) InvalidArg("num_a_components must be positive.");
  std::vector<Poly> out(num, Poly(d));
  for (int i = 0; i < num; ++i) {
    for (int j = 0; j < d; ++j) out[i][j] = prng.Next() & qmask;
  }
  return out;
}

// hint_computation.cc:78-160: rows x d hint matrix = DB x (stacked negacyclic
// blocks of the first-dimension a-components), via InnerProductFft in batches
// of safe_n polynomials. The spectra of the a-components are computed once
// (the reference recomputes identical spectra per row; results are equal).
std::vector<u64> ComputeHintMatrix(const std::vector<uint16_t>& db, int rows,
                                   int cols, const std::vector<Poly>& a,
                                   int db_bits) {
  const int n = static_cast<int>(a.size());
  const int d = static_cast<int>(a[0].size());
  if (cols != d * n) {
    InvalidArg(
        "Database columns must match d * n where d is the polynomial length "
        "and n is the number of polynomials.");
  }
  const int chunk_bits = 20;
  const double bits_left = 53.0 - (std::log2(static_cast<double>(d)) + 2.0 * chunk_bits);
  size_t safe_n = (bits_left >= 0.0)
                      ? static_cast<size_t>(std::pow(2.0, bits_left))
                      : 1;
  if (safe_n < 1) safe_n = 1;
  FftPlan plan(2 * d);
  std::vector<Poly> v_inv(n, Poly(d));
  for (int k = 0; k < n; ++k) {
    v_inv[k][0] = a[k][0];
    for (int j = 1; j < d; ++j) v_inv[k][j] = -a[k][d - j];
  }
  // InnerProductFft(u_batch, v_batch, ctx, 8*sizeof(S), 8*sizeof(T), 20):
  IpSetup s = InnerProductSetup(static_cast<int>(std::min<size_t>(safe_n, n)),
                                d, db_bits, 64, chunk_bits, false);
  std::vector<std::vector<CVec>> v_spec(n);
  for (int k = 0; k < n; ++k) v_spec[k] = ChunkSpectra(v_inv[k].data(), d, s.vr, plan);
  std::vector<u64> result(static_cast<size_t>(rows) * d, 0);
  Poly block(d);
  for (int i = 0; i < rows; ++i) {
    const size_t row_off = static_cast<size_t>(i) * cols;
    for (size_t bs = 0; bs < static_cast<size_t>(n); bs += safe_n) {
      const size_t bsz = std::min(safe_n, static_cast<size_t>(n) - bs);
      IpSetup sb = InnerProductSetup(static_cast<int>(bsz), d, db_bits, 64,
                                     chunk_bits, false);
      IpAccum acc(sb, d);
      for (size_t b = 0; b < bsz; ++b) {
        const size_t k = bs + b;
        for (int j = 0; j < d; ++j) block[j] = db[row_off + k * d + j];
        auto us = ChunkSpectra(block.data(), d, sb.ur, plan);
        acc.Mac(us, v_spec[k]);
      }
      Poly res = acc.Finish(plan, nullptr);
      for (int j = 0; j < d; ++j) result[static_cast<size_t>(i) * d + j] += res[j];
    }
  }
  return result;
}

// ---------------------------------------------------------------------------
// Client, crypto/encryption.cc and crypto/pir/client/pir_client.cc.

// encryption.cc:99-125: ternary coefficients {0, 1, q-1}. Uniform {0,1,2} by
// rejection from bytes.
Poly SampleSecretKey(int d, u64 qmask, SplitMix64& prng) {
  Poly s(d);
  for (int i = 0; i < d; ++i) {
    uint8_t r;
    do {
      r = prng.Next8();
    } while (r >= 255);
    const u64 val = r % 3;
    s[i] = (val == 1) ? 1 : (val == 2 ? qmask : 0);
  }
  return s;
}

// Centered binomial error with the given variance: sum of 2*variance
// differences of fair bits (the error distribution of encryption.cc:177-179).
i64 SampleError(int variance, SplitMix64& prng) {
  int pos = 0, neg = 0;
  int left = 2 * variance;
  while (left > 0) {
    const int take = std::min(left, 32);
    const u64 r = prng.Next();
    const u64 mask = (take == 32) ? 0xFFFFFFFFULL : ((1ULL << take) - 1);
    pos += __builtin_popcountll(r & mask);
    neg += __builtin_popcountll((r >> 32) & mask);
    left -= take;
  }
  return static_cast<i64>(pos) - static_cast<i64>(neg);
}

// encryption.cc:151-204: b_i = (e_i - a_i * s) mod q.
std::vector<Poly> GenerateRlweSamples(int d, u64 qmask, int variance,
                                      const Poly& s, const std::vector<Poly>& a,
                                      SplitMix64& prng) {
  std::vector<Poly> bs;
  for (const Poly& ai : a) {
    Poly as = MulTernary(ai, s);
    Poly b(d);
    for (int i = 0; i < d; ++i) {
      const i64 e = SampleError(variance, prng);
      b[i] = (static_cast<u64>(e) - as[i]) & qmask;
    }
    bs.push_back(std::move(b));
  }
  return bs;
}

// encryption.cc:224-260: b += round(m * q / p).
void EncryptInto(Poly& b, int log_q, u64 qmask, u64 p, const std::vector<u64>& m) {
  const u128 q128 = static_cast<u128>(1) << log_q;
  for (size_t i = 0; i < b.size(); ++i) {
    if (m[i] >= p) InvalidArg("Message components must be less than plaintext_modulus.");
    const u64 scaled = static_cast<u64>((static_cast<u128>(m[i]) * q128 + p / 2) / p);
    b[i] = (b[i] + scaled) & qmask;
  }
}

// encryption.cc:395-449: y_i = b_i + (m << (i*log_digit + offset_bits)).
std::vector<Poly> EncryptGadget(int log_q, u64 qmask, int log_digit,
                                int num_digits, const std::vector<Poly>& bs,
                                const Poly& message) {
  const int offset_bits = std::max(0, log_q - num_digits * log_digit);
  std::vector<Poly> out;
  for (int i = 0; i < num_digits; ++i) {
    const int shift = i * log_digit;
    Poly y(bs[i].size());
    for (size_t j = 0; j < y.size(); ++j) {
      u64 scaled = 0;
      if (shift + offset_bits < 64) scaled = message[j] << (shift + offset_bits);
      y[j] = (bs[i][j] + scaled) & qmask;
    }
    out.push_back(std::move(y));
  }
  return out;
}

struct ClientState {
  Poly s;
  SplitMix64 prng;
};

ClientState MakeClient(const Derived& D, uint64_t secret_seed) {
  ClientState c{Poly(), SplitMix64(secret_seed)};
  c.s = SampleSecretKey(D.d, D.qmask, c.prng);
  return c;
}

// ---- Bit-packed formats ----------------------------------------------------
// pir_server.cc:309-336 (BitPackSigned) and :369-399 (BitPackCoefficients),
// polynomial.cc:562-590 (ToProto): bits of value i occupy stream bits
// [i*bits, (i+1)*bits), LSB first within each byte. A signed value is first
// reinterpreted as the unsigned type of its own width (value_bits).
u64 FieldOf(u64 v, int is_signed, int value_bits, int bits) {
  if (is_signed && value_bits < 64) v &= (1ULL << value_bits) - 1;
  return bits >= 64 ? v : (v & ((1ULL << bits) - 1));
}

std::vector<uint8_t> BitPack(const u64* vals, size_t n, int bits, int is_signed,
                             int value_bits) {
  if (n == 0 || bits <= 0) return {};
  // BitPackCoefficients copies whole values when bits >= the type width; the
  // bitwise loop below produces the same little-endian bytes for 64.
  const int b = bits > 64 ? 64 : bits;
  std::vector<uint8_t> out((n * static_cast<size_t>(b) + 7) / 8, 0);
  u64 pos = 0;
  for (size_t i = 0; i < n; ++i) {
    u64 v = FieldOf(vals[i], is_signed, value_bits, b);
    int left = b;
    while (left > 0) {
      const size_t byte = pos / 8;
      const int off = static_cast<int>(pos % 8);
      const int take = std::min(left, 8 - off);
      out[byte] |= static_cast<uint8_t>((v & ((1u << take) - 1)) << off);
      v = take >= 64 ? 0 : (v >> take);
      pos += take;
      left -= take;
    }
  }
  return out;
}

// pir_server.cc:338-367 (BitUnpackSigned), :401-429 (BitUnpackCoefficients).
void BitUnpack(const uint8_t* data, size_t n, int bits, int is_signed, int value_bits,
               u64* out) {
  if (n == 0 || bits <= 0) {
    for (size_t i = 0; i < n; ++i) out[i] = 0;
    return;
  }
  const int b = bits > 64 ? 64 : bits;
  u64 pos = 0;
  for (size_t i = 0; i < n; ++i) {
    u64 v = 0;
    int left = b, shift = 0;
    while (left > 0) {
      const size_t byte = pos / 8;
      const int off = static_cast<int>(pos % 8);
      const int take = std::min(left, 8 - off);
      const u64 chunk = (data[byte] >> off) & ((1u << take) - 1);
      v |= chunk << shift;
      shift += take;
      pos += take;
      left -= take;
    }
    if (is_signed && b < 64 && ((v >> (b - 1)) & 1)) v |= ~((1ULL << b) - 1);
    if (value_bits < 64) v &= (1ULL << value_bits) - 1;  // static_cast to the value type
    if (is_signed && value_bits < 64 && ((v >> (value_bits - 1)) & 1)) {
      v |= ~((1ULL << value_bits) - 1);  // kept sign-extended in 64 bits
    }
    out[i] = v;
  }
}

// Byte writer / bounded reader of the server image (pir_server.cc:433-793).
struct ImageWriter {
  uint8_t* out;  // null: count only
  size_t n = 0;
  void Put(const void* p, size_t len) {
    if (out != nullptr && len > 0) std::memcpy(out + n, p, len);
    n += len;
  }
  template <typename T>
  void Val(T v) { Put(&v, sizeof(v)); }
};

// BufferReader of LoadFromBuffer: a read past the end leaves the offset and
// marks the reader bad; the loader then fails only when the offset is not at
// the end (pir_server.cc:778-780).
struct ImageReader {
  const uint8_t* buf;
  size_t size;
  size_t off = 0;
  bool ok = true;
  void Get(void* p, size_t len) {
    if (off + len > size) {
      ok = false;
      return;
    }
    if (len > 0) std::memcpy(p, buf + off, len);
    off += len;
  }
  template <typename T>
  T Val() {
    T v{};
    Get(&v, sizeof(v));
    return v;
  }
};

int ImageLogModulus(const Derived& D) { return D.log_q >= 64 ? 64 : D.log_q; }

}  // namespace

// ===========================================================================
// extern "C" API.

extern "C" {

void orc_set_matrix_threads(int n) { g_matrix_threads = n < 1 ? 1 : n; }

int orc_get_sizes(const orc_params* P, orc_sizes* out, char* err,
                  size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  out->d = D.d;
  out->t = D.t;
  out->ell = D.ell;
  out->k = D.k;
  out->e = D.e;
  out->r = D.r;
  out->S = D.S;
  out->K = D.K;
  out->M = D.M;
  out->db_words_per_row = (P->db_bytes == 1) ? D.K / 8 : D.K / 4;
  out->pack_cols = D.k * D.d;
  out->pack_words_per_row = (P->mat_bytes == 4) ? D.k * D.d / 2 : 0;
  ORC_GUARD_END(err, errlen)
}

void orc_low_bits(const uint64_t* in, int len, int amount, uint64_t* out) {
  Poly r = LowBits(Poly(in, in + len), amount);
  std::memcpy(out, r.data(), sizeof(u64) * len);
}

void orc_right_shift(const uint64_t* in, int len, int amount, uint64_t* out) {
  Poly r = RightShift(Poly(in, in + len), amount);
  std::memcpy(out, r.data(), sizeof(u64) * len);
}

void orc_rescale(const uint64_t* in, int len, int log_cur, int log_new,
                 uint64_t* out) {
  Poly r = Rescale(Poly(in, in + len), log_cur, log_new);
  std::memcpy(out, r.data(), sizeof(u64) * len);
}

int orc_gadget_inv(const uint64_t* in, int len, int log_modulus, int log_digit,
                   int num_digits, uint64_t* out, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  auto g = GadgetInv(Poly(in, in + len), log_modulus, log_digit, num_digits);
  for (int i = 0; i < num_digits; ++i) {
    std::memcpy(out + static_cast<size_t>(i) * len, g[i].data(), sizeof(u64) * len);
  }
  ORC_GUARD_END(err, errlen)
}

int orc_signed_gadget_inv(const uint64_t* in, int len, int log_modulus,
                          int log_digit, int num_digits, uint64_t* out,
                          char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  auto g = SignedGadgetInv(Poly(in, in + len), log_modulus, log_digit, num_digits);
  for (int i = 0; i < num_digits; ++i) {
    std::memcpy(out + static_cast<size_t>(i) * len, g[i].data(), sizeof(u64) * len);
  }
  ORC_GUARD_END(err, errlen)
}

int orc_automorph(const uint64_t* in, int len, int power, uint64_t* out,
                  char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  Poly r = Automorph(Poly(in, in + len), power);
  std::memcpy(out, r.data(), sizeof(u64) * len);
  ORC_GUARD_END(err, errlen)
}

static std::vector<Poly> SplitPolys(const uint64_t* p, int n, int d) {
  std::vector<Poly> out(n);
  for (int i = 0; i < n; ++i) {
    out[i].assign(p + static_cast<size_t>(i) * d, p + static_cast<size_t>(i + 1) * d);
  }
  return out;
}

int orc_inner_product_fft(const uint64_t* u, const uint64_t* v, int n, int d,
                          int u_bits, int v_bits, int chunk_bits,
                          int u_is_signed, uint64_t* out, double* max_round_err,
                          char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  double me = 0.0;
  Poly r = InnerProductFft(SplitPolys(u, n, d), SplitPolys(v, n, d), u_bits,
                           v_bits, chunk_bits, u_is_signed != 0, &me);
  std::memcpy(out, r.data(), sizeof(u64) * d);
  if (max_round_err != nullptr) *max_round_err = me;
  ORC_GUARD_END(err, errlen)
}

void orc_set_fft_contraction(int fused) { g_fft_fused = fused != 0; }

int orc_inner_product_fft_pairs(const uint64_t* u, const uint64_t* v, int n, int d,
                                int u_bits, int v_bits, int chunk_bits, int u_is_signed,
                                int fused, double* pre, int* pair_a, int* pair_b,
                                int* npairs, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  const auto U = SplitPolys(u, n, d);
  const auto V = SplitPolys(v, n, d);
  *npairs = fused ? InnerProductFftPairs<true>(U, V, u_bits, v_bits, chunk_bits,
                                               u_is_signed != 0, pre, pair_a, pair_b)
                  : InnerProductFftPairs<false>(U, V, u_bits, v_bits, chunk_bits,
                                                u_is_signed != 0, pre, pair_a, pair_b);
  ORC_GUARD_END(err, errlen)
}

int orc_inner_product_exact(const uint64_t* u, const uint64_t* v, int n, int d,
                            int u_bits, int v_bits, int chunk_bits,
                            int u_is_signed, uint64_t* out, char* err,
                            size_t errlen) {
  ORC_GUARD_BEGIN
  Poly r = InnerProductExact(SplitPolys(u, n, d), SplitPolys(v, n, d), u_bits,
                             v_bits, chunk_bits, u_is_signed != 0);
  std::memcpy(out, r.data(), sizeof(u64) * d);
  ORC_GUARD_END(err, errlen)
}


int orc_inner_product_fft_w(const uint64_t* u, const uint64_t* v, int n, int d,
                            int u_bits, int v_bits, int chunk_bits,
                            int u_is_signed, int width, uint64_t* out,
                            double* max_round_err, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  if (width != 32 && width != 64) InvalidArg("width must be 32 or 64.");
  double me = 0.0;
  Poly r = InnerProductFftW(SplitPolys(u, n, d), SplitPolys(v, n, d), u_bits, v_bits,
                            chunk_bits, u_is_signed != 0, width, &me);
  std::memcpy(out, r.data(), sizeof(u64) * d);
  if (max_round_err != nullptr) *max_round_err = me;
  ORC_GUARD_END(err, errlen)
}

int orc_inner_product_exact_w(const uint64_t* u, const uint64_t* v, int n, int d,
                              int u_bits, int v_bits, int chunk_bits,
                              int u_is_signed, int width, uint64_t* out,
                              char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  if (width != 32 && width != 64) InvalidArg("width must be 32 or 64.");
  Poly r = InnerProductExactW(SplitPolys(u, n, d), SplitPolys(v, n, d), u_bits, v_bits,
                              chunk_bits, u_is_signed != 0, width);
  std::memcpy(out, r.data(), sizeof(u64) * d);
  ORC_GUARD_END(err, errlen)
}

// The twiddle table of FftPlan (crypto/fft.h:71-80) as the oracle computes it
// (interleaved re, im).
void orc_fft_twiddles(int n, double* fwd, double* bwd) {
  FftPlan plan(static_cast<size_t>(n));
  for (int i = 0; i < n / 2; ++i) {
    fwd[2 * i] = plan.tw_fwd[i].real();
    fwd[2 * i + 1] = plan.tw_fwd[i].imag();
    bwd[2 * i] = plan.tw_bwd[i].real();
    bwd[2 * i + 1] = plan.tw_bwd[i].imag();
  }
}

// Raw std::mt19937_64 output (standard-specified), the generator of
// crypto/polynomial_test.cc:443,492,552.
void orc_mt19937_64(uint64_t seed, uint64_t* out, size_t n) {
  std::mt19937_64 gen(seed);
  for (size_t i = 0; i < n; ++i) out[i] = gen();
}

int orc_condense_u16(const uint16_t* data, int rows, int cols, uint64_t* out,
                     char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  // matrix.cc:107-123.
  if (cols % 4 != 0) InvalidArg("Columns must be a multiple of 4 to condense uint16_t.");
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols / 4; ++j) {
      u64 w = 0;
      for (int k = 0; k < 4; ++k) {
        w |= static_cast<u64>(data[static_cast<size_t>(i) * cols + j * 4 + k]) << (k * 16);
      }
      out[static_cast<size_t>(i) * (cols / 4) + j] = w;
    }
  }
  ORC_GUARD_END(err, errlen)
}

int orc_condense_u8(const uint8_t* data, int rows, int cols, uint64_t* out,
                    char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  // matrix.cc:90-106.
  if (cols % 8 != 0) InvalidArg("Columns must be a multiple of 8 to condense uint8_t.");
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols / 8; ++j) {
      u64 w = 0;
      for (int k = 0; k < 8; ++k) {
        w |= static_cast<u64>(data[static_cast<size_t>(i) * cols + j * 8 + k]) << (k * 8);
      }
      out[static_cast<size_t>(i) * (cols / 8) + j] = w;
    }
  }
  ORC_GUARD_END(err, errlen)
}

int orc_condense_i32(const int32_t* data, int rows, int cols, uint64_t* out,
                     char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  if (cols % 2 != 0) InvalidArg("Columns must be a multiple of 2 to condense int32_t.");
  std::vector<i64> v(static_cast<size_t>(rows) * cols);
  for (size_t i = 0; i < v.size(); ++i) v[i] = data[i];
  CondenseI32(v.data(), rows, cols, out);
  ORC_GUARD_END(err, errlen)
}

int orc_condensed_multiply_u16(const uint64_t* cdata, int rows, int cols,
                               const uint64_t* vec, int lanes, uint64_t* out,
                               char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  if (cols % (4 * lanes) != 0) {
    InvalidArg("Original columns must be a multiple of 4 * Lanes.");
  }
  CondensedMulU16Rows(cdata, cols, 0, rows, vec, out);
  ORC_GUARD_END(err, errlen)
}

int orc_condensed_multiply_u8(const uint64_t* cdata, int rows, int cols,
                              const uint64_t* vec, int lanes, uint64_t* out,
                              char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  if (cols % (8 * lanes) != 0) {
    InvalidArg("Original columns must be a multiple of 8 * Lanes.");
  }
  CondensedMulU8Rows(cdata, cols, 0, rows, vec, out);
  ORC_GUARD_END(err, errlen)
}

int orc_condensed_multiply_i32(const uint64_t* cdata, int rows, int cols,
                               const uint64_t* vec, int lanes, uint64_t* out,
                               char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  if (cols % (2 * lanes) != 0) {
    InvalidArg("Original columns must be a multiple of 2 * Lanes.");
  }
  CondensedMulI32Rows(cdata, cols, 0, rows, vec, out);
  ORC_GUARD_END(err, errlen)
}

void orc_multiply_i16(const int16_t* data, int rows, int cols,
                      const uint64_t* vec, uint64_t* out) {
  MulI16(data, rows, cols, vec, out);
}

void orc_condensed_multiply_u16_rows(const uint64_t* cdata, int cols, int row0,
                                     int nrows, const uint64_t* vec,
                                     uint64_t* out) {
  CondensedMulU16Rows(cdata, cols, row0, nrows, vec, out);
}

void orc_condensed_multiply_i32_rows(const uint64_t* cdata, int cols, int row0,
                                     int nrows, const uint64_t* vec,
                                     uint64_t* out) {
  CondensedMulI32Rows(cdata, cols, row0, nrows, vec, out);
}

int orc_finalize_matrix_pack(const orc_params* P, const uint64_t* b,
                             const uint64_t* b_agg_partial,
                             const uint64_t* y_vec_h, const uint64_t* t_vec_h,
                             const uint64_t* a_tilde_agg, uint64_t* out_a,
                             uint64_t* out_b, double* max_round_err, char* err,
                             size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  Poly ob;
  double me = 0.0;
  FinalizeMatrixPack(P, D.d, b, b_agg_partial, SplitPolys(y_vec_h, D.k, D.d),
                     SplitPolys(t_vec_h, D.k, D.d), &ob, &me);
  std::memcpy(out_a, a_tilde_agg, sizeof(u64) * D.d);
  std::memcpy(out_b, ob.data(), sizeof(u64) * D.d);
  if (max_round_err != nullptr) *max_round_err = me;
  ORC_GUARD_END(err, errlen)
}

int orc_matrix_pack(const orc_params* P, const uint64_t* b,
                    const uint64_t* y_vec_g, const uint64_t* y_vec_h,
                    const uint64_t* pack_matrix, const uint64_t* t_vec_h,
                    const uint64_t* a_tilde_agg, uint64_t* out_a,
                    uint64_t* out_b, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  const int cols = D.k * D.d;
  const PackView pv = MakePackView(P->mat_bytes, pack_matrix);
  Poly part(D.d);
  pv.Multiply(0, D.d, cols, y_vec_g, part.data());
  Poly ob;
  FinalizeMatrixPack(P, D.d, b, part.data(), SplitPolys(y_vec_h, D.k, D.d),
                     SplitPolys(t_vec_h, D.k, D.d), &ob, nullptr);
  std::memcpy(out_a, a_tilde_agg, sizeof(u64) * D.d);
  std::memcpy(out_b, ob.data(), sizeof(u64) * D.d);
  ORC_GUARD_END(err, errlen)
}

int orc_external_product(const orc_params* P, const uint64_t* ct_a,
                         const uint64_t* ct_b, const uint64_t* rgsw_a,
                         const uint64_t* rgsw_b, uint64_t* out_a,
                         uint64_t* out_b, double* max_round_err, char* err,
                         size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  Poly oa, ob;
  double me = 0.0;
  ExternalProduct(P, D.d, Poly(ct_a, ct_a + D.d), Poly(ct_b, ct_b + D.d),
                  SplitPolys(rgsw_a, 2 * D.e, D.d),
                  SplitPolys(rgsw_b, 2 * D.e, D.d), &oa, &ob, &me);
  std::memcpy(out_a, oa.data(), sizeof(u64) * D.d);
  std::memcpy(out_b, ob.data(), sizeof(u64) * D.d);
  if (max_round_err != nullptr) *max_round_err = me;
  ORC_GUARD_END(err, errlen)
}

// pir_server.cc:52-95 validation messages.
int orc_create_check(const orc_params* P, int num_db_matrices, int db_rows,
                     int db_cols, int num_outputs, int outputs_len,
                     int second_dim_a_len, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  if (num_db_matrices != D.ell) InvalidArg("db_matrices size mismatch.");
  if (db_cols < D.r) InvalidArg("DB matrix must have at least N/t columns.");
  if (db_rows != D.t * D.d) InvalidArg("DB matrix must have td rows.");
  if (num_outputs != D.ell) InvalidArg("preprocessed_outputs size mismatch.");
  if (outputs_len != D.t) InvalidArg("Preprocessed outputs must be of length t.");
  if (D.t > 1 && second_dim_a_len != 2 * D.e) {
    InvalidArg(
        "second_dim_query_a must have length 2 * "
        "PolyEvalGadgetParams().num_digits.");
  }
  ORC_GUARD_END(err, errlen)
}

// pir_server.cc:152-306.
int orc_process_response(const orc_params* P, const uint64_t* db_cdata,
                         int db_cols, const void* pack, const uint64_t* a_tilde,
                         const uint64_t* t_vec_h, const uint64_t* second_dim_a,
                         const uint64_t* first_dim, int fd_len,
                         const uint64_t* second_dim, int sd_polys,
                         const uint64_t* packing_key, int pk_polys,
                         uint64_t* out, double* stage_ms, double* max_round_err,
                         char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  const int d = D.d, t = D.t, ell = D.ell, k = D.k, e = D.e;
  double me = 0.0;
  // pir_server.cc:161-170.
  if (pk_polys != 2 * k) InvalidArg("Packing key must contain 2 * num_digits polynomials.");
  std::vector<Poly> y_vec_g = SplitPolys(packing_key, k, d);
  std::vector<Poly> y_vec_h = SplitPolys(packing_key + static_cast<size_t>(k) * d, k, d);
  // pir_server.cc:173-181.
  if (fd_len < D.r || fd_len > db_cols) {
    InvalidArg("First dimension query size (" + std::to_string(fd_len) +
               ") is invalid for DB columns (" + std::to_string(db_cols) + ").");
  }
  // pir_server.cc:188-205.
  std::vector<Poly> rgsw_a, rgsw_b;
  if (t > 1) {
    if (sd_polys != 2 * e) {
      InvalidArg("second_dimension_query size mismatch. Must be 2 * num_digits.");
    }
    rgsw_a = SplitPolys(second_dim_a, 2 * e, d);
    rgsw_b = SplitPolys(second_dim, 2 * e, d);
  }
  // pir_server.cc:215-229: one DB product with the zero-padded query.
  double t0 = NowMs();
  std::vector<u64> fd(db_cols, 0);
  std::copy(first_dim, first_dim + fd_len, fd.begin());
  std::vector<u64> b_prime(D.M);
  ParallelRows(D.M, [&](int r0, int nr) {
    if (P->db_bytes == 1) {
      CondensedMulU8Rows(db_cdata, db_cols, r0, nr, fd.data(), b_prime.data() + r0);
    } else {
      CondensedMulU16Rows(db_cdata, db_cols, r0, nr, fd.data(), b_prime.data() + r0);
    }
  });
  double t1 = NowMs();
  // pir_server.cc:237-254: one pack product.
  std::vector<u64> y(static_cast<size_t>(k) * d);
  for (int l = 0; l < k; ++l) {
    for (int m = 0; m < d; ++m) y[static_cast<size_t>(l) * d + m] = y_vec_g[l][m];
  }
  const PackView pv = MakePackView(P->mat_bytes, pack);
  std::vector<u64> b_part(D.M);
  ParallelRows(D.M, [&](int r0, int nr) {
    pv.Multiply(r0, nr, k * d, y.data(), b_part.data() + r0);
  });
  // CoeffType uint32_t: both products are Polynomial<uint32_t> sums, i.e.
  // the 64-bit sums mod 2^32 (wraparound arithmetic is a ring map).
  for (auto& x : b_prime) x &= D.wmask;
  for (auto& x : b_part) x &= D.wmask;
  double pack_ms = NowMs() - t1, eval_ms = 0.0;
  // pir_server.cc:256-303.
  for (int s = 0; s < ell; ++s) {
    double tp = NowMs();
    std::vector<Poly> cts_a(t), cts_b(t);
    for (int i = 0; i < t; ++i) {
      const size_t off = static_cast<size_t>(s) * t * d + static_cast<size_t>(i) * d;
      const size_t oi = static_cast<size_t>(s) * t + i;
      Poly ob;
      FinalizeMatrixPack(P, d, b_prime.data() + off, b_part.data() + off, y_vec_h,
                         SplitPolys(t_vec_h + oi * k * d, k, d), &ob, &me);
      cts_a[i].assign(a_tilde + oi * d, a_tilde + (oi + 1) * d);
      for (auto& x : cts_a[i]) x &= D.wmask;
      cts_b[i] = std::move(ob);
    }
    double te = NowMs();
    pack_ms += te - tp;
    // pir_server.cc:126-150 (Horner; the Add is not reduced).
    Poly ra = cts_a.back(), rb = cts_b.back();
    for (int i = t - 2; i >= 0; --i) {
      Poly pa, pb;
      ExternalProduct(P, d, ra, rb, rgsw_a, rgsw_b, &pa, &pb, &me);
      ra = Add(pa, cts_a[i]);
      rb = Add(pb, cts_b[i]);
      for (auto& x : ra) x &= D.wmask;  // uint32_t Add wraps at 2^32
      for (auto& x : rb) x &= D.wmask;
    }
    eval_ms += NowMs() - te;
    // pir_server.cc:288-297.
    // Rescale's result is cast to CoeffType (polynomial.cc:297-301).
    Poly oa = Rescale(ra, D.log_q, D.log_q1);
    Poly ob = Rescale(rb, D.log_q, D.log_q2);
    for (auto& x : oa) x &= D.wmask;
    for (auto& x : ob) x &= D.wmask;
    std::memcpy(out + static_cast<size_t>(s) * 2 * d, oa.data(), sizeof(u64) * d);
    std::memcpy(out + static_cast<size_t>(s) * 2 * d + d, ob.data(), sizeof(u64) * d);
  }
  if (stage_ms != nullptr) {
    stage_ms[0] = t1 - t0;
    stage_ms[1] = pack_ms;
    stage_ms[2] = eval_ms;
  }
  if (max_round_err != nullptr) *max_round_err = me;
  ORC_GUARD_END(err, errlen)
}

int orc_interpolate(const uint64_t* row, int d, int t, uint64_t p,
                    uint64_t* out, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  std::vector<u64> r(row, row + static_cast<size_t>(d) * t);
  std::vector<u64> o = Interpolate(r, d, t, p);
  std::memcpy(out, o.data(), sizeof(u64) * o.size());
  ORC_GUARD_END(err, errlen)
}

int orc_compute_hint_matrix(const uint16_t* db, int rows, int cols, int db_bits,
                            const uint64_t* a, int n, int d, uint64_t* out,
                            char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  if (rows <= 0) InvalidArg("Database matrix cannot be empty.");
  if (n <= 0) InvalidArg("a_components cannot be empty.");
  std::vector<uint16_t> m(db, db + static_cast<size_t>(rows) * cols);
  std::vector<Poly> as = SplitPolys(a, n, d);
  std::vector<u64> h = ComputeHintMatrix(m, rows, cols, as, db_bits);
  std::memcpy(out, h.data(), sizeof(u64) * h.size());
  ORC_GUARD_END(err, errlen)
}

int orc_sample_a_components(const orc_params* P, uint64_t public_seed,
                            uint64_t* first_dim_a, uint64_t* second_dim_a,
                            uint64_t* pack_a, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  SplitMix64 prng(public_seed);
  auto put = [&](const std::vector<Poly>& v, uint64_t* dst) {
    for (size_t j = 0; j < v.size(); ++j) {
      std::memcpy(dst + j * D.d, v[j].data(), sizeof(u64) * D.d);
    }
  };
  put(SampleAComponents(D.d, D.qmask, D.S, prng), first_dim_a);
  if (D.t > 1) {
    std::vector<Poly> sa = SampleAComponents(D.d, D.qmask, 2 * D.e, prng);
    if (second_dim_a != nullptr) put(sa, second_dim_a);
  }
  put(SampleAComponents(D.d, D.qmask, 2 * D.k, prng), pack_a);
  ORC_GUARD_END(err, errlen)
}

int orc_preprocess_matrix_pack(const orc_params* P, const uint64_t* a_is,
                               const uint64_t* w, void* pack,
                               uint64_t* a_tilde, uint64_t* t_vec_h,
                               char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  std::vector<Poly> a = SplitPolys(a_is, D.d, D.d);
  std::vector<Poly> ws = SplitPolys(w, 2 * D.k, D.d);
  std::vector<Poly> wg(ws.begin(), ws.begin() + D.k), wh(ws.begin() + D.k, ws.end());
  PreprocessPackOut pp = PreprocessPack(P, a, wg, wh);
  std::vector<i64> m = BuildPreprocessMatrix(pp, D.d, D.k, D.log_q, P->mat_bytes);
  StorePack(P, m, D.d, D.k * D.d, pack, 0);
  std::memcpy(a_tilde, pp.a_tilde_agg.data(), sizeof(u64) * D.d);
  for (int j = 0; j < D.k; ++j) {
    std::memcpy(t_vec_h + static_cast<size_t>(j) * D.d, pp.t_vec_h[j].data(), sizeof(u64) * D.d);
  }
  ORC_GUARD_END(err, errlen)
}

int orc_pack_reference(const orc_params* P, const uint64_t* a_is,
                       const uint64_t* w, const uint64_t* b,
                       const uint64_t* y_vec_g, const uint64_t* y_vec_h,
                       uint64_t* out_a, uint64_t* out_b, char* err,
                       size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  std::vector<Poly> a = SplitPolys(a_is, D.d, D.d);
  std::vector<Poly> ws = SplitPolys(w, 2 * D.k, D.d);
  std::vector<Poly> wg(ws.begin(), ws.begin() + D.k), wh(ws.begin() + D.k, ws.end());
  PreprocessPackOut pp = PreprocessPack(P, a, wg, wh);
  Poly ob;
  PackReference(P, pp, Poly(b, b + D.d), SplitPolys(y_vec_g, D.k, D.d),
                SplitPolys(y_vec_h, D.k, D.d), &ob);
  std::memcpy(out_a, pp.a_tilde_agg.data(), sizeof(u64) * D.d);
  std::memcpy(out_b, ob.data(), sizeof(u64) * D.d);
  ORC_GUARD_END(err, errlen)
}

// preprocessing_server.h:68-160 (+ InterpolateAndTranspose :176-209 and
// ComputePreprocessedOutputs :212-273).
int orc_preprocess(const orc_params* P, const uint16_t* db_raw,
                   uint64_t public_seed, uint64_t* db_cdata, void* pack,
                   uint64_t* a_tilde, uint64_t* t_vec_h,
                   uint64_t* second_dim_a, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  const int d = D.d, t = D.t, n = D.n, ell = D.ell, r = D.r;
  const u64 p = P->plaintext_modulus;
  SplitMix64 a_prng(public_seed);
  // Shard k of record idx is db_raw[idx*d*ell + k*d ... + d).
  std::vector<std::vector<uint16_t>> db_t(ell);
  for (int s = 0; s < ell; ++s) {
    // InterpolateAndTranspose: r rows of t*d values -> [t*d x r].
    std::vector<uint16_t>& mat = db_t[s];
    mat.assign(static_cast<size_t>(t) * d * r, 0);
    std::vector<u64> row(static_cast<size_t>(t) * d);
    for (int i = 0; i < r; ++i) {
      for (int j = 0; j < t; ++j) {
        const size_t idx = static_cast<size_t>(i) * t + j;  // record index
        for (int c = 0; c < d; ++c) {
          row[static_cast<size_t>(j) * d + c] =
              db_raw[idx * d * ell + static_cast<size_t>(s) * d + c];
        }
      }
      std::vector<u64> interp = Interpolate(row, d, t, p);
      for (size_t x = 0; x < interp.size(); ++x) {
        mat[x * r + i] = static_cast<uint16_t>(interp[x]);
      }
    }
  }
  (void)n;
  std::vector<Poly> first_a = SampleAComponents(d, D.qmask, D.S, a_prng);
  std::vector<Poly> second_a;
  if (t > 1) second_a = SampleAComponents(d, D.qmask, 2 * D.e, a_prng);
  std::vector<Poly> pack_a = SampleAComponents(d, D.qmask, 2 * D.k, a_prng);
  std::vector<Poly> wg(pack_a.begin(), pack_a.begin() + D.k);
  std::vector<Poly> wh(pack_a.begin() + D.k, pack_a.end());
  const int padded = D.S * d;
  const int db_bits = P->db_bytes * 8;
  for (int s = 0; s < ell; ++s) {
    std::vector<uint16_t> hint_db(static_cast<size_t>(t) * d * padded, 0);
    for (int i = 0; i < t * d; ++i) {
      for (int j = 0; j < r; ++j) {
        hint_db[static_cast<size_t>(i) * padded + j] = db_t[s][static_cast<size_t>(i) * r + j];
      }
    }
    std::vector<u64> hint = ComputeHintMatrix(hint_db, t * d, padded, first_a, db_bits);
    for (int i = 0; i < t; ++i) {
      std::vector<Poly> block(d);
      for (int rr = 0; rr < d; ++rr) {
        const size_t off = (static_cast<size_t>(i) * d + rr) * d;
        block[rr].assign(hint.begin() + off, hint.begin() + off + d);
      }
      PreprocessPackOut pp = PreprocessPack(P, block, wg, wh);
      std::vector<i64> m = BuildPreprocessMatrix(pp, d, D.k, D.log_q, P->mat_bytes);
      const size_t oi = static_cast<size_t>(s) * t + i;
      StorePack(P, m, d, D.k * d, pack, oi * d);
      std::memcpy(a_tilde + oi * d, pp.a_tilde_agg.data(), sizeof(u64) * d);
      for (int j = 0; j < D.k; ++j) {
        std::memcpy(t_vec_h + (oi * D.k + j) * d, pp.t_vec_h[j].data(), sizeof(u64) * d);
      }
    }
    // Condense (preprocessing_server.h:127-153): pad columns to K.
    const size_t wpr = (P->db_bytes == 1) ? D.K / 8 : D.K / 4;
    for (int i = 0; i < t * d; ++i) {
      u64* dst = db_cdata + (static_cast<size_t>(s) * t * d + i) * wpr;
      for (size_t w = 0; w < wpr; ++w) {
        u64 word = 0;
        const int per = (P->db_bytes == 1) ? 8 : 4;
        const int bits = (P->db_bytes == 1) ? 8 : 16;
        for (int x = 0; x < per; ++x) {
          const size_t col = w * per + x;
          const u64 val = (col < static_cast<size_t>(r))
                              ? db_t[s][static_cast<size_t>(i) * r + col]
                              : 0;
          word |= val << (x * bits);
        }
        dst[w] = word;
      }
    }
  }
  // CoeffType uint32_t: the material is Polynomial<uint32_t>. Every step
  // above is a ring operation mod 2^64 or reads only bits below log_q <= 32,
  // so the uint32_t material is this material mod 2^32 (a_tilde and
  // second_dim_a are already below q; t_vec_h's signed digits wrap).
  if (D.width < 64) {
    const size_t lt = static_cast<size_t>(D.ell) * D.t;
    for (size_t x = 0; x < lt * d; ++x) a_tilde[x] &= D.wmask;
    for (size_t x = 0; x < lt * D.k * d; ++x) t_vec_h[x] &= D.wmask;
  }
  if (t > 1 && second_dim_a != nullptr) {
    for (int j = 0; j < 2 * D.e; ++j) {
      std::memcpy(second_dim_a + static_cast<size_t>(j) * d, second_a[j].data(), sizeof(u64) * d);
    }
  }
  ORC_GUARD_END(err, errlen)
}

int orc_client_secret_key(const orc_params* P, uint64_t secret_seed,
                          uint64_t* out) {
  try {
    Derived D = Derive(P);
    ClientState c = MakeClient(D, secret_seed);
    std::memcpy(out, c.s.data(), sizeof(u64) * D.d);
  } catch (...) {
    return 3;
  }
  return 0;
}

// pir_client.cc:61-258.
int orc_client_request(const orc_params* P, uint64_t public_seed,
                       uint64_t secret_seed, int index, uint64_t* first_dim,
                       uint64_t* second_dim, uint64_t* packing_key, char* err,
                       size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  const int d = D.d, t = D.t;
  if (index < 0 || index >= D.n) InvalidArg("Index out of bounds.");
  ClientState c = MakeClient(D, secret_seed);
  SplitMix64 a_prng(public_seed);
  // First dimension (pir_client.cc:128-165).
  {
    const int row_index = index / t;
    const int i = row_index / d, j = row_index % d;
    std::vector<Poly> a_first = SampleAComponents(d, D.qmask, D.S, a_prng);
    std::vector<Poly> bs = GenerateRlweSamples(d, D.qmask, P->variance, c.s, a_first, c.prng);
    size_t pos = 0;
    for (int idx = 0; idx < D.S; ++idx) {
      std::vector<u64> msg(d, 0);
      if (idx == i) msg[j] = 1;
      EncryptInto(bs[idx], D.log_q, D.qmask, P->plaintext_modulus, msg);
      const int count = std::min(d, D.r - idx * d);
      std::memcpy(first_dim + pos, bs[idx].data(), sizeof(u64) * count);
      pos += count;
    }
  }
  // Second dimension (pir_client.cc:167-214).
  if (t > 1) {
    const int col_index = index % t;
    std::vector<Poly> a_second = SampleAComponents(d, D.qmask, 2 * D.e, a_prng);
    std::vector<Poly> bs = GenerateRlweSamples(d, D.qmask, P->variance, c.s, a_second, c.prng);
    Poly msg(d, 0);
    const int shift = (2 * d / t) * col_index;
    u64 mval = 1;
    int midx = shift;
    if (shift >= d) {
      midx = shift - d;
      mval = D.qmask;  // q - 1
    }
    msg[midx] = mval;
    // encryption.cc:451-514 monomial fast path: m*s by negacyclic shift.
    Poly ms(d);
    const bool is_neg = (mval == D.qmask);
    for (int jj = 0; jj < d; ++jj) {
      int target = jj + midx;
      const bool wrap = target >= d;
      if (wrap) target -= d;
      const bool negate = wrap ^ is_neg;
      const u64 val = c.s[jj];
      ms[target] = negate ? (static_cast<u64>(-val) & D.qmask) : val;
    }
    std::vector<Poly> ub(bs.begin(), bs.begin() + D.e), vb(bs.begin() + D.e, bs.end());
    std::vector<Poly> u = EncryptGadget(D.log_q, D.qmask, P->eval_log_digit, D.e, ub, ms);
    std::vector<Poly> v = EncryptGadget(D.log_q, D.qmask, P->eval_log_digit, D.e, vb, msg);
    for (int x = 0; x < D.e; ++x) {
      std::memcpy(second_dim + static_cast<size_t>(x) * d, u[x].data(), sizeof(u64) * d);
      std::memcpy(second_dim + static_cast<size_t>(D.e + x) * d, v[x].data(), sizeof(u64) * d);
    }
  }
  // Packing key (pir_client.cc:216-258).
  {
    std::vector<Poly> a_pack = SampleAComponents(d, D.qmask, 2 * D.k, a_prng);
    std::vector<Poly> bs = GenerateRlweSamples(d, D.qmask, P->variance, c.s, a_pack, c.prng);
    Poly sk_g = Automorph(c.s, 5);
    Poly sk_h = Automorph(c.s, 2 * d - 1);
    std::vector<Poly> bg(bs.begin(), bs.begin() + D.k), bh(bs.begin() + D.k, bs.end());
    std::vector<Poly> kg = EncryptGadget(D.log_q, D.qmask, P->pack_log_digit, D.k, bg, sk_g);
    std::vector<Poly> kh = EncryptGadget(D.log_q, D.qmask, P->pack_log_digit, D.k, bh, sk_h);
    for (int x = 0; x < D.k; ++x) {
      std::memcpy(packing_key + static_cast<size_t>(x) * d, kg[x].data(), sizeof(u64) * d);
      std::memcpy(packing_key + static_cast<size_t>(D.k + x) * d, kh[x].data(), sizeof(u64) * d);
    }
  }
  ORC_GUARD_END(err, errlen)
}

// pir_client.cc:260-281 + encryption.cc:330-374.
int orc_client_decrypt(const orc_params* P, uint64_t secret_seed,
                       const uint64_t* response, uint64_t* out_msg, char* err,
                       size_t errlen) {
  ORC_GUARD_BEGIN
  Derived D = Derive(P);
  const int d = D.d;
  ClientState c = MakeClient(D, secret_seed);
  const u64 p = P->plaintext_modulus;
  for (int s = 0; s < D.ell; ++s) {
    Poly a(response + static_cast<size_t>(s) * 2 * d, response + static_cast<size_t>(s) * 2 * d + d);
    Poly b(response + static_cast<size_t>(s) * 2 * d + d, response + static_cast<size_t>(s + 1) * 2 * d);
    Poly as = MulTernary(a, c.s);
    as = LowBits(as, D.log_q1);
    as = Rescale(as, D.log_q1, D.log_q2);
    Poly noisy = LowBits(Add(b, as), D.log_q2);
    const u64 m2mask = QMask(D.log_q2);
    const u128 q128 = static_cast<u128>(1) << D.log_q2;
    for (int i = 0; i < d; ++i) {
      const u64 val = noisy[i] & m2mask;
      out_msg[static_cast<size_t>(s) * d + i] =
          static_cast<u64>(((static_cast<u128>(val) * p + q128 / 2) / q128) % p);
    }
  }
  ORC_GUARD_END(err, errlen)
}

void orc_fill_u64(uint64_t seed, uint64_t* out, size_t n, uint64_t mask) {
  SplitMix64 g(seed);
  for (size_t i = 0; i < n; ++i) out[i] = g.Next() & mask;
}

void orc_fill_mod(uint64_t seed, uint64_t* out, size_t n, uint64_t modulus) {
  SplitMix64 g(seed);
  for (size_t i = 0; i < n; ++i) out[i] = g.Next() % modulus;
}

void orc_fill_db_u16(uint64_t seed, uint16_t* out, size_t n, uint64_t p) {
  SplitMix64 g(seed);
  for (size_t i = 0; i < n; ++i) out[i] = static_cast<uint16_t>(g.Next8() % p);
}


// polynomial.cc:562-590.
int orc_poly_to_proto(const uint64_t* coeffs, int len, int log_modulus,
                      uint8_t* out, size_t out_cap, size_t* out_len, char* err,
                      size_t errlen) {
  ORC_GUARD_BEGIN
  if (log_modulus <= 0 || log_modulus > 64) InvalidArg("Invalid log_modulus.");
  std::vector<uint8_t> b = BitPack(coeffs, static_cast<size_t>(len), log_modulus, 0, 64);
  if (b.size() > out_cap) InvalidArg("output buffer too small.");
  if (!b.empty()) std::memcpy(out, b.data(), b.size());
  *out_len = b.size();
  ORC_GUARD_END(err, errlen)
}

// polynomial.cc:208-252.
int orc_poly_from_proto(const uint8_t* bytes, size_t nbytes, int len,
                        int log_modulus, uint64_t* out, char* err,
                        size_t errlen) {
  ORC_GUARD_BEGIN
  if (len <= 0 || !IsPow2(static_cast<u64>(len))) InvalidArg("Invalid length.");
  if (log_modulus <= 0 || log_modulus > 64) InvalidArg("Invalid log_modulus.");
  if (nbytes == 0) InvalidArg("Empty serialized polynomial.");
  const size_t expected = (static_cast<size_t>(len) * log_modulus + 7) / 8;
  if (nbytes != expected) InvalidArg("Unexpected serialized size.");
  BitUnpack(bytes, static_cast<size_t>(len), log_modulus, 0, 64, out);
  ORC_GUARD_END(err, errlen)
}

size_t orc_bit_pack(const uint64_t* values, size_t n, int bits, int is_signed,
                    int value_bits, uint8_t* out) {
  std::vector<uint8_t> b = BitPack(values, n, bits, is_signed, value_bits);
  if (out != nullptr && !b.empty()) std::memcpy(out, b.data(), b.size());
  return b.size();
}

void orc_bit_unpack(const uint8_t* data, size_t n, int bits, int is_signed,
                    int value_bits, uint64_t* out) {
  BitUnpack(data, n, bits, is_signed, value_bits, out);
}

// pir_server.cc:433-592 on the combined material of a server built by
// PirServer::Create: DB condensed (u16/u8), pack int32 condensed or int16
// entries (not condensed: Matrix::CreateCondensed has no int16 form).
size_t orc_save_server(const orc_params* P, const uint64_t* db_cdata,
                       int db_cols, const void* pack, const uint64_t* a_tilde,
                       const uint64_t* t_vec_h, const uint64_t* second_dim_a,
                       uint8_t* out) {
  const Derived D = Derive(P);
  const int d = D.d, M = D.M, k = D.k, lt = D.ell * D.t;
  ImageWriter w{out};
  // 1. combined DB matrix (condensed words).
  const int per = P->db_bytes == 2 ? 4 : 8;
  const int32_t db_phys = (db_cols + per - 1) / per;
  w.Val<int32_t>(M);
  w.Val<int32_t>(db_phys);
  w.Val<uint8_t>(1);
  w.Val<int32_t>(db_cols);
  const uint64_t db_elems = static_cast<uint64_t>(M) * db_phys;
  w.Val<uint64_t>(db_elems);
  w.Put(db_cdata, db_elems * 8);
  // 2. combined pack matrix, bit-packed entries.
  const int32_t pcols = k * d;
  const bool i32 = P->mat_bytes == 4;
  w.Val<int32_t>(M);
  w.Val<int32_t>(pcols);
  w.Val<int32_t>(i32 ? (pcols + 1) / 2 : pcols);
  w.Val<uint8_t>(i32 ? 1 : 0);
  w.Val<int32_t>(pcols);
  const int32_t bits = P->pack_log_digit + BitWidth(static_cast<u64>(d - 1));
  w.Val<int32_t>(bits);
  // BitPackSigned casts each entry to its unsigned type before masking
  // (pir_server.cc:318-321): 32, 16 or 64 value bits.
  std::vector<u64> entries(static_cast<size_t>(M) * pcols);
  const int value_bits = P->mat_bytes == 8 ? 64 : (i32 ? 32 : 16);
  for (size_t i = 0; i < entries.size(); ++i) {
    if (P->mat_bytes == 8) {
      entries[i] = static_cast<const u64*>(pack)[i];
    } else {
      entries[i] = i32 ? static_cast<u64>(static_cast<i64>(static_cast<const int32_t*>(pack)[i]))
                       : static_cast<u64>(static_cast<i64>(static_cast<const int16_t*>(pack)[i]));
    }
  }
  std::vector<uint8_t> pk = BitPack(entries.data(), entries.size(), bits, 1, value_bits);
  w.Val<uint64_t>(pk.size());
  w.Put(pk.data(), pk.size());
  // 3. preprocessed outputs.
  const int32_t logm = ImageLogModulus(D);
  w.Val<int32_t>(logm);
  w.Val<int32_t>(D.ell);
  for (int sh = 0; sh < D.ell; ++sh) {
    w.Val<int32_t>(D.t);
    for (int i = 0; i < D.t; ++i) {
      const int oi = sh * D.t + i;
      w.Val<uint64_t>(d);
      std::vector<uint8_t> a = BitPack(a_tilde + static_cast<size_t>(oi) * d, d, logm, 0, 64);
      w.Val<uint64_t>(a.size());
      w.Put(a.data(), a.size());
      w.Val<uint64_t>(k);
      for (int j = 0; j < k; ++j) {
        w.Val<uint64_t>(d);
        std::vector<uint8_t> t =
            BitPack(t_vec_h + (static_cast<size_t>(oi) * k + j) * d, d, logm, 0, 64);
        w.Val<uint64_t>(t.size());
        w.Put(t.data(), t.size());
      }
    }
  }
  (void)lt;
  // 4. second-dimension a components (none when t == 1).
  const uint64_t nsda = D.t > 1 ? 2 * D.e : 0;
  w.Val<uint64_t>(nsda);
  for (uint64_t j = 0; j < nsda; ++j) {
    w.Val<uint64_t>(d);
    std::vector<uint8_t> b = BitPack(second_dim_a + j * d, d, logm, 0, 64);
    w.Val<uint64_t>(b.size());
    w.Put(b.data(), b.size());
  }
  return w.n;
}

// pir_server.cc:594-793, checked against P's shapes.
int orc_load_server(const orc_params* P, const uint8_t* buf, size_t size,
                    uint64_t* db_cdata, void* pack, uint64_t* a_tilde,
                    uint64_t* t_vec_h, uint64_t* second_dim_a, char* err,
                    size_t errlen) {
  ORC_GUARD_BEGIN
  const Derived D = Derive(P);
  const int d = D.d, M = D.M, k = D.k;
  ImageReader r{buf, size};
  const int32_t db_rows = r.Val<int32_t>();
  const int32_t db_phys = r.Val<int32_t>();
  const uint8_t db_cond = r.Val<uint8_t>();
  const int32_t db_orig = r.Val<int32_t>();
  const uint64_t db_elems = r.Val<uint64_t>();
  (void)db_cond;
  (void)db_orig;
  if (db_rows != M || db_elems != static_cast<uint64_t>(M) * db_phys) {
    InvalidArg("DB matrix shape does not match the parameters.");
  }
  r.Get(db_cdata, db_elems * 8);
  const int32_t prows = r.Val<int32_t>();
  const int32_t pcols = r.Val<int32_t>();
  (void)r.Val<int32_t>();
  (void)r.Val<uint8_t>();
  (void)r.Val<int32_t>();
  const int32_t bits = r.Val<int32_t>();
  const uint64_t plen = r.Val<uint64_t>();
  if (prows != M || pcols != k * d) InvalidArg("pack matrix shape does not match the parameters.");
  const size_t nent = static_cast<size_t>(prows) * pcols;
  if (plen < (nent * static_cast<size_t>(bits) + 7) / 8 || r.off + plen > size) {
    InvalidArg("pack matrix bytes are short.");
  }
  std::vector<u64> ent(nent);
  BitUnpack(buf + r.off, nent, bits, 1,
            P->mat_bytes == 8 ? 64 : (P->mat_bytes == 4 ? 32 : 16), ent.data());
  r.off += plen;
  for (size_t i = 0; i < nent; ++i) {
    if (P->mat_bytes == 8) {
      static_cast<u64*>(pack)[i] = ent[i];
    } else if (P->mat_bytes == 4) {
      static_cast<int32_t*>(pack)[i] = static_cast<int32_t>(ent[i]);
    } else {
      static_cast<int16_t*>(pack)[i] = static_cast<int16_t>(ent[i]);
    }
  }
  const int32_t logm = r.Val<int32_t>();
  const int32_t nsh = r.Val<int32_t>();
  if (nsh != D.ell) InvalidArg("preprocessed_outputs size mismatch.");
  auto poly = [&](uint64_t* dst) {
    const uint64_t len = r.Val<uint64_t>();
    const uint64_t nb = r.Val<uint64_t>();
    if (len != static_cast<uint64_t>(d) || r.off + nb > size ||
        nb < (len * static_cast<uint64_t>(logm) + 7) / 8) {
      InvalidArg("polynomial record does not match the parameters.");
    }
    BitUnpack(buf + r.off, len, logm, 0, 64, dst);
    r.off += nb;
  };
  for (int sh = 0; sh < nsh; ++sh) {
    const int32_t ns = r.Val<int32_t>();
    if (ns != D.t) InvalidArg("Preprocessed outputs must be of length t.");
    for (int i = 0; i < ns; ++i) {
      const int oi = sh * D.t + i;
      poly(a_tilde + static_cast<size_t>(oi) * d);
      const uint64_t tv = r.Val<uint64_t>();
      if (tv != static_cast<uint64_t>(k)) InvalidArg("t_vec_h must hold k polynomials.");
      for (int j = 0; j < k; ++j) poly(t_vec_h + (static_cast<size_t>(oi) * k + j) * d);
    }
  }
  const uint64_t nsda = r.Val<uint64_t>();
  if (nsda != static_cast<uint64_t>(D.t > 1 ? 2 * D.e : 0)) {
    InvalidArg("second_dim_query_a must have length 2 * PolyEvalGadgetParams().num_digits.");
  }
  for (uint64_t j = 0; j < nsda; ++j) poly(second_dim_a + j * d);
  if (!r.ok && r.off < size) Fail(13, "Error reading from buffer");
  ORC_GUARD_END(err, errlen)
}


// vpir_server.cc:101-195. The DB has ell*d rows and N = db_cols columns; the
// first dimension is N/d flattened polynomials.
int orc_vpir_process(const orc_params* P, int local_finalize,
                     const uint64_t* db_cdata, int db_cols, const void* pack,
                     const uint64_t* a_tilde, const uint64_t* t_vec_h,
                     const uint64_t* first_dim, int fd_polys,
                     const uint64_t* packing_key, int pk_polys, uint64_t* out,
                     char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  (void)a_tilde;
  if (P->interp_degree != 1) InvalidArg("vPIR parameters have interpolation degree 1.");
  const int d = 1 << P->log_degree, ell = P->entry_size_multiple;
  const int k = P->pack_num_digits, rows = ell * d;
  const int expected = local_finalize ? k : 2 * k;
  if (pk_polys != expected) {
    InvalidArg("Packing key must contain " + std::to_string(expected) + " polynomials.");
  }
  if (static_cast<long long>(fd_polys) * d != db_cols) {
    InvalidArg("Flattened first dimension query must match DB matrix columns (N).");
  }
  // Part 1: b' = DB * q.
  std::vector<u64> b_prime(rows);
  if (P->db_bytes == 1) {
    CondensedMulU8Rows(db_cdata, db_cols, 0, rows, first_dim, b_prime.data());
  } else {
    CondensedMulU16Rows(db_cdata, db_cols, 0, rows, first_dim, b_prime.data());
  }
  // Part 2: pack * y_vec_g.
  const PackView pv = MakePackView(P->mat_bytes, pack);
  std::vector<u64> b_part(rows);
  pv.Multiply(0, rows, k * d, packing_key, b_part.data());
  // Part 3: complete packing.
  std::vector<Poly> y_vec_h;
  if (!local_finalize) y_vec_h = SplitPolys(packing_key + static_cast<size_t>(k) * d, k, d);
  double me = 0.0;
  for (int s = 0; s < ell; ++s) {
    const size_t off = static_cast<size_t>(s) * d;
    Poly b(d);
    if (local_finalize) {
      for (int m = 0; m < d; ++m) b[m] = b_prime[off + m] + b_part[off + m];
      b = LowBits(b, P->log_q);
    } else {
      FinalizeMatrixPack(P, d, b_prime.data() + off, b_part.data() + off, y_vec_h,
                         SplitPolys(t_vec_h + off * k, k, d), &b, &me);
    }
    std::copy(b.begin(), b.end(), out + off);
  }
  ORC_GUARD_END(err, errlen)
}

// vpir_server.cc:197-287 and CreateHPrimeMatrix (:320-375): row k*d + r,
// column j*d + c of H' is t_vec_h[k][j][(r - c) mod d], negated when r < c,
// read as a signed log_digit-bit value.
int orc_vpir_process_matmul(const orc_params* P, int local_finalize,
                            const uint64_t* db_cdata, int db_cols,
                            const void* pack, const uint64_t* t_vec_h,
                            const uint64_t* first_dim, int fd_polys,
                            const uint64_t* packing_key, int pk_polys,
                            uint64_t* out, char* err, size_t errlen) {
  ORC_GUARD_BEGIN
  const int d = 1 << P->log_degree, ell = P->entry_size_multiple;
  const int k = P->pack_num_digits, rows = ell * d;
  const int expected = local_finalize ? k : 2 * k;
  if (pk_polys < expected) InvalidArg("Invalid packing key in request.");
  if (static_cast<long long>(fd_polys) * d != db_cols) {
    InvalidArg("Flattened query size must match combined DB matrix columns.");
  }
  std::vector<u64> b_prime(rows);
  if (P->db_bytes == 1) {
    CondensedMulU8Rows(db_cdata, db_cols, 0, rows, first_dim, b_prime.data());
  } else {
    CondensedMulU16Rows(db_cdata, db_cols, 0, rows, first_dim, b_prime.data());
  }
  const PackView pv = MakePackView(P->mat_bytes, pack);
  std::vector<u64> b_part(rows);
  pv.Multiply(0, rows, k * d, packing_key, b_part.data());
  std::vector<u64> h(rows, 0);
  if (!local_finalize) {
    const int shift = 64 - P->pack_log_digit;
    const u64* y_h = packing_key + static_cast<size_t>(k) * d;
    for (int s = 0; s < ell; ++s) {
      for (int r = 0; r < d; ++r) {
        u64 acc = 0;
        for (int j = 0; j < k; ++j) {
          const u64* tv = t_vec_h + (static_cast<size_t>(s) * k + j) * d;
          for (int c = 0; c < d; ++c) {
            const bool neg = r < c;
            const int m = neg ? r + d - c : r - c;
            i64 v = static_cast<i64>(tv[m] << shift) >> shift;
            if (neg) v = -v;
            // H' entries are MatCoeffType; the product accumulates in CoeffType.
            acc += static_cast<u64>(v) * y_h[static_cast<size_t>(j) * d + c];
          }
        }
        h[static_cast<size_t>(s) * d + r] = acc;
      }
    }
  }
  for (int s = 0; s < ell; ++s) {
    Poly b(d);
    for (int m = 0; m < d; ++m) {
      const size_t i = static_cast<size_t>(s) * d + m;
      b[m] = b_prime[i] + b_part[i] + h[i];
    }
    b = LowBits(b, P->log_q);
    std::copy(b.begin(), b.end(), out + static_cast<size_t>(s) * d);
  }
  ORC_GUARD_END(err, errlen)
}

}  // extern "C"
