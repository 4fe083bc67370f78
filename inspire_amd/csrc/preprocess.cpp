// Offline preprocessing on the GPU: PreprocessingServer::Preprocess
// (crypto/pir/server/preprocessing_server.h:68-273) and its three stages,
// Interpolate (crypto/db_row_interpolation.cc:115-228), ComputeHintMatrix
// (crypto/hint_computation.cc:78-160) and PreprocessMatrixPack
// (crypto/lwes_to_rlwe.cc:163-446).
//
// Every FFT product goes through the same bit-exact transform and
// accumulate-invert-round kernels as the online path (fft_kernels.hip), in the
// reference's batching and term order, so the material is bit-identical to
// the reference's. The integer stages are streaming passes
// (preprocess_kernels.hip). The a-components the reference draws from its
// ChaCha PRNG (preprocessing_server.h:98-111) are inputs here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "preprocess_internal.h"
#include "server_internal.h"

namespace pirg {
namespace {

int Log2i(long long x) {
  int l = 0;
  while ((1LL << l) < x) ++l;
  return l;
}

bool Pow2(long long x) { return x > 0 && (x & (x - 1)) == 0; }

// The FFT kernels hold a 2d-point transform in LDS: d <= 4096.
constexpr int kMaxLogDegree = 12;

// absl::bit_width(d - 1) of PreprocessPack (lwes_to_rlwe.cc:244).
int BitWidth(uint32_t x) { return x == 0 ? 0 : 32 - __builtin_clz(x); }

int Brev(int x, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1) << (bits - 1 - b);
  return r;
}

// ComputePower (lwes_to_rlwe.cc:77-83): 5^k mod two_n.
std::vector<int> Powers5(int count, int two_n) {
  std::vector<int> p(std::max(count, 1));
  long long r = 1;
  for (int i = 0; i < count; ++i) {
    p[i] = static_cast<int>(r);
    r = (r * 5) % two_n;
  }
  return p;
}

// Inverse of an odd x modulo 2^m (Newton iteration, exact in 64 bits).
int OddInverse(int x, int two_d) {
  u64 y = static_cast<u64>(x);
  for (int i = 0; i < 6; ++i) y *= 2 - static_cast<u64>(x) * y;
  return static_cast<int>(y & static_cast<u64>(two_d - 1));
}

// ModInverse (db_row_interpolation.cc:85-111).
bool ModInverse(long long a, long long m, uint32_t* out) {
  long long t = 0, nt = 1, r = m, nr = a;
  while (nr != 0) {
    const long long q = r / nr;
    long long tmp = t - q * nt;
    t = nt;
    nt = tmp;
    tmp = r - q * nr;
    r = nr;
    nr = tmp;
  }
  if (r > 1) return false;
  if (t < 0) t += m;
  *out = static_cast<uint32_t>(t);
  return true;
}

// FFT context for one ring degree: stream, twiddles, dims. Given no stream,
// it makes its own (destroyed, after a synchronize, before the twiddles are
// freed).
struct FftCtx {
  int d = 0, log_d = 0;
  FftDims F{};
  hipStream_t s = nullptr;
  DevBuf tw_fwd, tw_bwd;
  CallStream own;
  Error Init(int d_, hipStream_t stream) {
    d = d_;
    log_d = Log2i(d);
    F = FftDims{d, log_d, 2 * d, log_d + 1};
    if (stream == nullptr) {
      PIRG_HIP(own.create());
      stream = own.s;
    }
    s = stream;
    PIRG_HIP(ConfigureFftKernels(F.N));
    std::vector<double2> fwd, bwd;
    Twiddles(F.N, &fwd, &bwd);
    Error er = tw_fwd.alloc(fwd.size() * sizeof(double2));
    if (!er.ok()) return er;
    if (!(er = tw_bwd.alloc(bwd.size() * sizeof(double2))).ok()) return er;
    PIRG_HIP(CopySync(tw_fwd.p, fwd.data(), fwd.size() * sizeof(double2), hipMemcpyHostToDevice,
                      s));
    PIRG_HIP(CopySync(tw_bwd.p, bwd.data(), bwd.size() * sizeof(double2), hipMemcpyHostToDevice,
                      s));
    return Ok();
  }
};

// A device array of pointers, built on the host.
template <typename T>
Error UploadTable(const std::vector<T*>& v, DevBuf* buf, hipStream_t s) {
  Error er = buf->alloc(std::max<size_t>(v.size(), 1) * sizeof(T*));
  if (!er.ok()) return er;
  if (!v.empty()) {
    PIRG_HIP(CopySync(buf->p, v.data(), v.size() * sizeof(T*), hipMemcpyHostToDevice, s));
  }
  return Ok();
}

// Bytes of scratch the hint computation may hold at once.
size_t HintBudget() {
  size_t b = static_cast<size_t>(6) << 30;
  if (const char* env = std::getenv("PIRG_HINT_BUDGET_BYTES")) {
    const long long v = std::atoll(env);
    if (v > 0) b = static_cast<size_t>(v);
  }
  return b;
}

// Interpolate every row group of one shard into `out` ([groups][t][d]).
// `raw` holds records of ell shards (record i of shard s at (i*ell + s)*d).
Error InterpolateRows(const void* raw, int elem_bytes, long long groups, int d, int t, int ell,
                      int shard, uint32_t p, void* scratch, void* out, hipStream_t s) {
  uint32_t t_inv = 0;
  if (!ModInverse(static_cast<long long>(t % p), p, &t_inv)) {
    return Invalid("a is not invertible modulo m (not coprime).");
  }
  const int log_t = Log2i(t);
  if (t <= 16) {
    InterpStage A{};
    A.src = raw;
    A.dst = out;
    A.elem_bytes = elem_bytes;
    A.groups = groups;
    A.d = d;
    A.log_d = Log2i(d);
    A.t = t;
    A.log_t = log_t;
    A.first = 1;
    A.ell = ell;
    A.shard = shard;
    A.p = p;
    A.t_inv = t_inv;
    PIRG_HIP(LaunchInterpFused(A, s));
    return Ok();
  }
  // Larger t: one launch per stage, alternating between `scratch` and `out`
  // so the last lands in `out`.
  void* bufs[2] = {scratch, out};
  int cur = (log_t % 2 == 1) ? 1 : 0;
  const void* src = raw;
  for (int len = 2, st = 0; len <= t; len <<= 1, ++st) {
    InterpStage A{};
    A.src = src;
    A.dst = bufs[cur];
    A.elem_bytes = elem_bytes;
    A.groups = groups;
    A.d = d;
    A.log_d = Log2i(d);
    A.t = t;
    A.log_t = log_t;
    A.len = len;
    A.first = st == 0;
    A.ell = ell;
    A.shard = shard;
    A.p = p;
    A.t_inv = len == t ? t_inv : 0;
    PIRG_HIP(LaunchInterpStage(A, s));
    src = bufs[cur];
    cur ^= 1;
  }
  return Ok();
}

// ComputeHintMatrix (hint_computation.cc:78-160) of `rows` DB rows (row x at
// db + x*stride elements, columns >= valid read as zero) against n
// a-components (host, n x d): out[x] (device, rows x d).
Error HintMatrix(FftCtx& C, const void* db, int elem_bytes, long long stride, long long valid,
                 int rows, const u64* a_host, int n, u64* out) {
  const int d = C.d;
  const int chunk_bits = 20;
  // hint_computation.cc:100-106: polynomials per InnerProductFft call.
  const double bits_left = 53.0 - (std::log2(static_cast<double>(d)) + 2.0 * chunk_bits);
  size_t safe_n = bits_left >= 0.0 ? static_cast<size_t>(std::pow(2.0, bits_left)) : 1;
  if (safe_n < 1) safe_n = 1;
  const int bs = static_cast<int>(std::min<size_t>(safe_n, static_cast<size_t>(n)));
  ChunkRule ur{}, vr{};
  Error er = MakeRules(bs, d, 8 * elem_bytes, 64, chunk_bits, false, &ur, &vr);
  if (!er.ok()) return er;
  const int nbf = n / bs;                 // full batches
  const int tail = n - nbf * bs;          // polynomials of the last, short batch
  const int nb = nbf + (tail > 0 ? 1 : 0);
  if (tail > 0) {
    ChunkRule ut{}, vt{};
    if (!(er = MakeRules(tail, d, 8 * elem_bytes, 64, chunk_bits, false, &ut, &vt)).ok()) {
      return er;
    }
  }
  const PairList pairs = MakePairs(ur, vr);
  const int nu = ur.num_chunks, nv = vr.num_chunks, N = C.F.N;

  // v_k = a_k[0] - sum_j a_k[d - j] X^j (:108-117), spectra of its chunks.
  std::vector<u64> vinv(static_cast<size_t>(n) * d);
  for (int k = 0; k < n; ++k) {
    const u64* a = a_host + static_cast<size_t>(k) * d;
    u64* v = vinv.data() + static_cast<size_t>(k) * d;
    v[0] = a[0];
    for (int j = 1; j < d; ++j) v[j] = 0ULL - a[d - j];
  }
  DevBuf vdev, vspec;
  if (!(er = vdev.alloc(vinv.size() * 8)).ok()) return er;
  PIRG_HIP(CopySync(vdev.p, vinv.data(), vinv.size() * 8, hipMemcpyHostToDevice, C.s));
  if (!(er = vspec.alloc(static_cast<size_t>(n) * nv * N * sizeof(double2))).ok()) return er;
  PIRG_HIP(LaunchSpectra(vdev.as<u64>(), n, 1, d, vr, C.F, C.tw_fwd.as<double2>(),
                         vspec.as<double2>(), C.s));

  // Row blocks sized to the scratch budget.
  const size_t per_row = static_cast<size_t>(n) * nu * N * sizeof(double2) +
                         static_cast<size_t>(nb) * pairs.count * d * 8;
  int R = static_cast<int>(std::min<size_t>(rows, std::max<size_t>(1, HintBudget() / per_row)));
  DevBuf uspec, part;
  if (!(er = uspec.alloc(static_cast<size_t>(R) * n * nu * N * sizeof(double2))).ok()) return er;
  if (!(er = part.alloc(static_cast<size_t>(R) * nb * pairs.count * d * 8)).ok()) return er;
  // IpKernel tables: output (x, batch) reads the batch's u and v spectra and
  // writes its pair partials; rows are the slow index, so a short last block
  // uses a prefix.
  const double2* U = uspec.as<double2>();
  const double2* V = vspec.as<double2>();
  u64* Pp = part.as<u64>();
  std::vector<const double2*> fu, fv, tu, tv;
  std::vector<u64*> fo, to;
  for (int x = 0; x < R; ++x) {
    for (int b = 0; b < nb; ++b) {
      const int k0 = b * bs;
      const double2* up = U + (static_cast<size_t>(x) * n + k0) * nu * N;
      const double2* vp = V + static_cast<size_t>(k0) * nv * N;
      u64* op = Pp + (static_cast<size_t>(x) * nb + b) * pairs.count * d;
      if (b < nbf) {
        fu.push_back(up);
        fv.push_back(vp);
        fo.push_back(op);
      } else {
        tu.push_back(up);
        tv.push_back(vp);
        to.push_back(op);
      }
    }
  }
  DevBuf tfu, tfv, tfo, ttu, ttv, tto;
  if (!(er = UploadTable(fu, &tfu, C.s)).ok() || !(er = UploadTable(fv, &tfv, C.s)).ok() ||
      !(er = UploadTable(fo, &tfo, C.s)).ok() || !(er = UploadTable(tu, &ttu, C.s)).ok() ||
      !(er = UploadTable(tv, &ttv, C.s)).ok() || !(er = UploadTable(to, &tto, C.s)).ok()) {
    return er;
  }
  for (int x0 = 0; x0 < rows; x0 += R) {
    const int nr = std::min(R, rows - x0);
    const void* src = static_cast<const char*>(db) + static_cast<size_t>(x0) * stride * elem_bytes;
    PIRG_HIP(LaunchDbSpectra(src, elem_bytes, stride, valid, nr, n, ur, C.F,
                             C.tw_fwd.as<double2>(), uspec.as<double2>(), C.s));
    if (nbf > 0) {
      PIRG_HIP(LaunchIp(tfu.as<const double2*>(), tfv.as<const double2*>(), tfo.as<u64*>(),
                        nr * nbf, nbf, bs, nu, nv, pairs, C.F, C.tw_bwd.as<double2>(), nullptr,
                        C.s));
    }
    if (tail > 0) {
      PIRG_HIP(LaunchIp(ttu.as<const double2*>(), ttv.as<const double2*>(), tto.as<u64*>(), nr,
                        1, tail, nu, nv, pairs, C.F, C.tw_bwd.as<double2>(), nullptr, C.s));
    }
    PIRG_HIP(LaunchSumParts(part.as<u64>(), nb * pairs.count, d, nr,
                            out + static_cast<size_t>(x0) * d, C.s));
  }
  PIRG_HIP(hipStreamSynchronize(C.s));
  return Ok();
}

// PreprocessMatrixPack (lwes_to_rlwe.cc:421-446) for nb blocks at once:
// block b's LWE a-vectors are rows [b*d, b*d + d) of a_is (device, d words
// each); w (device) holds w_vec_g then w_vec_h (2k x d). Outputs (device):
// pack rows [b*d, b*d + d) of k*d int32 / int16 entries, a_tilde[b], t_vec_h[b].
Error MatrixPackPre(FftCtx& C, const pirg_params& P, const u64* a_is, int nb, const u64* w,
                    void* pack, u64* a_tilde, u64* t_vec_h) {
  const int d = C.d, log_d = C.log_d, N = C.F.N;
  const int k = P.pack_num_digits, ld = P.pack_log_digit, log_q = P.log_q;
  const u64 qmask = log_q >= 64 ? ~0ULL : ((1ULL << log_q) - 1);
  if (ld < 1 || ld > 63) return Invalid("log_digit must be in [1, 63] on the GPU path.");
  ChunkRule ur{}, vr{};
  Error er = MakeRules(k, d, ld, log_q, ChunkFor(k, d), true, &ur, &vr);
  if (!er.ok()) return er;
  const PairList pairs = MakePairs(ur, vr);
  const int nu = ur.num_chunks, nv = vr.num_chunks, np = pairs.count;
  hipStream_t s = C.s;
  const size_t dd = static_cast<size_t>(d) * d;

  // Steps 1-3: a_hat_agg' = EvaluateNttInPlace(X^k p_k) (:247-299).
  DevBuf A0, A1;
  if (!(er = A0.alloc(static_cast<size_t>(nb) * dd * 8)).ok()) return er;
  if (!(er = A1.alloc(static_cast<size_t>(nb) * dd * 8)).ok()) return er;
  PIRG_HIP(LaunchPackNttInit(a_is, nb, d, log_d, BitWidth(static_cast<uint32_t>(d - 1)), qmask,
                             A0.as<u64>(), s));
  u64* a_cur = A0.as<u64>();
  u64* a_nxt = A1.as<u64>();
  for (int len = 2; len <= d; len <<= 1) {
    PIRG_HIP(LaunchRingNttStage(a_cur, a_nxt, 8, nb, log_d, len, s));
    std::swap(a_cur, a_nxt);
  }
  const u64* Ap = a_cur;

  // Step 4: the two reduction chains per block (:301-361).
  const int nch = 2 * nb;
  DevBuf st, ud, vw, part, T0, T1;
  if (!(er = st.alloc(static_cast<size_t>(nch) * d * 8)).ok()) return er;
  if (!(er = ud.alloc(static_cast<size_t>(nch) * k * nu * N * sizeof(double2))).ok()) return er;
  if (!(er = vw.alloc(static_cast<size_t>(2) * k * nv * N * sizeof(double2))).ok()) return er;
  if (!(er = part.alloc(static_cast<size_t>(nch) * np * d * 8)).ok()) return er;
  if (!(er = T0.alloc(static_cast<size_t>(nb) * k * dd * 4)).ok()) return er;
  if (!(er = T1.alloc(static_cast<size_t>(nb) * k * dd * 4)).ok()) return er;
  // BuildPreprocessMatrix's t array starts as zero polynomials (:178); the
  // loop below fills all but the two slots of power +-5^(d/2-1).
  PIRG_HIP(hipMemsetAsync(T0.p, 0, T0.bytes, s));
  std::vector<const double2*> su, sv, fu, fv;
  std::vector<u64*> so, fo;
  const double2* Ud = ud.as<double2>();
  const double2* Vw = vw.as<double2>();
  for (int c = 0; c < nch; ++c) {
    su.push_back(Ud + static_cast<size_t>(c) * k * nu * N);
    sv.push_back(Vw + static_cast<size_t>(c & 1) * k * nv * N);
    so.push_back(part.as<u64>() + static_cast<size_t>(c) * np * d);
  }
  for (int b = 0; b < nb; ++b) {
    fu.push_back(Ud + static_cast<size_t>(b) * k * nu * N);
    fv.push_back(Vw);
    fo.push_back(part.as<u64>() + static_cast<size_t>(b) * np * d);
  }
  DevBuf tsu, tsv, tso, tfu, tfv, tfo;
  if (!(er = UploadTable(su, &tsu, C.s)).ok() || !(er = UploadTable(sv, &tsv, C.s)).ok() ||
      !(er = UploadTable(so, &tso, C.s)).ok() || !(er = UploadTable(fu, &tfu, C.s)).ok() ||
      !(er = UploadTable(fv, &tfv, C.s)).ok() || !(er = UploadTable(fo, &tfo, C.s)).ok()) {
    return er;
  }
  const int two_d = 2 * d;
  const int half = d / 2;
  const std::vector<int> pw = Powers5(half, two_d);
  auto gh_of = [&](int pg) {
    return static_cast<int>((static_cast<long long>(pg) * (two_d - 1)) % two_d);
  };
  // t_g = a_hat_agg[d/2 - 1], t_gh = a_hat_agg[d - 1] (:301-302).
  {
    const int pg = pw[half - 1];
    PIRG_HIP(LaunchChainCombine(Ap, d, (pg - 1) / 2, (gh_of(pg) - 1) / 2, nullptr, 0, nch,
                                st.as<u64>(), s));
  }
  ChainSpectraArgs A{};
  A.state = st.as<u64>();
  A.k = k;
  A.log_digit = ld;
  A.to_shift = log_q - k * ld;
  A.qmask = qmask;
  A.ur = ur;
  A.uspec = ud.as<double2>();
  A.vr = vr;
  A.vspec = vw.as<double2>();
  for (int i = half - 2; i >= 0; --i) {
    const int pg = pw[i];
    const int pgh = gh_of(pg);
    A.n_dig = nch;
    A.dig_first = 0;
    A.dig_step = 1;
    A.digits32 = T0.as<uint32_t>();
    A.slot_g = Brev((pg - 1) / 2, log_d);
    A.slot_gh = Brev((pgh - 1) / 2, log_d);
    A.digits64 = nullptr;
    A.w = w;
    A.n_wtypes = 2;
    A.pinv[0] = OddInverse(pg, two_d);
    A.pinv[1] = OddInverse(pgh, two_d);
    PIRG_HIP(LaunchChainSpectra(A, C.F, C.tw_fwd.as<double2>(), s));
    PIRG_HIP(LaunchIp(tsu.as<const double2*>(), tsv.as<const double2*>(), tso.as<u64*>(), nch, 1,
                      k, nu, nv, pairs, C.F, C.tw_bwd.as<double2>(), nullptr, s));
    PIRG_HIP(LaunchChainCombine(Ap, d, (pg - 1) / 2, (pgh - 1) / 2, part.as<u64>(), np, nch,
                                st.as<u64>(), s));
  }
  // t_vec_h = SignedGadgetInv(t_gh), a_tilde = LowBits(t_g + <t_vec_h, w_vec_h>).
  A.n_dig = nb;
  A.dig_first = 1;
  A.dig_step = 2;
  A.digits32 = nullptr;
  A.digits64 = t_vec_h;
  A.w = w + static_cast<size_t>(k) * d;
  A.n_wtypes = 1;
  A.pinv[0] = A.pinv[1] = 1;
  PIRG_HIP(LaunchChainSpectra(A, C.F, C.tw_fwd.as<double2>(), s));
  PIRG_HIP(LaunchIp(tfu.as<const double2*>(), tfv.as<const double2*>(), tfo.as<u64*>(), nb, 1, k,
                    nu, nv, pairs, C.F, C.tw_bwd.as<double2>(), nullptr, s));
  PIRG_HIP(LaunchChainFinal(st.as<u64>(), part.as<u64>(), np, nb, d, qmask, a_tilde, s));

  // BuildPreprocessMatrix (:163-235) on the low 32 bits: the entries keep
  // only the low 16 / 32 bits, and every step is a ring operation.
  uint32_t* t_cur = T0.as<uint32_t>();
  uint32_t* t_nxt = T1.as<uint32_t>();
  for (int len = 2; len <= d; len <<= 1) {
    PIRG_HIP(LaunchRingNttStage(t_cur, t_nxt, 4, static_cast<long long>(nb) * k, log_d, len, s));
    std::swap(t_cur, t_nxt);
  }
  PIRG_HIP(LaunchMatrixOut(t_cur, nb, k, d, P.mat_bytes, pack, s));
  PIRG_HIP(hipStreamSynchronize(s));
  return Ok();
}

// Device-resident result of Preprocess, in the combined layouts PirServer
// keeps (db: M x K elements, pack: M rows of k*d entries, a_tilde: lt x d,
// t_vec_h: lt x k x d).
struct Material {
  DevBuf db, pack, a_tilde, t_vec_h;
  int K = 0;
};

struct StageTimes {
  double interpolate_ms = 0, hints_ms = 0, pack_ms = 0, total_ms = 0;
};

Error CheckParams(const pirg_params& P, pirg_server* S) {
  Error er = DeriveServer(P, S);
  if (!er.ok()) return er;
  if (S->log_d > kMaxLogDegree) return Invalid("ring degree too large for the GPU path.");
  if (P.mat_bytes == 8) {
    return Invalid("GPU preprocessing writes int32_t or int16_t pack matrices (mat_bytes 4 or 2).");
  }
  return Ok();
}

// The plaintext modulus as InterpolateAndTranspose passes it to Interpolate
// (a DbValueType, preprocessing_server.h:196).
uint32_t DbModulus(const pirg_params& P) {
  return P.db_bytes == 2 ? static_cast<uint16_t>(P.plaintext_modulus)
                         : static_cast<uint8_t>(P.plaintext_modulus);
}

Error RunPreprocess(const pirg_params& P, const void* db_raw, const u64* first_a,
                    const u64* pack_a, Material* M, StageTimes* times) {
  pirg_server S;
  Error er = CheckParams(P, &S);
  if (!er.ok()) return er;
  const int d = S.d, t = S.t, ell = S.ell, k = S.k, r = S.r;
  const int eb = P.db_bytes;
  const int nfirst = (r + d - 1) / d;  // NumFirstDimSamples: PaddedCols / d
  const int simd = eb == 1 ? 64 : 32;
  const int K = (r + simd - 1) / simd * simd;
  const uint32_t p = DbModulus(P);
  if (t > 1 && (p % 2 == 0)) return Invalid("plaintext_modulus must be odd.");
  if (!(er = CheckDevice()).ok()) return er;
  M->K = K;
  hipStream_t s = nullptr;
  PIRG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  // ev[0] start, ev[1] interpolation done, ev[2] / ev[3] around each shard's
  // hint stage, ev[4] after its pack stage.
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  struct EventGuard {
    hipEvent_t* e;
    ~EventGuard() {
      for (int i = 0; i < 5; ++i) {
        if (e[i] != nullptr) (void)hipEventDestroy(e[i]);
      }
    }
  } eguard{ev};
  for (auto& e : ev) PIRG_HIP(hipEventCreate(&e));
  FftCtx C;
  if (!(er = C.Init(d, s)).ok()) return er;

  const size_t raw_elems = static_cast<size_t>(P.num_entries) * d * ell;
  const size_t shard_rows = static_cast<size_t>(t) * d;
  const size_t Mrows = static_cast<size_t>(ell) * shard_rows;
  if (!(er = M->db.alloc(Mrows * K * eb)).ok()) return er;
  if (!(er = M->pack.alloc(Mrows * k * d * P.mat_bytes)).ok()) return er;
  if (!(er = M->a_tilde.alloc(static_cast<size_t>(ell) * t * d * 8)).ok()) return er;
  if (!(er = M->t_vec_h.alloc(static_cast<size_t>(ell) * t * k * d * 8)).ok()) return er;

  {
    // Stage 1: InterpolateAndTranspose per shard (:176-209), written in the
    // condensed layout (columns padded to K with zeros, :127-153). The timers
    // start once the raw DB is in device memory.
    DevBuf raw, i0, i1;
    if (!(er = raw.alloc(raw_elems * eb)).ok()) return er;
    PIRG_HIP(hipMemcpyAsync(raw.p, db_raw, raw_elems * eb, hipMemcpyHostToDevice, s));
    PIRG_HIP(hipEventRecord(ev[0], s));
    if (t > 1) {
      if (!(er = i0.alloc(static_cast<size_t>(r) * shard_rows * eb)).ok()) return er;
      if (!(er = i1.alloc(static_cast<size_t>(r) * shard_rows * eb)).ok()) return er;
    }
    for (int sh = 0; sh < ell; ++sh) {
      char* dst = M->db.as<char>() + static_cast<size_t>(sh) * shard_rows * K * eb;
      if (t > 1) {
        er = InterpolateRows(raw.p, eb, r, d, t, ell, sh, p, i0.p, i1.p, s);
        if (!er.ok()) return er;
        PIRG_HIP(LaunchTranspose(i1.p, static_cast<long long>(shard_rows), r,
                                 static_cast<int>(shard_rows), eb, dst, K, s));
      } else {
        PIRG_HIP(LaunchTranspose(raw.as<char>() + static_cast<size_t>(sh) * d * eb,
                                 static_cast<long long>(ell) * d, r, d, eb, dst, K, s));
      }
    }
    PIRG_HIP(hipEventRecord(ev[1], s));
    PIRG_HIP(hipStreamSynchronize(s));
  }
  // Stages 2 and 3 per shard: the hint matrix, then PreprocessMatrixPack on
  // each of its t blocks of d rows (:212-273).
  DevBuf hint, wdev;
  if (!(er = hint.alloc(shard_rows * d * 8)).ok()) return er;
  if (!(er = wdev.alloc(static_cast<size_t>(2) * k * d * 8)).ok()) return er;
  PIRG_HIP(CopySync(wdev.p, pack_a, static_cast<size_t>(2) * k * d * 8, hipMemcpyHostToDevice, s));
  float interp_ms = 0, hint_ms = 0, pack_ms = 0, total_ms = 0;
  PIRG_HIP(hipEventElapsedTime(&interp_ms, ev[0], ev[1]));
  for (int sh = 0; sh < ell; ++sh) {
    const char* dbs = M->db.as<char>() + static_cast<size_t>(sh) * shard_rows * K * eb;
    PIRG_HIP(hipEventRecord(ev[2], s));
    er = HintMatrix(C, dbs, eb, K, r, static_cast<int>(shard_rows), first_a, nfirst,
                    hint.as<u64>());
    if (!er.ok()) return er;
    PIRG_HIP(hipEventRecord(ev[3], s));
    char* pk = M->pack.as<char>() + static_cast<size_t>(sh) * shard_rows * k * d * P.mat_bytes;
    er = MatrixPackPre(C, P, hint.as<u64>(), t, wdev.as<u64>(), pk,
                       M->a_tilde.as<u64>() + static_cast<size_t>(sh) * t * d,
                       M->t_vec_h.as<u64>() + static_cast<size_t>(sh) * t * k * d);
    if (!er.ok()) return er;
    PIRG_HIP(hipEventRecord(ev[4], s));
    PIRG_HIP(hipEventSynchronize(ev[4]));
    float a = 0, b = 0;
    PIRG_HIP(hipEventElapsedTime(&a, ev[2], ev[3]));
    PIRG_HIP(hipEventElapsedTime(&b, ev[3], ev[4]));
    hint_ms += a;
    pack_ms += b;
  }
  PIRG_HIP(hipEventElapsedTime(&total_ms, ev[0], ev[4]));
  if (times != nullptr) {
    times->interpolate_ms = interp_ms;
    times->hints_ms = hint_ms;
    times->pack_ms = pack_ms;
    times->total_ms = total_ms;
  }
  return Ok();
}

Error DownloadMaterial(const pirg_params& P, const Material& M, int ell, int t, int d, int k,
                       uint64_t* db_cdata, void* pack, uint64_t* a_tilde, uint64_t* t_vec_h) {
  const size_t rows = static_cast<size_t>(ell) * t * d;
  CallStream cs;
  PIRG_HIP(cs.create());
  if (db_cdata != nullptr) {
    PIRG_HIP(CopySync(db_cdata, M.db.p, rows * M.K * P.db_bytes, hipMemcpyDeviceToHost, cs.s));
  }
  if (pack != nullptr) {
    PIRG_HIP(CopySync(pack, M.pack.p, rows * k * d * P.mat_bytes, hipMemcpyDeviceToHost, cs.s));
  }
  if (a_tilde != nullptr) {
    PIRG_HIP(CopySync(a_tilde, M.a_tilde.p, static_cast<size_t>(ell) * t * d * 8,
                      hipMemcpyDeviceToHost, cs.s));
  }
  if (t_vec_h != nullptr) {
    PIRG_HIP(CopySync(t_vec_h, M.t_vec_h.p, static_cast<size_t>(ell) * t * k * d * 8,
                      hipMemcpyDeviceToHost, cs.s));
    // CoeffType uint32_t: t_vec_h's signed digits as Polynomial<uint32_t>
    // holds them (the other outputs are already below q <= 2^32).
    if (P.coeff_bits == 32) {
      const size_t n = static_cast<size_t>(ell) * t * k * d;
      for (size_t x = 0; x < n; ++x) t_vec_h[x] &= 0xFFFFFFFFULL;
    }
  }
  return Ok();
}

void WriteTimes(const StageTimes& T, double* stage_ms) {
  if (stage_ms == nullptr) return;
  stage_ms[0] = T.interpolate_ms;
  stage_ms[1] = T.hints_ms;
  stage_ms[2] = T.pack_ms;
  stage_ms[3] = T.total_ms;
}

}  // namespace
}  // namespace pirg

using namespace pirg;

extern "C" {

int pirg_interpolate(const void* rows, int db_bytes, long long nrows, int d, int t,
                     uint32_t plaintext_modulus, void* out, char* err, size_t errlen) {
  try {
    if (db_bytes != 1 && db_bytes != 2) return Report(Invalid("db_bytes must be 1 or 2."), err, errlen);
    if (!Pow2(d)) return Report(Invalid("ring_degree must be a power of two."), err, errlen);
    if (!Pow2(t)) {
      return Report(Invalid("interpolation_degree must be a power of two."), err, errlen);
    }
    if (t > 2 * d) {
      return Report(
          Invalid("interpolation_degree must be less than or equal to 2 * ring_degree."), err,
          errlen);
    }
    if (nrows < 0) return Report(Invalid("nrows must be non-negative."), err, errlen);
    if (nrows > 0 && (rows == nullptr || out == nullptr)) {
      return Report(Invalid("rows and out must not be null."), err, errlen);
    }
    const size_t bytes = static_cast<size_t>(nrows) * t * d * db_bytes;
    if (t == 1 || nrows == 0) {
      if (bytes > 0) std::memcpy(out, rows, bytes);
      return 0;
    }
    const uint32_t p = plaintext_modulus;
    if (p % 2 == 0) return Report(Invalid("plaintext_modulus must be odd."), err, errlen);
    if (p > (db_bytes == 2 ? 0xFFFFu : 0xFFu)) {
      return Report(Invalid("plaintext_modulus does not fit the DB value type."), err, errlen);
    }
    Error er = CheckDevice();
    if (!er.ok()) return Report(er, err, errlen);
    DevBuf src, a, b;
    if (!(er = src.alloc(bytes)).ok() || !(er = a.alloc(bytes)).ok() ||
        !(er = b.alloc(bytes)).ok()) {
      return Report(er, err, errlen);
    }
    CallStream cs;
    hipError_t e = cs.create();
    if (e != hipSuccess) return Report(HipErr(e, "hipStreamCreateWithFlags"), err, errlen);
    e = CopySync(src.p, rows, bytes, hipMemcpyHostToDevice, cs.s);
    if (e != hipSuccess) return Report(HipErr(e, "hipMemcpyAsync"), err, errlen);
    er = InterpolateRows(src.p, db_bytes, nrows, d, t, 1, 0, p, a.p, b.p, cs.s);
    if (!er.ok()) return Report(er, err, errlen);
    e = CopySync(out, b.p, bytes, hipMemcpyDeviceToHost, cs.s);
    if (e != hipSuccess) return Report(HipErr(e, "hipMemcpy"), err, errlen);
    return 0;
  } catch (const std::exception& ex) {
    return Report(Internal(ex.what()), err, errlen);
  }
}

int pirg_compute_hint_matrix(const void* db, int db_bytes, int rows, int cols,
                             const uint64_t* a_components, int n, int d, uint64_t* out,
                             char* err, size_t errlen) {
  try {
    if (rows <= 0) return Report(Invalid("Database matrix cannot be empty."), err, errlen);
    if (n <= 0) return Report(Invalid("a_components cannot be empty."), err, errlen);
    if (static_cast<long long>(cols) != static_cast<long long>(d) * n) {
      return Report(Invalid("Database columns must match d * n where d is the polynomial length "
                            "and n is the number of polynomials."),
                    err, errlen);
    }
    if (db_bytes != 1 && db_bytes != 2) return Report(Invalid("db_bytes must be 1 or 2."), err, errlen);
    if (db == nullptr || a_components == nullptr || out == nullptr) {
      return Report(Invalid("db, a_components and out must not be null."), err, errlen);
    }
    if (!Pow2(d) || d < 2 || Log2i(d) > kMaxLogDegree) {
      return Report(Invalid("d must be a power of two in [2, 4096] on the GPU path."), err,
                    errlen);
    }
    Error er = CheckDevice();
    if (!er.ok()) return Report(er, err, errlen);
    FftCtx C;
    if (!(er = C.Init(d, nullptr)).ok()) return Report(er, err, errlen);
    DevBuf dbd, o;
    const size_t dbb = static_cast<size_t>(rows) * cols * db_bytes;
    if (!(er = dbd.alloc(dbb)).ok() || !(er = o.alloc(static_cast<size_t>(rows) * d * 8)).ok()) {
      return Report(er, err, errlen);
    }
    hipError_t e = CopySync(dbd.p, db, dbb, hipMemcpyHostToDevice, C.s);
    if (e != hipSuccess) return Report(HipErr(e, "hipMemcpyAsync"), err, errlen);
    er = HintMatrix(C, dbd.p, db_bytes, cols, cols, rows, a_components, n, o.as<u64>());
    if (!er.ok()) return Report(er, err, errlen);
    e = CopySync(out, o.p, static_cast<size_t>(rows) * d * 8, hipMemcpyDeviceToHost, C.s);
    if (e != hipSuccess) return Report(HipErr(e, "hipMemcpy"), err, errlen);
    return 0;
  } catch (const std::exception& ex) {
    return Report(Internal(ex.what()), err, errlen);
  }
}

int pirg_preprocess_matrix_pack(const pirg_params* params, int nblocks, const uint64_t* a_is,
                                const uint64_t* w, void* pack, uint64_t* a_tilde,
                                uint64_t* t_vec_h, char* err, size_t errlen) {
  try {
    if (params == nullptr || a_is == nullptr || w == nullptr || pack == nullptr ||
        a_tilde == nullptr || t_vec_h == nullptr) {
      return Report(Invalid("null argument."), err, errlen);
    }
    pirg_server S;
    Error er = CheckParams(*params, &S);
    if (!er.ok()) return Report(er, err, errlen);
    if (nblocks <= 0) return Report(Invalid("nblocks must be positive."), err, errlen);
    if (!(er = CheckDevice()).ok()) return Report(er, err, errlen);
    const int d = S.d, k = S.k;
    FftCtx C;
    if (!(er = C.Init(d, nullptr)).ok()) return Report(er, err, errlen);
    const size_t ab = static_cast<size_t>(nblocks) * d * d * 8;
    const size_t pb = static_cast<size_t>(nblocks) * d * k * d * params->mat_bytes;
    DevBuf ad, wd, pd, atd, thd;
    if (!(er = ad.alloc(ab)).ok() || !(er = wd.alloc(static_cast<size_t>(2) * k * d * 8)).ok() ||
        !(er = pd.alloc(pb)).ok() ||
        !(er = atd.alloc(static_cast<size_t>(nblocks) * d * 8)).ok() ||
        !(er = thd.alloc(static_cast<size_t>(nblocks) * k * d * 8)).ok()) {
      return Report(er, err, errlen);
    }
    hipError_t e = CopySync(ad.p, a_is, ab, hipMemcpyHostToDevice, C.s);
    if (e == hipSuccess) {
      e = CopySync(wd.p, w, static_cast<size_t>(2) * k * d * 8, hipMemcpyHostToDevice, C.s);
    }
    if (e != hipSuccess) return Report(HipErr(e, "hipMemcpyAsync"), err, errlen);
    er = MatrixPackPre(C, *params, ad.as<u64>(), nblocks, wd.as<u64>(), pd.p, atd.as<u64>(),
                       thd.as<u64>());
    if (!er.ok()) return Report(er, err, errlen);
    e = CopySync(pack, pd.p, pb, hipMemcpyDeviceToHost, C.s);
    if (e == hipSuccess) {
      e = CopySync(a_tilde, atd.p, static_cast<size_t>(nblocks) * d * 8, hipMemcpyDeviceToHost,
                   C.s);
    }
    if (e == hipSuccess) {
      e = CopySync(t_vec_h, thd.p, static_cast<size_t>(nblocks) * k * d * 8,
                   hipMemcpyDeviceToHost, C.s);
    }
    if (e != hipSuccess) return Report(HipErr(e, "hipMemcpyAsync"), err, errlen);
    return 0;
  } catch (const std::exception& ex) {
    return Report(Internal(ex.what()), err, errlen);
  }
}

int pirg_preprocess(const pirg_params* params, const void* db_raw, const uint64_t* first_dim_a,
                    const uint64_t* pack_a, uint64_t* db_cdata, int* db_cols, void* pack,
                    uint64_t* a_tilde, uint64_t* t_vec_h, double* stage_ms, char* err,
                    size_t errlen) {
  try {
    if (params == nullptr || db_raw == nullptr || first_dim_a == nullptr || pack_a == nullptr) {
      return Report(Invalid("params, db_raw, first_dim_a and pack_a must not be null."), err,
                    errlen);
    }
    Material M;
    StageTimes T;
    Error er = RunPreprocess(*params, db_raw, first_dim_a, pack_a, &M, &T);
    if (!er.ok()) return Report(er, err, errlen);
    const int d = 1 << params->log_degree;
    er = DownloadMaterial(*params, M, params->entry_size_multiple, params->interp_degree, d,
                          params->pack_num_digits, db_cdata, pack, a_tilde, t_vec_h);
    if (!er.ok()) return Report(er, err, errlen);
    if (db_cols != nullptr) *db_cols = M.K;
    WriteTimes(T, stage_ms);
    return 0;
  } catch (const std::exception& ex) {
    return Report(Internal(ex.what()), err, errlen);
  }
}

int pirg_server_create_preprocessed(const pirg_params* params, const void* db_raw,
                                    const uint64_t* first_dim_a, const uint64_t* second_dim_a,
                                    const uint64_t* pack_a, double* stage_ms, pirg_server** out,
                                    char* err, size_t errlen) {
  try {
    if (params == nullptr || db_raw == nullptr || first_dim_a == nullptr || pack_a == nullptr ||
        out == nullptr) {
      return Report(Invalid("params, db_raw, first_dim_a, pack_a and out must not be null."),
                    err, errlen);
    }
    if (params->interp_degree > 1 && second_dim_a == nullptr) {
      return Report(Invalid("second_dim_a must not be null when interpolation_degree > 1."), err,
                    errlen);
    }
    Material M;
    StageTimes T;
    Error er = RunPreprocess(*params, db_raw, first_dim_a, pack_a, &M, &T);
    if (!er.ok()) return Report(er, err, errlen);
    std::unique_ptr<pirg_server> S;
    er = CreateFromDevice(params, M.db.as<uint64_t>(), M.K, M.pack.p, M.a_tilde.as<uint64_t>(),
                          M.t_vec_h.as<uint64_t>(), second_dim_a, &S);
    if (!er.ok()) return Report(er, err, errlen);
    WriteTimes(T, stage_ms);
    *out = S.release();
    return 0;
  } catch (const std::exception& ex) {
    return Report(Internal(ex.what()), err, errlen);
  }
}

}  // extern "C"
