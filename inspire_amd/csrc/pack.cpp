// Batched LWE-to-RLWE packing: MatrixPack (crypto/lwes_to_rlwe.cc:481-508)
// for many independent jobs whose PreprocessMatrixPackOutput material
// (crypto/lwes_to_rlwe.h:42-47) stays resident in HBM.
//
// Job j's pack matrix P_j (d x k*d) is row block j of one tall int8-limb
// matrix; the product b_part_j = P_j * y_vec_g_j is a grouped narrow scan
// (each d-row block against its own key), and FinalizeMatrixPack
// (lwes_to_rlwe.cc:448-479) runs for every job as in ProcessResponse with
// t = 1 and no modulus switch. A job's LWE count m_j <= d only decides how
// much of its b vector is non-zero (the reference packs exactly d LWEs,
// lwes_to_rlwe.cc:428, so shorter inputs are zero-padded); the whole batch
// is one hipGraph launch.
#include <cstring>

#include "server_internal.h"

using namespace pirg;

struct pirg_packer {
  pirg_params P{};
  int device = 0;
  int d = 0, log_d = 0, N = 0, k = 0, J = 0, log_q = 0;
  FftDims F{};
  ChunkRule fin_u{}, fin_v{};
  PairList fin_pairs{};
  LimbLayout L{};
  ScanPlan plan{};
  DevBuf tiles, a_tilde, t_spec, tw_fwd, tw_bwd, margin;
  DevBuf b_dev, keys_dev, bt, corr, zero, partial, yh_spec, fin_partial, ct_b, st, out,
      fin_tab;
  HostBuf h_b, h_keys, h_out;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraphExec_t graph = nullptr;
  bool use_graph = true;
  double total_ms = 0, scan_ms = 0;
  long long runs = 0;

  ~pirg_packer() {
    if (stream != nullptr) (void)hipStreamSynchronize(stream);
    if (graph != nullptr) (void)hipGraphExecDestroy(graph);
    for (auto& e : ev) {
      if (e != nullptr) (void)hipEventDestroy(e);
    }
    if (stream != nullptr) (void)hipStreamDestroy(stream);
  }
};

namespace pirg {
namespace {

Error CheckPackParams(const pirg_params& P, int njobs) {
  if (njobs <= 0) return Invalid("njobs must be positive.");
  if (P.log_degree < 6 || P.log_degree > 12) {
    return Invalid("log_degree must be in [6, 12] for batched packing.");
  }
  if (P.log_q < 1 || P.log_q > 64) return Invalid("log_q must be in [1, 64].");
  if (P.coeff_bits != 0 && P.coeff_bits != 32 && P.coeff_bits != 64) {
    return Invalid("coeff_bits must be 32 or 64.");
  }
  if (P.coeff_bits == 32 && P.log_q > 32) {
    return Invalid("log_q exceeds the coefficient width (32 bits).");
  }
  if (P.pack_log_digit <= 0 || P.pack_num_digits <= 0) {
    return Invalid("pack gadget parameters must be positive.");
  }
  if (P.mat_bytes != 4 && P.mat_bytes != 2) return Invalid("mat_bytes must be 4 or 2.");
  return Ok();
}

// The batch on the packer's stream; `timed` adds the scan-kernel events.
Error Enqueue(pirg_packer* K, bool timed) {
  hipStream_t st = K->stream;
  const int d = K->d, k = K->k, J = K->J;
  // b_part = P_j * concat(y_vec_g_j) for every job (lwes_to_rlwe.cc:490-500).
  PIRG_HIP(LaunchQueryDigits(K->keys_dev.as<u64>(), 2 * k * d, k * d, K->L, K->plan,
                             K->bt.as<int8_t>(), st));
  PIRG_HIP(LaunchVectorSums(K->keys_dev.as<u64>(), J, 2 * k * d, k * d, K->L.c_off,
                            K->corr.as<u64>() + J, K->corr.as<u64>(), st));
  if (timed) PIRG_HIP(hipEventRecord(K->ev[1], st));
  PIRG_HIP(LaunchScan(K->tiles.as<int8_t>(), K->L, K->plan, K->bt.as<int8_t>(),
                      K->partial.as<u64>(), st));
  if (timed) PIRG_HIP(hipEventRecord(K->ev[2], st));
  // FinalizeMatrixPack's inner product <t_vec_h, y_vec_h> (lwes_to_rlwe.cc:472-474).
  PIRG_HIP(LaunchSpectra(K->keys_dev.as<u64>() + static_cast<size_t>(k) * d, J * k, k,
                         2LL * k * d, K->fin_v, K->F, K->tw_fwd.as<double2>(),
                         K->yh_spec.as<double2>(), st));
  void* const* tab = K->fin_tab.as<void* const>();
  PIRG_HIP(LaunchIp(reinterpret_cast<const double2* const*>(tab),
                    reinterpret_cast<const double2* const*>(tab + J),
                    reinterpret_cast<u64* const*>(tab + 2 * J), J, 1, k, K->fin_u.num_chunks,
                    K->fin_v.num_chunks, K->fin_pairs, K->F, K->tw_bwd.as<double2>(),
                    K->margin.as<unsigned long long>(), st));
  // b_agg = LowBits(LowBits(b + b_part) + inner); (a_tilde_agg, b_agg).
  FinalizeArgs FA{};
  FA.db_partial = K->b_dev.as<u64>();
  FA.ksplit_db = 1;
  FA.corr_db = K->zero.as<u64>();
  FA.pk_partial = K->partial.as<u64>();
  FA.ksplit_pk = K->plan.ksplit;
  FA.corr_pk = K->corr.as<u64>();
  FA.corr_rows = d;
  FA.fin_partial = K->fin_partial.as<u64>();
  FA.npairs = K->fin_pairs.count;
  FA.a_tilde = K->a_tilde.as<u64>();
  FA.ct_b = K->ct_b.as<u64>();
  FA.st_a = K->st.as<u64>();
  FA.st_b = K->st.as<u64>();
  FA.out = K->out.as<u64>();
  FA.nq = 1;
  FA.M = J * d;
  FA.d = d;
  FA.t = 1;
  FA.ell = J;
  FA.log_q = K->log_q;
  // No modulus switch: Rescale(x, log_q, log_q) is the identity on reduced x.
  FA.log_q1 = K->log_q;
  FA.log_q2 = K->log_q;
  FA.cmask = ~0ULL;  // b_agg is reduced mod q; no Rescale
  PIRG_HIP(LaunchFinalizeCombine(FA, st));
  return Ok();
}

Error Create(const pirg_params* params, int njobs, const void* const* pack,
             const uint64_t* const* a_tilde, const uint64_t* const* t_vec_h,
             std::unique_ptr<pirg_packer>* out) {
  Error er = CheckPackParams(*params, njobs);
  if (!er.ok()) return er;
  if (pack == nullptr || a_tilde == nullptr || t_vec_h == nullptr) {
    return Invalid("pack, a_tilde and t_vec_h must hold njobs entries.");
  }
  er = CheckDevice();
  if (!er.ok()) return er;
  std::unique_ptr<pirg_packer> K(new pirg_packer());
  K->P = *params;
  K->J = njobs;
  K->log_d = params->log_degree;
  K->d = 1 << K->log_d;
  K->N = 2 * K->d;
  K->k = params->pack_num_digits;
  K->log_q = params->log_q;
  const int d = K->d, k = K->k, J = K->J;
  K->F = FftDims{d, K->log_d, K->N, K->log_d + 1};
  er = MakeRules(k, d, params->pack_log_digit, params->log_q, ChunkFor(k, d), true, &K->fin_u,
                 &K->fin_v);
  if (!er.ok()) return er;
  K->fin_pairs = MakePairs(K->fin_u, K->fin_v, params->coeff_bits == 32 ? 32 : 64);
  PIRG_HIP(hipGetDevice(&K->device));
  PIRG_HIP(hipStreamCreateWithFlags(&K->stream, hipStreamNonBlocking));
  for (auto& e : K->ev) PIRG_HIP(hipEventCreate(&e));
  PIRG_HIP(ConfigureFftKernels(K->N));
  PIRG_HIP(ConfigureScanKernels());

  // Pack matrices -> one tall int8-limb matrix of J*d rows.
  const int cols = k * d;
  const int planes = params->mat_bytes == 4 ? 4 : 2;
  const u64 c_off = params->mat_bytes == 4 ? 128ULL * (1 + 256 + 65536) : 128ULL;
  K->L = MakeLayout(J * d, cols, planes, c_off);
  K->plan = MakeGroupedScanPlan(K->L, J, d, K->log_q);
  er = K->tiles.alloc(TileBufferBytes(K->L));
  if (!er.ok()) return er;
  const size_t rowbytes = params->mat_bytes == 4 ? static_cast<size_t>(cols) * 4
                                                 : static_cast<size_t>(cols) * 2;
  pirg_packer* Kp = K.get();
  for (int j = 0; j < J; ++j) {
    if (pack[j] == nullptr) return Invalid("pack[" + std::to_string(j) + "] is null.");
    er = UploadTiles(pack[j], rowbytes, d, false, K->stream,
                     [&](const void* src, int r0, int nr) {
                       return LaunchPackToTiles(src, Kp->P.mat_bytes, J * d, cols, j * d + r0,
                                                nr, Kp->L, Kp->tiles.as<int8_t>(), Kp->stream);
                     });
    if (!er.ok()) return er;
  }
  // Resident polynomials and the spectra of t_vec_h.
  er = K->a_tilde.alloc(static_cast<size_t>(J) * d * 8);
  if (!er.ok()) return er;
  DevBuf th;
  er = th.alloc(static_cast<size_t>(J) * k * d * 8);
  if (!er.ok()) return er;
  for (int j = 0; j < J; ++j) {
    PIRG_HIP(CopySync(K->a_tilde.as<u64>() + static_cast<size_t>(j) * d, a_tilde[j],
                      static_cast<size_t>(d) * 8, hipMemcpyHostToDevice, K->stream));
    PIRG_HIP(CopySync(th.as<u64>() + static_cast<size_t>(j) * k * d, t_vec_h[j],
                      static_cast<size_t>(k) * d * 8, hipMemcpyHostToDevice, K->stream));
  }
  std::vector<double2> fwd, bwd;
  Twiddles(K->N, &fwd, &bwd);
  if ((er = K->tw_fwd.alloc(fwd.size() * sizeof(double2))).code) return er;
  if ((er = K->tw_bwd.alloc(bwd.size() * sizeof(double2))).code) return er;
  PIRG_HIP(CopySync(K->tw_fwd.p, fwd.data(), fwd.size() * sizeof(double2), hipMemcpyHostToDevice,
                    K->stream));
  PIRG_HIP(CopySync(K->tw_bwd.p, bwd.data(), bwd.size() * sizeof(double2), hipMemcpyHostToDevice,
                    K->stream));
  if ((er = K->margin.alloc(8)).code) return er;
  PIRG_HIP(FillSync(K->margin.p, 0, 8, K->stream));
  const int nu = K->fin_u.num_chunks, nv = K->fin_v.num_chunks;
  if ((er = K->t_spec.alloc(static_cast<size_t>(J) * k * nu * K->N * sizeof(double2))).code) return er;
  PIRG_HIP(LaunchSpectra(th.as<u64>(), J * k, 1, d, K->fin_u, K->F, K->tw_fwd.as<double2>(),
                         K->t_spec.as<double2>(), K->stream));

  // Workspace.
  const size_t Md = static_cast<size_t>(J) * d;
  struct Item {
    DevBuf* b;
    size_t bytes;
  } items[] = {
      {&K->b_dev, Md * 8},
      {&K->keys_dev, static_cast<size_t>(J) * 2 * k * d * 8},
      {&K->bt, K->plan.btile_bytes},
      {&K->corr, (static_cast<size_t>(J) + VectorSumScratchWords(J)) * 8},
      {&K->zero, 8},
      {&K->partial, static_cast<size_t>(K->plan.ksplit) * Md * 8},
      {&K->yh_spec, static_cast<size_t>(J) * k * nv * K->N * sizeof(double2)},
      {&K->fin_partial, static_cast<size_t>(J) * K->fin_pairs.count * d * 8},
      {&K->ct_b, Md * 8},
      {&K->st, static_cast<size_t>(d) * 8},
      {&K->out, Md * 2 * 8},
      {&K->fin_tab, static_cast<size_t>(J) * 3 * sizeof(void*)},
  };
  for (auto& it : items) {
    if ((er = it.b->alloc(it.bytes)).code) return er;
  }
  PIRG_HIP(FillSync(K->zero.p, 0, 8, K->stream));
  if ((er = K->h_b.alloc(Md * 8)).code) return er;
  if ((er = K->h_keys.alloc(static_cast<size_t>(J) * 2 * k * d * 8)).code) return er;
  if ((er = K->h_out.alloc(Md * 2 * 8)).code) return er;
  std::vector<const void*> tab(static_cast<size_t>(J) * 3);
  for (int o = 0; o < J; ++o) {
    tab[o] = K->t_spec.as<double2>() + static_cast<size_t>(o) * k * nu * K->N;
    tab[J + o] = K->yh_spec.as<double2>() + static_cast<size_t>(o) * k * nv * K->N;
    tab[2 * J + o] = K->fin_partial.as<u64>() + static_cast<size_t>(o) * K->fin_pairs.count * d;
  }
  PIRG_HIP(CopySync(K->fin_tab.p, tab.data(), tab.size() * sizeof(void*), hipMemcpyHostToDevice,
                    K->stream));
  PIRG_HIP(hipStreamSynchronize(K->stream));
  *out = std::move(K);
  return Ok();
}

Error Run(pirg_packer* K, const uint64_t* const* b, const int* m, const uint64_t* const* keys,
          uint64_t* const* outs) {
  const int d = K->d, k = K->k, J = K->J;
  if (b == nullptr || m == nullptr || keys == nullptr || outs == nullptr) {
    return Invalid("b, m, keys and out must hold njobs entries.");
  }
  u64* hb = K->h_b.as<u64>();
  u64* hk = K->h_keys.as<u64>();
  for (int j = 0; j < J; ++j) {
    if (m[j] < 0 || m[j] > d) {
      return Invalid("job " + std::to_string(j) + ": LWE count " + std::to_string(m[j]) +
                     " exceeds the ring degree " + std::to_string(d) + ".");
    }
    if ((m[j] > 0 && b[j] == nullptr) || keys[j] == nullptr) {
      return Invalid("job " + std::to_string(j) + ": null input.");
    }
    u64* dst = hb + static_cast<size_t>(j) * d;
    if (m[j] > 0) std::memcpy(dst, b[j], static_cast<size_t>(m[j]) * 8);
    std::memset(dst + m[j], 0, static_cast<size_t>(d - m[j]) * 8);
    std::memcpy(hk + static_cast<size_t>(j) * 2 * k * d, keys[j],
                static_cast<size_t>(2 * k) * d * 8);
  }
  hipStream_t st = K->stream;
  PIRG_HIP(hipMemcpyAsync(K->b_dev.p, hb, static_cast<size_t>(J) * d * 8, hipMemcpyHostToDevice, st));
  PIRG_HIP(hipMemcpyAsync(K->keys_dev.p, hk, static_cast<size_t>(J) * 2 * k * d * 8,
                          hipMemcpyHostToDevice, st));
  PIRG_HIP(hipEventRecord(K->ev[0], st));
  if (K->use_graph) {
    if (K->graph == nullptr) {
      PIRG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      Error er = Enqueue(K, false);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(st, &g);
      if (!er.ok()) {
        if (g != nullptr) (void)hipGraphDestroy(g);
        return er;
      }
      PIRG_HIP(ce);
      const hipError_t ie = hipGraphInstantiate(&K->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      PIRG_HIP(ie);
    }
    PIRG_HIP(hipGraphLaunch(K->graph, st));
  } else {
    Error er = Enqueue(K, true);
    if (!er.ok()) return er;
  }
  PIRG_HIP(hipEventRecord(K->ev[3], st));
  PIRG_HIP(hipMemcpyAsync(K->h_out.p, K->out.p, static_cast<size_t>(J) * 2 * d * 8,
                          hipMemcpyDeviceToHost, st));
  PIRG_HIP(hipStreamSynchronize(st));
  float total = 0, scan = 0;
  PIRG_HIP(hipEventElapsedTime(&total, K->ev[0], K->ev[3]));
  if (!K->use_graph) PIRG_HIP(hipEventElapsedTime(&scan, K->ev[1], K->ev[2]));
  K->total_ms += total;
  K->scan_ms += scan;
  K->runs += 1;
  for (int j = 0; j < J; ++j) {
    std::memcpy(outs[j], K->h_out.as<u64>() + static_cast<size_t>(j) * 2 * d,
                static_cast<size_t>(2) * d * 8);
  }
  return Ok();
}

}  // namespace
}  // namespace pirg

extern "C" {

int pirg_packer_create(const pirg_params* params, int njobs, const void* const* pack,
                       const uint64_t* const* a_tilde, const uint64_t* const* t_vec_h,
                       pirg_packer** out, char* err, size_t errlen) {
  if (params == nullptr || out == nullptr) return Report(Invalid("null argument."), err, errlen);
  std::unique_ptr<pirg_packer> K;
  Error er = Create(params, njobs, pack, a_tilde, t_vec_h, &K);
  if (!er.ok()) return Report(er, err, errlen);
  *out = K.release();
  return 0;
}

int pirg_packer_run(pirg_packer* K, const uint64_t* const* b, const int* m,
                    const uint64_t* const* keys, uint64_t* const* out, char* err,
                    size_t errlen) {
  if (K == nullptr) return Report(Invalid("invalid packer."), err, errlen);
  Error er = Run(K, b, m, keys, out);
  return er.ok() ? 0 : Report(er, err, errlen);
}

int pirg_packer_use_graph(pirg_packer* K, int enable) {
  if (K == nullptr) return PIRG_INVALID_ARGUMENT;
  K->use_graph = enable != 0;
  return 0;
}

int pirg_packer_times(pirg_packer* K, double* total_ms, double* scan_ms, long long* runs,
                      int reset) {
  if (K == nullptr) return PIRG_INVALID_ARGUMENT;
  if (total_ms) *total_ms = K->total_ms;
  if (scan_ms) *scan_ms = K->scan_ms;
  if (runs) *runs = K->runs;
  if (reset) {
    K->total_ms = K->scan_ms = 0;
    K->runs = 0;
  }
  return 0;
}

void pirg_packer_destroy(pirg_packer* K) { delete K; }

}  // extern "C"
