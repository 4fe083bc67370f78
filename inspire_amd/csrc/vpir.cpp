// vPIR online path (crypto/vpir/server/vpir_server.cc:45-195) over the PIR
// answer machinery: VpirServer's material is a PirServer's with one
// interpolation slice (t = 1), its DB product and pack product are the same
// matrix-vector products, and its b response is FinalizeMatrixPack's b
// (vpir_server.cc:161-192) with no modulus switch -- or, with local finalize,
// LowBits(DB*q + pack*y_g) (:148-160), i.e. FinalizeMatrixPack without the
// t_vec_h inner product. A pirg_vpir_server owns such a t = 1 server (with
// log q1 = log q2 = log q, so the switch is the identity) and converts
// requests and responses.
#include <cstring>
#include <vector>

#include "pirgpu.h"
#include "server_internal.h"

using namespace pirg;

struct pirg_vpir_server {
  pirg_server* inner = nullptr;
  int ell = 0, d = 0, k = 0, n = 0;
  bool local_finalize = false;
  std::vector<u64> zero_h;  // k x d zeros: the y_vec_h slot when finalizing locally
  std::vector<u64> out_buf;  // ell x 2 x d per request of a batch

  ~pirg_vpir_server() {
    if (inner != nullptr) pirg_server_destroy(inner);
  }
};

namespace {

// vpir_server.cc:107-124 (ProcessResponse's request checks, in order).
Error ValidateVpirRequest(const pirg_vpir_server* V, const pirg_vpir_request_view& r) {
  const size_t expected_keys = V->local_finalize ? V->k : 2 * V->k;
  if (r.packing_key_polys != expected_keys) {
    return Invalid("Packing key must contain " + std::to_string(expected_keys) +
                   " polynomials.");
  }
  if (r.first_dim_polys * static_cast<size_t>(V->d) != static_cast<size_t>(V->n)) {
    return Invalid("Flattened first dimension query must match DB matrix columns (N).");
  }
  return Ok();
}

// The inner server's request: the flattened first dimension, no second
// dimension, and 2k packing-key polynomials (zeros for y_vec_h when local).
struct InnerRequests {
  std::vector<pirg_request_view> views;
  std::vector<std::vector<u64>> keys;
};

void MakeInner(const pirg_vpir_server* V, const pirg_vpir_request_view* reqs,
               const std::vector<int>& idx, InnerRequests* out) {
  out->views.resize(idx.size());
  if (V->local_finalize) out->keys.resize(idx.size());
  const size_t kd = static_cast<size_t>(V->k) * V->d;
  for (size_t j = 0; j < idx.size(); ++j) {
    const pirg_vpir_request_view& r = reqs[idx[j]];
    pirg_request_view& v = out->views[j];
    v.first_dim = r.first_dim;
    v.first_dim_len = static_cast<size_t>(V->n);
    v.second_dim = nullptr;
    v.second_dim_polys = 0;
    if (V->local_finalize) {
      std::vector<u64>& key = out->keys[j];
      key.assign(2 * kd, 0);
      std::memcpy(key.data(), r.packing_key, kd * 8);
      v.packing_key = key.data();
    } else {
      v.packing_key = r.packing_key;
    }
    v.packing_key_polys = static_cast<size_t>(2 * V->k);
  }
}

// b of every shard from the inner response (ell x [a, b] x d).
void TakeB(const pirg_vpir_server* V, const u64* inner_out, uint64_t* out) {
  for (int s = 0; s < V->ell; ++s) {
    std::memcpy(out + static_cast<size_t>(s) * V->d,
                inner_out + (static_cast<size_t>(s) * 2 + 1) * V->d,
                static_cast<size_t>(V->d) * 8);
  }
}

}  // namespace

extern "C" {

int pirg_vpir_server_create(const pirg_params* params, int local_finalize,
                            const uint64_t* db_cdata, int db_rows, int db_cols,
                            const void* pack, int num_outputs, const uint64_t* a_tilde,
                            const uint64_t* t_vec_h, pirg_vpir_server** out, char* err,
                            size_t errlen) {
  if (params == nullptr || out == nullptr) {
    return Report(Invalid("params and out must not be null."), err, errlen);
  }
  if (params->interp_degree != 1) {
    return Report(Invalid("vPIR parameters have interpolation degree 1."), err, errlen);
  }
  if (params->log_degree < 1 || params->log_degree > 12) {
    return Report(Invalid("log_degree must be in [1, 12] on the GPU path (2d-point FFT in LDS)."),
                  err, errlen);
  }
  const int d = 1 << params->log_degree;
  const int ell = params->entry_size_multiple;
  if (db_rows != ell * d) return Report(Invalid("DB matrix must have ell * d rows."), err, errlen);
  if (db_cols != params->num_entries) {
    return Report(Invalid("DB matrix must have N columns."), err, errlen);
  }
  if (num_outputs != ell) {
    return Report(Invalid("preprocessed_outputs size mismatch."), err, errlen);
  }
  pirg_params P = *params;
  P.log_q1 = P.log_q;
  P.log_q2 = P.log_q;
  if (P.eval_num_digits <= 0) P.eval_num_digits = 1;
  if (P.eval_log_digit <= 0) P.eval_log_digit = 1;
  std::unique_ptr<pirg_vpir_server> V(new pirg_vpir_server());
  V->ell = ell;
  V->d = d;
  V->k = params->pack_num_digits;
  V->n = params->num_entries;
  V->local_finalize = local_finalize != 0;
  const int rc = pirg_server_create_flat(&P, db_cdata, db_cols, pack, a_tilde, t_vec_h,
                                         nullptr, &V->inner, err, errlen);
  if (rc != 0) return rc;
  if (V->local_finalize) {
    V->inner->local_finalize = true;
    V->inner->deferred = Ok();  // no FFT inner product, so no chunk rule to satisfy
  }
  *out = V.release();
  return 0;
}

void pirg_vpir_server_destroy(pirg_vpir_server* server) { delete server; }

int pirg_vpir_process_batch(pirg_vpir_server* V, int nreq, const pirg_vpir_request_view* reqs,
                            uint64_t* const* outs, int* statuses, char* err, size_t errlen) {
  if (V == nullptr || nreq < 0) return Report(Invalid("invalid server or batch size."), err, errlen);
  std::vector<int> idx;
  Error first_bad;
  for (int i = 0; i < nreq; ++i) {
    Error er = ValidateVpirRequest(V, reqs[i]);
    statuses[i] = er.code;
    if (er.ok()) {
      idx.push_back(i);
    } else if (first_bad.ok()) {
      first_bad = er;
    }
  }
  if (!first_bad.ok()) Report(first_bad, err, errlen);
  if (idx.empty()) return 0;
  InnerRequests in;
  MakeInner(V, reqs, idx, &in);
  const size_t per = static_cast<size_t>(V->ell) * 2 * V->d;
  V->out_buf.resize(per * idx.size());
  std::vector<uint64_t*> inner_outs(idx.size());
  for (size_t j = 0; j < idx.size(); ++j) inner_outs[j] = V->out_buf.data() + j * per;
  std::vector<int> st(idx.size(), 0);
  const int rc = pirg_process_batch(V->inner, static_cast<int>(idx.size()), in.views.data(),
                                    inner_outs.data(), st.data(), err, errlen);
  if (rc != 0) return rc;
  for (size_t j = 0; j < idx.size(); ++j) {
    statuses[idx[j]] = st[j];
    if (st[j] == 0) TakeB(V, inner_outs[j], outs[idx[j]]);
  }
  return 0;
}

int pirg_vpir_process(pirg_vpir_server* V, const pirg_vpir_request_view* req, uint64_t* out,
                      char* err, size_t errlen) {
  if (V == nullptr || req == nullptr) return Report(Invalid("null argument."), err, errlen);
  Error er = ValidateVpirRequest(V, *req);
  if (!er.ok()) return Report(er, err, errlen);
  int status = 0;
  uint64_t* outs[1] = {out};
  const int rc = pirg_vpir_process_batch(V, 1, req, outs, &status, err, errlen);
  return rc != 0 ? rc : status;
}

int pirg_vpir_batch_stage(pirg_vpir_server* V, int nreq, const pirg_vpir_request_view* reqs,
                          char* err, size_t errlen) {
  if (V == nullptr || nreq < 0) return Report(Invalid("invalid server or batch size."), err, errlen);
  std::vector<int> idx(nreq);
  for (int i = 0; i < nreq; ++i) {
    Error er = ValidateVpirRequest(V, reqs[i]);
    if (!er.ok()) return Report(er, err, errlen);
    idx[i] = i;
  }
  InnerRequests in;
  MakeInner(V, reqs, idx, &in);
  return pirg_batch_stage(V->inner, nreq, in.views.data(), err, errlen);
}

int pirg_vpir_batch_run(pirg_vpir_server* V, int nreq, char* err, size_t errlen) {
  if (V == nullptr) return Report(Invalid("null server."), err, errlen);
  return pirg_batch_run(V->inner, nreq, err, errlen);
}

int pirg_vpir_batch_fetch(pirg_vpir_server* V, int nreq, uint64_t* const* outs, char* err,
                          size_t errlen) {
  if (V == nullptr || nreq < 0) return Report(Invalid("invalid server or batch size."), err, errlen);
  const size_t per = static_cast<size_t>(V->ell) * 2 * V->d;
  V->out_buf.resize(per * static_cast<size_t>(nreq));
  std::vector<uint64_t*> inner_outs(nreq);
  for (int j = 0; j < nreq; ++j) inner_outs[j] = V->out_buf.data() + j * per;
  const int rc = pirg_batch_fetch(V->inner, nreq, inner_outs.data(), err, errlen);
  if (rc != 0) return rc;
  for (int j = 0; j < nreq; ++j) TakeB(V, inner_outs[j], outs[j]);
  return 0;
}

int pirg_vpir_autotune(pirg_vpir_server* V, const int* widths, int nwidths, char* report,
                       size_t report_len, char* err, size_t errlen) {
  if (V == nullptr) return Report(Invalid("null server."), err, errlen);
  const int rc = pirg_server_autotune(V->inner, widths, nwidths, report, report_len);
  if (rc != 0 && err != nullptr && errlen > 0) {
    std::snprintf(err, errlen, "autotune failed (status %d)", rc);
  }
  return rc;
}

int pirg_vpir_use_graphs(pirg_vpir_server* V, int enable) {
  return V == nullptr ? 3 : pirg_server_use_graphs(V->inner, enable);
}

int pirg_vpir_time_graph_scans(pirg_vpir_server* V, int enable) {
  return V == nullptr ? 3 : pirg_server_time_graph_scans(V->inner, enable);
}

int pirg_vpir_scan_times(pirg_vpir_server* V, double* db_scan_ms, double* pack_scan_ms,
                         long long* launches, int reset) {
  return V == nullptr ? 3
                      : pirg_server_scan_times(V->inner, db_scan_ms, pack_scan_ms, launches,
                                               reset);
}

}  // extern "C"
