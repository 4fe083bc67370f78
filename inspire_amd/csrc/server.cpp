// Host side of libpirgpu: the C ABI of include/pirgpu.h.
//
// pirg_server mirrors PirServer<Db, uint64_t, Mat>
// (crypto/pir/server/pir_server.h:41-111): Create validates and takes the
// preprocessed material (here: into HBM), ProcessResponse answers a request.
// Every validation message is the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "server_internal.h"

extern "C" void pirg_internal_twiddles(int n, double* fwd, double* bwd);

namespace pirg {

int Report(const Error& e, char* err, size_t errlen) {
  if (err != nullptr && errlen > 0) std::snprintf(err, errlen, "%s", e.msg.c_str());
  return e.code;
}

void DevBuf::reset() {
  if (p != nullptr) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

Error DevBuf::alloc(size_t n) {
  reset();
  if (n == 0) n = 8;
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    return Error{PIRG_RESOURCE_EXHAUSTED, "device allocation of " + std::to_string(n) +
                                              " bytes failed: " + hipGetErrorString(e)};
  }
  bytes = n;
  return Ok();
}

void HostBuf::reset() {
  if (p != nullptr) (void)hipHostFree(p);
  p = nullptr;
  bytes = 0;
}

Error HostBuf::alloc(size_t n) {
  reset();
  if (n == 0) n = 8;
  // Coherent: kernels that read the buffer (host_io batches) see what the
  // host wrote before the launch, never a line cached by an earlier read,
  // and their writes reach host memory directly. Copy engines and the host
  // are unaffected.
  hipError_t e = hipHostMalloc(&p, n, hipHostMallocCoherent);
  if (e != hipSuccess) {
    p = nullptr;
    return Error{PIRG_RESOURCE_EXHAUSTED, "pinned host allocation failed"};
  }
  bytes = n;
  return Ok();
}

Slot::~Slot() {
  if (stream != nullptr) (void)hipStreamSynchronize(stream);
  DropGraphs();
  for (auto& x : ev) {
    if (x != nullptr) (void)hipEventDestroy(x);
  }
  for (auto& x : kev) {
    if (x != nullptr) (void)hipEventDestroy(x);
  }
  if (done != nullptr) (void)hipEventDestroy(done);
  if (graph_done != nullptr) (void)hipEventDestroy(graph_done);
  if (side != nullptr) (void)hipStreamSynchronize(side);
  if (fork_ev != nullptr) (void)hipEventDestroy(fork_ev);
  if (join_ev != nullptr) (void)hipEventDestroy(join_ev);
  if (pk_join_ev != nullptr) (void)hipEventDestroy(pk_join_ev);
  if (hfork_ev != nullptr) (void)hipEventDestroy(hfork_ev);
  if (hjoin_ev != nullptr) (void)hipEventDestroy(hjoin_ev);
  if (side != nullptr) (void)hipStreamDestroy(side);
  if (stream != nullptr) (void)hipStreamDestroy(stream);
}

// Streams keep the default priority.
Error Slot::Init() {
  PIRG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  for (auto& x : ev) PIRG_HIP(hipEventCreate(&x));
  for (auto& x : kev) PIRG_HIP(hipEventCreate(&x));
  PIRG_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
  PIRG_HIP(hipEventCreateWithFlags(&graph_done, hipEventDisableTiming));
  PIRG_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  PIRG_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
  PIRG_HIP(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
  PIRG_HIP(hipEventCreateWithFlags(&pk_join_ev, hipEventDisableTiming));
  PIRG_HIP(hipEventCreateWithFlags(&hfork_ev, hipEventDisableTiming));
  PIRG_HIP(hipEventCreateWithFlags(&hjoin_ev, hipEventDisableTiming));
  return Ok();
}

void Slot::DropGraphs() {
  for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  graphs.clear();
}

namespace {

bool IsPow2(u64 x) { return x != 0 && (x & (x - 1)) == 0; }

}  // namespace

Error CheckDevice() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    return Error{PIRG_UNAVAILABLE, "no HIP device available"};
  }
  return Ok();
}

// crypto/lwes_to_rlwe.cc:60-74 and crypto/external_product.cc:100-102.
int ChunkFor(int n_terms, int d) {
  const int max_allowed = static_cast<int>(
      (53.0 - std::log2(static_cast<double>(n_terms)) -
       std::log2(static_cast<double>(d))) /
      2.0);
  return std::min(20, max_allowed);
}

// crypto/polynomial_fft.cc:90-107.
Error MakeRules(int n_terms, int d, int u_bits, int v_bits, int chunk,
                bool u_signed, ChunkRule* ur, ChunkRule* vr) {
  if (u_bits <= 0 || v_bits <= 0 || chunk <= 0 || chunk > 63) {
    return Invalid("Invalid bits or chunk_bits.");
  }
  if (std::log2(static_cast<double>(n_terms)) + std::log2(static_cast<double>(d)) +
          2.0 * chunk >
      53.0) {
    return Invalid("Potential precision overflow in accumulation.");
  }
  const int max_chunks = (64 + chunk - 1) / chunk;
  ur->chunk_bits = vr->chunk_bits = chunk;
  ur->num_chunks = std::min((u_bits + chunk - 1) / chunk, max_chunks);
  vr->num_chunks = std::min((v_bits + chunk - 1) / chunk, max_chunks);
  ur->bits = u_bits;
  vr->bits = v_bits;
  ur->is_signed = u_signed ? 1 : 0;
  vr->is_signed = 0;
  return Ok();
}

}  // namespace pirg

using namespace pirg;

namespace pirg {

// pir_params.cc:29-68 and the GPU path's own limits.
Error DeriveServer(const pirg_params& P, pirg_server* S) {
  if (P.num_entries <= 0) return Invalid("num_entries must be strictly positive.");
  if (P.entry_size_multiple <= 0) {
    return Invalid("entry_size_multiple must be strictly positive.");
  }
  if (P.interp_degree <= 0 || !IsPow2(static_cast<u64>(P.interp_degree))) {
    return Invalid("interpolation_degree must be a power of two.");
  }
  if (P.num_entries % P.interp_degree != 0) {
    return Invalid("interpolation_degree (" + std::to_string(P.interp_degree) +
                   ") must divide num_entries (" + std::to_string(P.num_entries) +
                   ").");
  }
  if (P.log_degree < 1 || P.log_degree > 12) {
    return Invalid("log_degree must be in [1, 12] on the GPU path (2d-point FFT in LDS).");
  }
  const int d = 1 << P.log_degree;
  if (P.interp_degree > 2 * d) {
    return Invalid("interpolation_degree (" + std::to_string(P.interp_degree) +
                   ") must be less than or equal to 2 * ring degree (" +
                   std::to_string(2 * d) + ").");
  }
  if (P.log_q < 1 || P.log_q > 64 || P.log_q1 < 1 || P.log_q1 > 64 ||
      P.log_q2 < 1 || P.log_q2 > 64) {
    return Invalid("log moduli must be in [1, 64].");
  }
  if (P.coeff_bits != 0 && P.coeff_bits != 32 && P.coeff_bits != 64) {
    return Invalid("coeff_bits must be 32 or 64.");
  }
  const int width = P.coeff_bits == 32 ? 32 : 64;
  if (P.log_q > width) {
    return Invalid("log_q exceeds the coefficient width (" + std::to_string(width) + " bits).");
  }
  if (P.db_bytes != 1 && P.db_bytes != 2) return Invalid("db_bytes must be 1 or 2.");
  if (P.mat_bytes != 2 && P.mat_bytes != 4 && P.mat_bytes != 8) {
    return Invalid("mat_bytes must be 2, 4 or 8.");
  }
  if (P.mat_bytes == 8 && width != 64) {
    return Invalid("mat_bytes 8 (a uint64_t pack matrix) needs 64-bit coefficients.");
  }
  if (P.pack_num_digits <= 0 || P.eval_num_digits <= 0) {
    return Invalid("log_digit and num_digits must be positive.");
  }
  S->P = P;
  S->d = d;
  S->log_d = P.log_degree;
  S->N = 2 * d;
  S->log_N = P.log_degree + 1;
  S->F = FftDims{d, P.log_degree, 2 * d, P.log_degree + 1};
  S->t = P.interp_degree;
  S->ell = P.entry_size_multiple;
  S->k = P.pack_num_digits;
  S->e = P.eval_num_digits;
  S->r = P.num_entries / P.interp_degree;
  S->M = S->ell * S->t * d;
  S->log_q = P.log_q;
  S->qmask = P.log_q >= 64 ? ~0ULL : ((1ULL << P.log_q) - 1);
  S->width = width;
  S->cmask = width >= 64 ? ~0ULL : ((1ULL << width) - 1);
  S->mat_src_bytes = P.mat_bytes;
  // FinalizeMatrixPack's InnerProduct (lwes_to_rlwe.cc:472-474).
  Error er = MakeRules(S->k, d, P.pack_log_digit, P.log_q, ChunkFor(S->k, d), true,
                       &S->fin_u, &S->fin_v);
  if (!er.ok()) {
    S->deferred = er;
  } else {
    S->fin_pairs = MakePairs(S->fin_u, S->fin_v, width);
  }
  if (S->t > 1 && S->deferred.ok()) {
    // ExternalProduct (external_product.cc:100-111).
    er = MakeRules(2 * S->e, d, P.eval_log_digit, P.log_q, ChunkFor(2 * S->e, d),
                   false, &S->ext_u, &S->ext_v);
    if (!er.ok()) {
      S->deferred = er;
    } else {
      S->ext_pairs = MakePairs(S->ext_u, S->ext_v, width);
    }
  }
  if (S->fin_pairs.count > 64 || S->ext_pairs.count > 64) {
    return Invalid("too many FFT chunk pairs for the GPU path.");
  }
  return Ok();
}

// crypto/fft.h:71-80, from the g++-built table (twiddles.cpp), reordered by
// stage for the kernels: stage s (1..log2 n) occupies entries
// [2^(s-1) - 1, 2^s - 1), and entry 2^(s-1) - 1 + k is tw[k * (n >> s)], the
// twiddle the reference's stage s reads for offset k. n - 1 entries.
void Twiddles(int n, std::vector<double2>* fwd, std::vector<double2>* bwd) {
  std::vector<double> f(n), b(n);
  pirg_internal_twiddles(n, f.data(), b.data());
  fwd->assign(n - 1, make_double2(0.0, 0.0));
  bwd->assign(n - 1, make_double2(0.0, 0.0));
  for (int s = 1; (1 << s) <= n; ++s) {
    const int base = (1 << (s - 1)) - 1;
    const int step = n >> s;
    for (int k = 0; k < (1 << (s - 1)); ++k) {
      const int i = k * step;
      (*fwd)[base + k] = make_double2(f[2 * i], f[2 * i + 1]);
      (*bwd)[base + k] = make_double2(b[2 * i], b[2 * i + 1]);
    }
  }
}

LimbLayout MakeLayout(int rows, int cols, int planes, u64 c_off) {
  LimbLayout L;
  L.rows = rows;
  L.cols = cols;
  L.strips = (rows + 15) / 16;
  L.kblocks = (cols + 63) / 64;
  L.planes = planes;
  L.c_off = c_off;
  return L;
}

namespace {

// Pack rows [0, M) from a bit-packed image: row chunks of the stream are
// staged, unpacked (sign-extended) into int32/int16 rows on the GPU and
// converted to tiles.
Error UploadPackedPack(pirg_server* S, const PackedPack& pk, int pk_cols) {
  const int M = S->M;
  const int eb = S->P.mat_bytes == 4 ? 4 : 2;
  const size_t row_bits = static_cast<size_t>(pk_cols) * pk.bits;
  size_t chunk_bytes = static_cast<size_t>(256) << 20;  // decoded rows per chunk
  if (const char* env = std::getenv("PIRG_UPLOAD_CHUNK_BYTES")) {
    const long long v = std::atoll(env);
    if (v > 0) chunk_bytes = static_cast<size_t>(v);
  }
  int rows_per = static_cast<int>(
      std::max<size_t>(16, chunk_bytes / (static_cast<size_t>(pk_cols) * eb)));
  rows_per = std::max(16, rows_per / 16 * 16);
  const int chunk_rows = std::min(rows_per, M);
  const size_t max_bytes = (static_cast<size_t>(chunk_rows) * row_bits + 7) / 8 + 8;
  DevBuf words, rowsbuf;
  Error er = words.alloc((max_bytes + 7) / 8 * 8 + 8);
  if (!er.ok()) return er;
  er = rowsbuf.alloc(static_cast<size_t>(chunk_rows) * pk_cols * eb);
  if (!er.ok()) return er;
  hipStream_t st = S->main.stream;
  for (int r0 = 0; r0 < M; r0 += rows_per) {
    const int nr = std::min(rows_per, M - r0);
    const size_t bit_a = static_cast<size_t>(r0) * row_bits;
    const size_t bit_b = static_cast<size_t>(r0 + nr) * row_bits;
    const size_t byte_a = bit_a / 8;
    const size_t byte_b = (bit_b + 7) / 8;
    if (byte_b > pk.nbytes) return Invalid("pack matrix bytes are short.");
    PIRG_HIP(hipMemsetAsync(words.p, 0, words.bytes, st));
    PIRG_HIP(hipMemcpyAsync(words.p, pk.bytes + byte_a, byte_b - byte_a, hipMemcpyHostToDevice,
                            st));
    PIRG_HIP(LaunchBitUnpack(words.as<u64>(), 0, static_cast<long long>(bit_a % 8), 1,
                             static_cast<long long>(nr) * pk_cols, pk.bits, true, eb, rowsbuf.p,
                             st));
    PIRG_HIP(LaunchPackToTiles(rowsbuf.p, S->P.mat_bytes, M, pk_cols, r0, nr, S->pkL,
                               S->pk_tiles.as<int8_t>(), st));
    PIRG_HIP(hipStreamSynchronize(st));
  }
  return Ok();
}

// Polynomial sources may be host or device pointers (hipMemcpyDefault).
Error BuildResident(pirg_server* S, const std::vector<const uint64_t*>& db_parts,
                    int db_cols, const std::vector<const void*>& pack_parts,
                    const std::vector<const uint64_t*>& a_parts,
                    const std::vector<const uint64_t*>& th_parts,
                    const uint64_t* second_dim_a, const PackedPack* packed = nullptr,
                    bool parts_on_device = false) {
  const int d = S->d;
  S->K = db_cols;
  PIRG_HIP(hipGetDevice(&S->device));
  PIRG_HIP(hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, S->device));
  {
    Error ie = S->main.Init();
    if (!ie.ok()) return ie;
  }
  PIRG_HIP(ConfigureScanKernels());
  PIRG_HIP(ConfigureFftKernels(S->N));

  // DB planes.
  const int db_planes = S->P.db_bytes == 2 ? 2 : 1;
  const u64 db_off = S->P.db_bytes == 2 ? 32896ULL : 128ULL;
  S->dbL = MakeLayout(S->M, db_cols, db_planes, db_off);
  Error er = S->db_tiles.alloc(TileBufferBytes(S->dbL));
  if (!er.ok()) return er;
  const int wpr = S->P.db_bytes == 2 ? db_cols / 4 : db_cols / 8;
  const int rows_per_part = S->M / static_cast<int>(db_parts.size());
  for (size_t part = 0; part < db_parts.size(); ++part) {
    const int base = static_cast<int>(part) * rows_per_part;
    er = UploadTiles(db_parts[part], static_cast<size_t>(wpr) * 8, rows_per_part,
                     parts_on_device, S->main.stream, [&](const void* src, int r0, int nr) {
                       return LaunchDbToTiles(static_cast<const u64*>(src), S->P.db_bytes,
                                              S->M, db_cols, base + r0, nr, S->dbL,
                                              S->db_tiles.as<int8_t>(), S->main.stream);
                     });
    if (!er.ok()) return er;
  }

  // Pack planes: int32 -> 3 offset bytes + signed top byte; int16 -> 1 + 1.
  const int pk_cols = S->k * d;
  const int pk_planes = S->P.mat_bytes == 4 ? 4 : 2;
  const u64 pk_off = S->P.mat_bytes == 4 ? 128ULL * (1 + 256 + 65536) : 128ULL;
  S->pkL = MakeLayout(S->M, pk_cols, pk_planes, pk_off);
  er = S->pk_tiles.alloc(TileBufferBytes(S->pkL));
  if (!er.ok()) return er;
  const size_t pk_rowbytes = S->P.mat_bytes == 4 ? static_cast<size_t>(pk_cols) * 4
                                                 : static_cast<size_t>(pk_cols) * 2;
  if (packed != nullptr) {
    er = UploadPackedPack(S, *packed, pk_cols);
    if (!er.ok()) return er;
  }
  const int pk_rows_per_part =
      packed != nullptr ? 0 : S->M / static_cast<int>(pack_parts.size());
  for (size_t part = 0; packed == nullptr && part < pack_parts.size(); ++part) {
    const int base = static_cast<int>(part) * pk_rows_per_part;
    er = UploadTiles(pack_parts[part], pk_rowbytes, pk_rows_per_part, parts_on_device,
                     S->main.stream,
                     [&](const void* src, int r0, int nr) {
                       return LaunchPackToTiles(src, S->P.mat_bytes, S->M, pk_cols,
                                                base + r0, nr, S->pkL,
                                                S->pk_tiles.as<int8_t>(), S->main.stream);
                     });
    if (!er.ok()) return er;
  }

  // Small resident polynomials.
  const int lt = S->ell * S->t;
  er = S->a_tilde.alloc(static_cast<size_t>(lt) * d * 8);
  if (!er.ok()) return er;
  for (int oi = 0; oi < lt; ++oi) {
    PIRG_HIP(CopySync(S->a_tilde.as<u64>() + static_cast<size_t>(oi) * d, a_parts[oi],
                      static_cast<size_t>(d) * 8, hipMemcpyDefault, S->main.stream));
  }
  DevBuf& th = S->t_raw;
  er = th.alloc(static_cast<size_t>(lt) * S->k * d * 8);
  if (!er.ok()) return er;
  for (int oi = 0; oi < lt; ++oi) {
    PIRG_HIP(CopySync(th.as<u64>() + static_cast<size_t>(oi) * S->k * d, th_parts[oi],
                      static_cast<size_t>(S->k) * d * 8, hipMemcpyDefault, S->main.stream));
  }
  if (S->t > 1) {
    er = S->sda_raw.alloc(static_cast<size_t>(2 * S->e) * d * 8);
    if (!er.ok()) return er;
    PIRG_HIP(CopySync(S->sda_raw.p, second_dim_a, static_cast<size_t>(2 * S->e) * d * 8,
                      hipMemcpyDefault, S->main.stream));
  }

  // Twiddles.
  std::vector<double2> fwd, bwd;
  Twiddles(S->N, &fwd, &bwd);
  er = S->tw_fwd.alloc(fwd.size() * sizeof(double2));
  if (!er.ok()) return er;
  er = S->tw_bwd.alloc(bwd.size() * sizeof(double2));
  if (!er.ok()) return er;
  PIRG_HIP(CopySync(S->tw_fwd.p, fwd.data(), fwd.size() * sizeof(double2), hipMemcpyHostToDevice,
                    S->main.stream));
  PIRG_HIP(CopySync(S->tw_bwd.p, bwd.data(), bwd.size() * sizeof(double2), hipMemcpyHostToDevice,
                    S->main.stream));
  er = S->margin.alloc(8);
  if (!er.ok()) return er;
  PIRG_HIP(FillSync(S->margin.p, 0, 8, S->main.stream));

  // Resident spectra: t_vec_h chunks (finalize u operand) and second_dim_a
  // chunks (external product v operand of the a-component).
  if (S->deferred.ok()) {
    er = S->t_spec.alloc(static_cast<size_t>(lt) * S->k * S->fin_u.num_chunks * S->N *
                         sizeof(double2));
    if (!er.ok()) return er;
    PIRG_HIP(LaunchSpectra(th.as<u64>(), lt * S->k, 1, d, S->fin_u, S->F,
                           S->tw_fwd.as<double2>(), S->t_spec.as<double2>(), S->main.stream));
    if (S->t > 1) {
      const DevBuf& sda = S->sda_raw;
      er = S->sda_spec.alloc(static_cast<size_t>(2 * S->e) * S->ext_v.num_chunks * S->N *
                             sizeof(double2));
      if (!er.ok()) return er;
      PIRG_HIP(LaunchSpectra(sda.as<u64>(), 2 * S->e, 1, d, S->ext_v, S->F,
                             S->tw_fwd.as<double2>(), S->sda_spec.as<double2>(), S->main.stream));
      PIRG_HIP(hipStreamSynchronize(S->main.stream));
    }
    PIRG_HIP(hipStreamSynchronize(S->main.stream));
  }
  return Ok();
}

}  // namespace

size_t ResponseWords(const pirg_server* S) { return static_cast<size_t>(S->ell) * 2 * S->d; }

ScanPlan PlanFor(const pirg_server* S, bool db, int nq) {
  const auto& tuned = db ? S->tuned_db : S->tuned_pk;
  auto it = tuned.find(nq);
  if (it != tuned.end()) return it->second;
  return MakeScanPlan(db ? S->dbL : S->pkL, nq, S->log_q, S->wide_min, S->digit_major, S->tall,
                      S->pack_stream);
}

namespace {

// Device buffer sizes of a workspace for `cap` requests, in the order of
// Workspace's members fd .. ext_tab.
std::vector<size_t> WorkspaceSizes(const pirg_server* S, int cap) {
  const size_t d = S->d, N = S->N;
  const int lt = S->ell * S->t;
  const size_t nq = cap;
  // Scan buffers: the largest plan over every batch size up to cap.
  size_t bt_db = 0, bt_pk = 0, part_db = 0, part_pk = 0;
  for (int n = 1; n <= cap; ++n) {
    const ScanPlan pd = PlanFor(S, true, n);
    const ScanPlan pp = PlanFor(S, false, n);
    bt_db = std::max(bt_db, pd.btile_bytes);
    bt_pk = std::max(bt_pk, pp.btile_bytes);
    part_db = std::max(part_db, static_cast<size_t>(pd.ksplit) * n);
    part_pk = std::max(part_pk, static_cast<size_t>(pp.ksplit) * n);
  }
  // Column-sum scratch of LaunchQueryDigitsSums for each product's layout.
  const size_t corr_db_words =
      nq + static_cast<size_t>(cap) * std::max(kVectorSumChunks, DigitSumChunks(S->dbL));
  const size_t corr_pk_words =
      nq + static_cast<size_t>(cap) * std::max(kVectorSumChunks, DigitSumChunks(S->pkL));
  const int nv_f = S->fin_v.num_chunks;
  const int nu_e = S->t > 1 ? S->ext_u.num_chunks : 0;
  const int nv_e = S->t > 1 ? S->ext_v.num_chunks : 0;
  return {
      nq * RequestWords(S) * 8,
      bt_db,
      bt_pk,
      corr_db_words * 8,
      corr_pk_words * 8,
      part_db * S->M * 8,
      part_pk * S->M * 8,
      nq * S->k * nv_f * N * sizeof(double2),
      nq * 2 * S->e * nv_e * N * sizeof(double2),
      nq * S->ell * 2 * S->e * nu_e * N * sizeof(double2),
      nq * lt * S->fin_pairs.count * d * 8,
      nq * S->ell * 2 * std::max(1, S->ext_pairs.count) * d * 8,
      nq * lt * d * 8,
      nq * S->ell * d * 8,
      nq * S->ell * d * 8,
      nq * S->ell * 2 * d * 8,
      nq * lt * 3 * sizeof(void*),
      nq * S->ell * 2 * 3 * sizeof(void*),
      static_cast<size_t>(kIpSplitMaxJobs) * N * sizeof(double2),
  };
}

}  // namespace

size_t WorkspaceBytes(const pirg_server* S, int cap) {
  size_t total = 0;
  for (size_t b : WorkspaceSizes(S, cap)) total += b;
  return total;
}

Error EnsureWorkspace(pirg_server* S, Slot& slot, int cap) {
  Workspace& W = slot.ws;
  if (W.cap >= cap) return Ok();
  // Captured graphs hold the old buffer addresses.
  (void)hipStreamSynchronize(slot.stream);
  slot.DropGraphs();
  const size_t d = S->d, N = S->N;
  const int lt = S->ell * S->t;
  const size_t nq = cap;
  const int nu_f = S->fin_u.num_chunks, nv_f = S->fin_v.num_chunks;
  const int nu_e = S->t > 1 ? S->ext_u.num_chunks : 0;
  const int nv_e = S->t > 1 ? S->ext_v.num_chunks : 0;
  const std::vector<size_t> sizes = WorkspaceSizes(S, cap);
  DevBuf* bufs[] = {&W.req,                                      &W.bt_db,      &W.bt_pk,
                    &W.corr_db,  &W.corr_pk,    &W.db_partial,  &W.pk_partial, &W.yh_spec,
                    &W.sdb_spec, &W.d_spec,     &W.fin_partial, &W.ext_partial, &W.ct_b,
                    &W.st_a,     &W.st_b,       &W.out,         &W.fin_tab,    &W.ext_tab,
                    &W.mac_acc};
  W.cap = 0;
  for (size_t i = 0; i < sizes.size(); ++i) {
    Error er = bufs[i]->alloc(sizes[i]);
    if (!er.ok()) return er;
  }
  // Zeroed once: a batch padded past its requests (the serving loop rounds
  // batch sizes up) then reads zero records, or earlier valid ones.
  PIRG_HIP(FillSync(W.req.p, 0, sizes[0], slot.stream));
  Error er = W.h_req.alloc(nq * RequestWords(S) * 8);
  if (!er.ok()) return er;
  er = W.h_out.alloc(nq * S->ell * 2 * d * 8);
  if (!er.ok()) return er;

  // IpKernel pointer tables (q-major, so any prefix of queries is valid).
  std::vector<const void*> tab(nq * lt * 3);
  const size_t n_fin = nq * lt;
  for (size_t o = 0; o < n_fin; ++o) {
    const size_t q = o / lt, oi = o % lt;
    tab[o] = S->t_spec.as<double2>() + oi * S->k * nu_f * N;
    tab[n_fin + o] = W.yh_spec.as<double2>() + q * S->k * nv_f * N;
    tab[2 * n_fin + o] = W.fin_partial.as<u64>() + o * S->fin_pairs.count * d;
  }
  PIRG_HIP(CopySync(W.fin_tab.p, tab.data(), tab.size() * sizeof(void*), hipMemcpyHostToDevice,
                    slot.stream));
  if (S->t > 1) {
    const size_t n_ext = nq * S->ell * 2;
    std::vector<const void*> et(n_ext * 3);
    for (size_t o = 0; o < n_ext; ++o) {
      const size_t qs = o / 2, which = o % 2, q = qs / S->ell;
      et[o] = W.d_spec.as<double2>() + qs * 2 * S->e * nu_e * N;
      et[n_ext + o] = which == 0 ? S->sda_spec.as<double2>()
                                 : W.sdb_spec.as<double2>() + q * 2 * S->e * nv_e * N;
      et[2 * n_ext + o] = W.ext_partial.as<u64>() + o * S->ext_pairs.count * d;
    }
    PIRG_HIP(CopySync(W.ext_tab.p, et.data(), et.size() * sizeof(void*), hipMemcpyHostToDevice,
                      slot.stream));
  }
  W.cap = cap;
  return Ok();
}

// pir_server.cc:161-194 request validation, in the reference's order.
Error ValidateRequest(const pirg_server* S, size_t fd_len, size_t sd_polys,
                      size_t pk_polys) {
  if (pk_polys != static_cast<size_t>(2 * S->k)) {
    return Invalid("Packing key must contain 2 * num_digits polynomials.");
  }
  if (fd_len < static_cast<size_t>(S->r) || fd_len > static_cast<size_t>(S->K)) {
    return Invalid("First dimension query size (" + std::to_string(fd_len) +
                   ") is invalid for DB columns (" + std::to_string(S->K) + ").");
  }
  if (S->t > 1 && sd_polys != static_cast<size_t>(2 * S->e)) {
    return Invalid("second_dimension_query size mismatch. Must be 2 * num_digits.");
  }
  if (!S->deferred.ok()) return S->deferred;
  return Ok();
}

size_t RecordSdOffset(const pirg_server* S) { return static_cast<size_t>(S->K); }
size_t RecordPkOffset(const pirg_server* S) {
  return static_cast<size_t>(S->K) + (S->t > 1 ? static_cast<size_t>(2 * S->e) * S->d : 0);
}
size_t RequestWords(const pirg_server* S) {
  return RecordPkOffset(S) + static_cast<size_t>(2 * S->k) * S->d;
}

void WriteRecord(const pirg_server* S, const pirg_request_view& R, u64* dst) {
  const size_t d = S->d;
  std::memcpy(dst, R.first_dim, R.first_dim_len * 8);
  std::memset(dst + R.first_dim_len, 0, (S->K - R.first_dim_len) * 8);
  if (S->t > 1) {
    std::memcpy(dst + RecordSdOffset(S), R.second_dim, static_cast<size_t>(2 * S->e) * d * 8);
  }
  std::memcpy(dst + RecordPkOffset(S), R.packing_key, static_cast<size_t>(2 * S->k) * d * 8);
}

Error StageRecords(pirg_server* S, Slot& slot, const u64* const* records, int nq) {
  const size_t rw = RequestWords(S);
  u64* dev = slot.ws.req.as<u64>();
  // Records that follow each other in host memory go as one copy.
  for (int i = 0; i < nq;) {
    int j = i + 1;
    while (j < nq && records[j] == records[j - 1] + rw) ++j;
    PIRG_HIP(hipMemcpyAsync(dev + static_cast<size_t>(i) * rw, records[i],
                            static_cast<size_t>(j - i) * rw * 8, hipMemcpyHostToDevice,
                            slot.stream));
    i = j;
  }
  return Ok();
}

namespace {

// Words [lo, hi) of request R's record, written to rec + lo.
void WriteRecordPart(const pirg_server* S, const pirg_request_view& R, u64* rec, size_t lo,
                     size_t hi) {
  struct Seg {
    const u64* src;  // nullptr: zeros
    size_t len;
  };
  const size_t d = S->d;
  const size_t sd_len = S->t > 1 ? static_cast<size_t>(2 * S->e) * d : 0;
  const Seg segs[4] = {{R.first_dim, R.first_dim_len},
                       {nullptr, static_cast<size_t>(S->K) - R.first_dim_len},
                       {R.second_dim, sd_len},
                       {R.packing_key, static_cast<size_t>(2 * S->k) * d}};
  size_t off = 0;
  for (const Seg& g : segs) {
    const size_t a = std::max(lo, off), b = std::min(hi, off + g.len);
    if (a < b) {
      if (g.src != nullptr) {
        std::memcpy(rec + a, g.src + (a - off), (b - a) * 8);
      } else {
        std::memset(rec + a, 0, (b - a) * 8);
      }
    }
    off += g.len;
  }
}

// Worker threads for host copies into pinned buffers, started once per
// process: one thread copies about 10 GB/s from memory the caches do not
// hold, so a single 1.4 MB request takes over 100 us alone, and a thread
// started per call would cost as much as it saves.
class HostCopyPool {
 public:
  static HostCopyPool& Get() {
    static HostCopyPool pool;
    return pool;
  }

  // Runs fn(i) for every i < n on the pool's threads and the caller, and
  // returns when all have run. Concurrent calls share the threads.
  void Run(size_t n, const std::function<void(size_t)>& fn) {
    auto job = std::make_shared<Job>();
    job->fn = &fn;
    job->n = n;
    {
      std::lock_guard<std::mutex> lock(mu_);
      job_ = job;
      ++gen_;
    }
    cv_.notify_all();
    Work(*job);
    while (job->done.load(std::memory_order_acquire) < n) std::this_thread::yield();
    std::lock_guard<std::mutex> lock(mu_);
    if (job_ == job) job_.reset();
  }

 private:
  // A worker that takes a job after its last index was claimed runs nothing,
  // so fn (the caller's) is called only while Run waits.
  struct Job {
    const std::function<void(size_t)>* fn = nullptr;
    size_t n = 0;
    std::atomic<size_t> next{0}, done{0};
  };

  HostCopyPool() {
    const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
    const unsigned workers = std::min(7u, hw - 1);
    for (unsigned i = 0; i < workers; ++i) threads_.emplace_back([this] { Loop(); });
  }
  ~HostCopyPool() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      quit_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }

  static void Work(Job& j) {
    for (;;) {
      const size_t i = j.next.fetch_add(1, std::memory_order_relaxed);
      if (i >= j.n) return;
      (*j.fn)(i);
      j.done.fetch_add(1, std::memory_order_release);
    }
  }

  void Loop() {
    unsigned long long seen = 0;
    for (;;) {
      std::shared_ptr<Job> j;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return quit_ || gen_ != seen; });
        if (quit_) return;
        seen = gen_;
        j = job_;
      }
      if (j) Work(*j);
    }
  }

  std::mutex mu_;
  std::condition_variable cv_;
  std::shared_ptr<Job> job_;
  unsigned long long gen_ = 0;
  bool quit_ = false;
  std::vector<std::thread> threads_;
};

}  // namespace

void WriteRecords(pirg_server* S, Slot& slot, const pirg_request_view* reqs,
                  const std::vector<int>& idx) {
  const size_t nq = idx.size();
  const size_t rw = RequestWords(S);
  u64* h = slot.ws.h_req.as<u64>();
  const size_t total = nq * rw;
  // 128 KB pieces, each within one record or across two, on the copy pool
  // from 256 KB up.
  constexpr size_t kPiece = size_t(16) << 10;
  auto piece = [&](size_t c) {
    const size_t lo = c * kPiece, hi = std::min(total, lo + kPiece);
    for (size_t w = lo; w < hi;) {
      const size_t q = w / rw, r0 = w - q * rw;
      const size_t r1 = std::min(rw, r0 + (hi - w));
      WriteRecordPart(S, reqs[idx[q]], h + q * rw, r0, r1);
      w += r1 - r0;
    }
  };
  const size_t pieces = (total + kPiece - 1) / kPiece;
  if (total * 8 < (size_t(256) << 10)) {
    for (size_t c = 0; c < pieces; ++c) piece(c);
  } else {
    HostCopyPool::Get().Run(pieces, piece);
  }
}

// Copies validated requests (reqs[idx[i]]) as records into the pinned
// staging buffer and on to HBM.
Error StageRequests(pirg_server* S, Slot& slot, const pirg_request_view* reqs,
                    const std::vector<int>& idx) {
  WriteRecords(S, slot, reqs, idx);
  const size_t total = RequestWords(S) * 8 * idx.size();
  PIRG_HIP(hipMemcpyAsync(slot.ws.req.p, slot.ws.h_req.p, total, hipMemcpyHostToDevice,
                          slot.stream));
  return Ok();
}

namespace {

// ProcessResponse for nq staged requests (pir_server.cc:207-305), enqueued on
// the slot's stream. `timed` records the stage and scan-kernel events (it is
// off while the sequence is captured into a graph).
// Whether a batch with these (MFMA-bound) plans runs its DB and pack scans as
// one launch: the kinds LaunchScanPair fuses, and a DB scan of at most four
// rounds of workgroups (see EnqueueBatch).
bool FuseScans(const pirg_server* S, int nq, const ScanPlan& pd, const ScanPlan& pp) {
  if (!pd.wide || !S->fuse_scans || !ScanPairFusable(S->dbL, pd, S->pkL, pp) ||
      ScanWorkgroups(S->dbL, pd) > 4LL * S->cus) {
    return false;
  }
  // A tuned width keeps the autotuner's measured choice.
  const auto it = S->tuned_fuse.find(nq);
  return it == S->tuned_fuse.end() || it->second;
}

// FinalizeMatrixPack's inner-product tiles when a request has many (shard, i)
// blocks: requests x blocks per XCD.
constexpr int kFinTileRequests = 8;
constexpr int kFinTileBlocks = 8;

Error EnqueueBatch(pirg_server* S, Slot& slot, int nq, bool timed, bool host_io) {
  Workspace& W = slot.ws;
  hipStream_t st = slot.stream;
  const int d = S->d;
  const int lt = S->ell * S->t;
  unsigned long long* margin = S->margin.as<unsigned long long>();
  // Stage boundaries: eager runs record them; under capture the two inner
  // boundaries become event-record nodes of the graph (RunBatch records the
  // outer two around each replay), so replays report stage times too.
  auto mark = [&](hipEvent_t e) -> hipError_t {
    if (timed) return hipEventRecord(e, st);
    if (e == slot.ev[0] || e == slot.ev[3]) return hipSuccess;
    return hipEventRecordWithFlags(e, st, hipEventRecordExternal);
  };
  // Scan brackets: eager runs record them directly; under capture with
  // graph_scan_events they become event-record nodes of the graph.
  const bool scan_nodes = !timed && slot.graph_scan_events;
  auto kmark = [&](hipEvent_t e, hipStream_t s) -> hipError_t {
    if (timed) return hipEventRecord(e, s);
    return scan_nodes ? hipEventRecordWithFlags(e, s, hipEventRecordExternal) : hipSuccess;
  };
  PIRG_HIP(mark(slot.ev[0]));
  if (host_io) {
    PIRG_HIP(LaunchCopyWords(W.h_req.as<u64>(), W.req.as<u64>(),
                             static_cast<size_t>(nq) * RequestWords(S), st));
  }
  // DB product (pir_server.cc:215-229).
  const ScanPlan pd = PlanFor(S, true, nq);
  const ScanPlan pp = PlanFor(S, false, nq);
  // Request records (RequestWords apart): first_dim, second_dim, packing key.
  const int rw = static_cast<int>(RequestWords(S));
  const u64* rec_fd = W.req.as<u64>();
  const u64* rec_sd = W.req.as<u64>() + RecordSdOffset(S);
  const u64* rec_pk = W.req.as<u64>() + RecordPkOffset(S);
  // Pack product (pir_server.cc:237-248), y = concat(y_vec_g): its query
  // digits, then the scan.
  auto pack_digits = [&](hipStream_t s) -> Error {
    PIRG_HIP(LaunchQueryDigitsSums(rec_pk, rw, S->k * d, S->pkL, pp, W.bt_pk.as<int8_t>(),
                                   S->pkL.c_off, W.corr_pk.as<u64>() + nq, W.corr_pk.as<u64>(),
                                   s));
    return Ok();
  };
  auto pack_scan = [&]() -> Error {
    PIRG_HIP(kmark(slot.kev[2], st));
    PIRG_HIP(LaunchScan(S->pk_tiles.as<int8_t>(), S->pkL, pp, W.bt_pk.as<int8_t>(),
                        W.pk_partial.as<u64>(), st));
    PIRG_HIP(kmark(slot.kev[3], st));
    return Ok();
  };
  // FinalizeMatrixPack's inner products for every (shard, i)
  // (pir_server.cc:256-279), from the request's y_vec_h alone. A
  // local-finalize server (vPIR, vpir_server.cc:148-160) adds none.
  auto finalize_ip = [&](hipStream_t s) -> Error {
    if (S->local_finalize) return Ok();
    PIRG_HIP(LaunchSpectra(rec_pk + static_cast<size_t>(S->k) * d, nq * S->k, S->k, rw,
                           S->fin_v, S->F, S->tw_fwd.as<double2>(), W.yh_spec.as<double2>(),
                           s));
    const size_t n_fin_cap = static_cast<size_t>(W.cap) * lt;
    void* const* tab = W.fin_tab.as<void* const>();
    // Many (shard, i) blocks per request (vPIR's 256 shards): tiles of 8
    // requests x 8 blocks share their y_vec_h and t_vec_h spectra through one
    // L2 (6 MiB per 128 jobs at d = 2048, against 16.5 MiB for 128 jobs of
    // one request), and a single request's jobs spread over every XCD.
    const int tile_r = std::min(nq, kFinTileRequests);
    const int tile_s = lt >= 32 ? kFinTileBlocks : 0;
    PIRG_HIP(LaunchIp(reinterpret_cast<const double2* const*>(tab),
                      reinterpret_cast<const double2* const*>(tab + n_fin_cap),
                      reinterpret_cast<u64* const*>(tab + 2 * n_fin_cap),
                      nq * lt, lt, S->k, S->fin_u.num_chunks, S->fin_v.num_chunks,
                      S->fin_pairs, S->F, S->tw_bwd.as<double2>(), margin, s,
                      W.mac_acc.as<double2>(), tile_r, tile_s));
    return Ok();
  };
  // EvalPoly's RGSW spectra (pir_server.cc:126-150), from the request alone.
  auto second_dim_spectra = [&](hipStream_t s) -> Error {
    PIRG_HIP(LaunchSpectra(rec_sd, nq * 2 * S->e, 2 * S->e, rw, S->ext_v, S->F,
                           S->tw_fwd.as<double2>(), W.sdb_spec.as<double2>(), s));
    return Ok();
  };
  // While the DB scan is bound by HBM (a NarrowScan plan), the
  // latency-bound kernels that need only the request (the packing key's
  // digits, FinalizeMatrixPack's inner products, the second dimension's
  // spectra) run beside it on the side stream, forked from and joined back
  // into the slot's stream; the pack scan, bound by HBM too, follows the DB
  // scan. An MFMA-bound scan keeps every CU busy: then all of it stays in line.
  // The side branch forks here (it needs only the request records) but its
  // kernels are issued after the DB scan's: issued first, its seven small
  // launches ran before the DB scan started, ~80 us at one request.
  const bool fork = !pd.wide && S->fork_branches;
  if (fork) PIRG_HIP(hipEventRecord(slot.fork_ev, st));
  // The DB product's offset correction (the query's column sums) is read only
  // by FinalizeCombine: with a side branch it is computed there, and the DB
  // scan follows its query digits at once.
  const bool sums_on_side = fork && !pd.digit_major;
  if (sums_on_side) {
    PIRG_HIP(LaunchQueryDigits(rec_fd, rw, S->K, S->dbL, pd, W.bt_db.as<int8_t>(), st));
  } else {
    PIRG_HIP(LaunchQueryDigitsSums(rec_fd, rw, S->K, S->dbL, pd, W.bt_db.as<int8_t>(),
                                   S->dbL.c_off, W.corr_db.as<u64>() + nq, W.corr_db.as<u64>(),
                                   st));
  }
  // Two MFMA-bound scans of the kinds LaunchScanPair fuses run as one launch
  // (the pack product's query digits first) when the DB product takes few
  // rounds of workgroups: the pack product's then fill the CUs that its last,
  // partial round leaves idle (64 queries at n = 2^20, d = 4096: 2.7 rounds,
  // 4.21 -> 4.11 ms per batch). Over many rounds the idle CUs only let the
  // others clock higher (256 queries: 10.7 rounds, 14.57 against 14.62 ms),
  // so those batches keep the two launches. The scan brackets of a fused
  // batch both cover the one launch.
  const bool fused = !fork && FuseScans(S, nq, pd, pp);
  if (fused) {
    Error er = pack_digits(st);
    if (!er.ok()) return er;
    PIRG_HIP(kmark(slot.kev[0], st));
    PIRG_HIP(LaunchScanPair(S->db_tiles.as<int8_t>(), S->dbL, pd, W.bt_db.as<int8_t>(),
                            W.db_partial.as<u64>(), S->pk_tiles.as<int8_t>(), S->pkL, pp,
                            W.bt_pk.as<int8_t>(), W.pk_partial.as<u64>(), st));
    PIRG_HIP(kmark(slot.kev[1], st));
    PIRG_HIP(kmark(slot.kev[2], st));
    PIRG_HIP(kmark(slot.kev[3], st));
  } else {
    PIRG_HIP(kmark(slot.kev[0], st));
    PIRG_HIP(LaunchScan(S->db_tiles.as<int8_t>(), S->dbL, pd, W.bt_db.as<int8_t>(),
                        W.db_partial.as<u64>(), st));
    PIRG_HIP(kmark(slot.kev[1], st));
  }
  if (fork) {
    PIRG_HIP(hipStreamWaitEvent(slot.side, slot.fork_ev, 0));
    Error er = pack_digits(slot.side);
    if (!er.ok()) return er;
    PIRG_HIP(hipEventRecord(slot.pk_join_ev, slot.side));
    if (sums_on_side) {
      PIRG_HIP(LaunchVectorSums(rec_fd, nq, rw, S->K, S->dbL.c_off, W.corr_db.as<u64>() + nq,
                                W.corr_db.as<u64>(), slot.side));
    }
    er = finalize_ip(slot.side);
    if (er.ok() && S->t > 1) er = second_dim_spectra(slot.side);
    if (!er.ok()) return er;
    PIRG_HIP(hipEventRecord(slot.join_ev, slot.side));
  }
  PIRG_HIP(mark(slot.ev[1]));
  {
    Error er;
    if (fused) {
      er = finalize_ip(st);
    } else if (fork) {
      PIRG_HIP(hipStreamWaitEvent(st, slot.pk_join_ev, 0));
      er = pack_scan();
      if (er.ok()) PIRG_HIP(hipStreamWaitEvent(st, slot.join_ev, 0));
    } else {
      er = pack_digits(st);
      if (er.ok()) er = pack_scan();
      if (er.ok()) er = finalize_ip(st);
    }
    if (!er.ok()) return er;
  }
  FinalizeArgs FA{};
  FA.db_partial = W.db_partial.as<u64>();
  FA.ksplit_db = pd.ksplit;
  FA.corr_db = W.corr_db.as<u64>();
  FA.pk_partial = W.pk_partial.as<u64>();
  FA.ksplit_pk = pp.ksplit;
  FA.corr_pk = W.corr_pk.as<u64>();
  FA.fin_partial = W.fin_partial.as<u64>();
  FA.npairs = S->local_finalize ? 0 : S->fin_pairs.count;
  FA.a_tilde = S->a_tilde.as<u64>();
  FA.ct_b = W.ct_b.as<u64>();
  FA.st_a = W.st_a.as<u64>();
  FA.st_b = W.st_b.as<u64>();
  FA.out = W.out.as<u64>();
  FA.nq = nq;
  FA.M = S->M;
  FA.d = d;
  FA.t = S->t;
  FA.ell = S->ell;
  FA.log_q = S->log_q;
  FA.log_q1 = S->P.log_q1;
  FA.log_q2 = S->P.log_q2;
  FA.cmask = S->cmask;
  PIRG_HIP(LaunchFinalizeCombine(FA, st));
  PIRG_HIP(mark(slot.ev[2]));
  // EvalPoly (pir_server.cc:126-150) and the modulus switch.
  if (S->t > 1) {
    if (!fork) {
      Error er = second_dim_spectra(st);
      if (!er.ok()) return er;
    }
    const size_t n_ext_cap = static_cast<size_t>(W.cap) * S->ell * 2;
    void* const* et = W.ext_tab.as<void* const>();
    const int d_n = S->N;
    const int nu_e = S->ext_u.num_chunks;
    // Requests are independent, so the chain of half the batch's requests
    // runs beside the other half's: one half's digit transforms (FFT-bound)
    // then overlap the other's inner products (bound by their loads), where
    // in line every kernel waits for the one before. Each half must keep its
    // inner products off the shared accumulation workspace of small
    // launches (more than kIpSplitMaxJobs jobs).
    const int half = nq / 2;
    const bool split = S->horner_split &&
                       static_cast<long long>(half) * S->ell * 2 * S->ext_pairs.count >
                           kIpSplitMaxJobs;
    const int nparts = split ? 2 : 1;
    if (split) {
      PIRG_HIP(hipEventRecord(slot.hfork_ev, st));
      PIRG_HIP(hipStreamWaitEvent(slot.side, slot.hfork_ev, 0));
    }
    for (int part = 0; part < nparts; ++part) {
      hipStream_t ps = part == 0 ? st : slot.side;
      const int q0 = part == 0 ? 0 : half;
      const int pq = split ? (part == 0 ? half : nq - half) : nq;
      const size_t qo = static_cast<size_t>(q0) * S->ell;  // first (request, shard)
      const size_t to = qo * 2;                            // its first output
      for (int i = S->t - 2; i >= 0; --i) {
        DigitsArgs DA{};
        DA.st_a = W.st_a.as<u64>() + qo * d;
        DA.st_b = W.st_b.as<u64>() + qo * d;
        DA.nq = pq;
        DA.ell = S->ell;
        DA.d = d;
        DA.e = S->e;
        DA.log_q = S->log_q;
        DA.log_digit = S->P.eval_log_digit;
        DA.rule = S->ext_u;
        DA.dspec = W.d_spec.as<double2>() + qo * 2 * S->e * nu_e * static_cast<size_t>(d_n);
        // After the first step the digits kernel forms the state itself from
        // the previous step's partials; only the last step runs ExtCombine
        // (which also switches the modulus).
        DA.prev_step = i == S->t - 2 ? -1 : i + 1;
        DA.ext_partial = W.ext_partial.as<u64>() + to * S->ext_pairs.count * d;
        DA.npairs = S->ext_pairs.count;
        DA.t = S->t;
        DA.a_tilde = S->a_tilde.as<u64>();
        DA.ct_b = W.ct_b.as<u64>() + qo * S->t * d;
        PIRG_HIP(LaunchDigitsSpectra(DA, S->F, S->tw_fwd.as<double2>(), ps));
        PIRG_HIP(LaunchIp(reinterpret_cast<const double2* const*>(et + to),
                          reinterpret_cast<const double2* const*>(et + n_ext_cap + to),
                          reinterpret_cast<u64* const*>(et + 2 * n_ext_cap + to),
                          pq * S->ell * 2, S->ell * 2, 2 * S->e, S->ext_u.num_chunks,
                          S->ext_v.num_chunks, S->ext_pairs, S->F, S->tw_bwd.as<double2>(),
                          margin, ps, W.mac_acc.as<double2>()));
        if (i == 0) {
          ExtCombineArgs EA{};
          EA.ext_partial = W.ext_partial.as<u64>() + to * S->ext_pairs.count * d;
          EA.npairs = S->ext_pairs.count;
          EA.a_tilde = S->a_tilde.as<u64>();
          EA.ct_b = W.ct_b.as<u64>() + qo * S->t * d;
          EA.st_a = W.st_a.as<u64>() + qo * d;
          EA.st_b = W.st_b.as<u64>() + qo * d;
          EA.out = W.out.as<u64>() + to * d;
          EA.nq = pq;
          EA.ell = S->ell;
          EA.t = S->t;
          EA.d = d;
          EA.step_i = i;
          EA.log_q = S->log_q;
          EA.log_q1 = S->P.log_q1;
          EA.log_q2 = S->P.log_q2;
          EA.cmask = S->cmask;
          PIRG_HIP(LaunchExtCombine(EA, ps));
        }
      }
    }
    if (split) {
      PIRG_HIP(hipEventRecord(slot.hjoin_ev, slot.side));
      PIRG_HIP(hipStreamWaitEvent(st, slot.hjoin_ev, 0));
    }
  }
  if (host_io) {
    PIRG_HIP(LaunchCopyWords(W.out.as<u64>(), W.h_out.as<u64>(),
                             static_cast<size_t>(nq) * ResponseWords(S), st));
  }
  PIRG_HIP(mark(slot.ev[3]));
  return Ok();
}

}  // namespace

// Graphs of a slot, by batch size and whether they move their own host I/O.
int GraphKey(int nq, bool host_io) { return 2 * nq + (host_io ? 1 : 0); }

Error PrepareGraph(pirg_server* S, Slot& slot, int nq, bool host_io) {
  if (nq <= 0 || nq > slot.ws.cap) return Invalid("batch size exceeds the slot's workspace.");
  if (slot.graphs.count(GraphKey(nq, host_io)) != 0) return Ok();
  PIRG_HIP(hipStreamBeginCapture(slot.stream, hipStreamCaptureModeThreadLocal));
  Error er = EnqueueBatch(S, slot, nq, false, host_io);
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(slot.stream, &graph);
  if (!er.ok()) {
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    return er;
  }
  PIRG_HIP(ce);
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  PIRG_HIP(ie);
  slot.graphs.emplace(GraphKey(nq, host_io), exec);
  return Ok();
}

Error RunBatch(pirg_server* S, Slot& slot, int nq, bool use_graph, bool host_io) {
  if (nq <= 0 || nq > slot.ws.cap) return Invalid("batch size exceeds the slot's workspace.");
  if (!use_graph) {
    Error er = EnqueueBatch(S, slot, nq, true, host_io);
    if (!er.ok()) return er;
    slot.timing_pending = true;
    slot.graph_run = false;
    return Ok();
  }
  Error pe = PrepareGraph(S, slot, nq, host_io);
  if (!pe.ok()) return pe;
  auto it = slot.graphs.find(GraphKey(nq, host_io));
  PIRG_HIP(hipEventRecord(slot.ev[0], slot.stream));
  PIRG_HIP(hipGraphLaunch(it->second, slot.stream));
  PIRG_HIP(hipEventRecord(slot.ev[3], slot.stream));
  slot.timing_pending = true;
  slot.graph_run = true;
  return Ok();
}

void CollectTiming(pirg_server* S, Slot& slot) {
  if (!slot.timing_pending) return;
  if (hipEventSynchronize(slot.ev[3]) != hipSuccess) return;
  slot.timing_pending = false;
  if (slot.graph_run) {
    // A replay's stage boundaries: ev[0] / ev[3] around the launch, ev[1] /
    // ev[2] nodes inside the graph; its scan brackets are nodes too when it
    // was captured with graph_scan_events.
    float a = 0, b = 0, c = 0;
    if (hipEventElapsedTime(&a, slot.ev[0], slot.ev[1]) != hipSuccess ||
        hipEventElapsedTime(&b, slot.ev[1], slot.ev[2]) != hipSuccess ||
        hipEventElapsedTime(&c, slot.ev[2], slot.ev[3]) != hipSuccess) {
      return;
    }
    float ka = 0, kb = 0;
    const bool scans = slot.graph_scan_events &&
                       hipEventElapsedTime(&ka, slot.kev[0], slot.kev[1]) == hipSuccess &&
                       hipEventElapsedTime(&kb, slot.kev[2], slot.kev[3]) == hipSuccess;
    std::lock_guard<std::mutex> lock(S->mu);
    S->db_ms += a;
    S->pack_ms += b;
    S->eval_ms += c;
    if (scans) {
      S->db_scan_ms += ka;
      S->pk_scan_ms += kb;
      S->scans += 1;
    }
    return;
  }
  float a = 0, b = 0, c = 0;
  (void)hipEventElapsedTime(&a, slot.ev[0], slot.ev[1]);
  (void)hipEventElapsedTime(&b, slot.ev[1], slot.ev[2]);
  (void)hipEventElapsedTime(&c, slot.ev[2], slot.ev[3]);
  float ka = 0, kb = 0;
  (void)hipEventElapsedTime(&ka, slot.kev[0], slot.kev[1]);
  (void)hipEventElapsedTime(&kb, slot.kev[2], slot.kev[3]);
  std::lock_guard<std::mutex> lock(S->mu);
  S->db_ms += a;
  S->pack_ms += b;
  S->eval_ms += c;
  S->db_scan_ms += ka;
  S->pk_scan_ms += kb;
  S->scans += 1;
}

Error EnqueueFetch(pirg_server* S, Slot& slot, int nq) {
  Workspace& W = slot.ws;
  PIRG_HIP(hipMemcpyAsync(W.h_out.p, W.out.p, static_cast<size_t>(nq) * ResponseWords(S) * 8,
                          hipMemcpyDeviceToHost, slot.stream));
  return Ok();
}

namespace {

// Waits for the batch and copies its responses out of the pinned output
// buffer; `copy` first enqueues the device-to-host copy (a host_io batch
// has written the buffer itself).
Error FetchOutputs(pirg_server* S, Slot& slot, int nq, uint64_t* const* outs,
                   const std::vector<int>* idx, bool copy = true) {
  if (copy) {
    Error er = EnqueueFetch(S, slot, nq);
    if (!er.ok()) return er;
  }
  PIRG_HIP(hipStreamSynchronize(slot.stream));
  CollectTiming(S, slot);
  const size_t per = ResponseWords(S);
  for (int i = 0; i < nq; ++i) {
    const int dst = idx ? (*idx)[i] : i;
    std::memcpy(outs[dst], slot.ws.h_out.as<u64>() + static_cast<size_t>(i) * per, per * 8);
  }
  return Ok();
}

// pir_server.cc:52-124 validation (Create).
Error ValidateView(const pirg_server* S, const pirg_preprocessed_view* v) {
  if (v->num_db_matrices != S->ell) return Invalid("db_matrices size mismatch.");
  const int db_cols = v->db_cols[0];
  if (db_cols < S->r) return Invalid("DB matrix must have at least N/t columns.");
  for (int k = 0; k < S->ell; ++k) {
    if (v->db_rows[k] != S->t * S->d) return Invalid("DB matrix must have td rows.");
    if (v->db_cols[k] != db_cols) {
      return Invalid("All DB matrices must have the same number of columns.");
    }
  }
  if (v->num_outputs != S->ell) return Invalid("preprocessed_outputs size mismatch.");
  for (int k = 0; k < S->ell; ++k) {
    if (v->outputs_len[k] != S->t) return Invalid("Preprocessed outputs must be of length t.");
  }
  if (S->t > 1 && v->second_dim_a_len != 2 * S->e) {
    return Invalid(
        "second_dim_query_a must have length 2 * PolyEvalGadgetParams().num_digits.");
  }
  const int per = S->P.db_bytes == 2 ? 4 : 8;
  if (db_cols % per != 0) {
    return Invalid("DB columns must be a multiple of " + std::to_string(per) +
                   " in the condensed layout.");
  }
  return Ok();
}

// Overrides of the default scan plans, for tests and tuning: PIRG_WIDE_MIN
// sets the batch size from which the DB and pack products use the MFMA-bound
// WideScan, PIRG_DIGIT_MAJOR=0 keeps WideScan on the query-major B layout and
// PIRG_TALL=0 keeps digit-major scans on the 8-strip kernel (tests cover every
// kernel at the same batch sizes with them); PIRG_HORNER_SPLIT=0 runs
// EvalPoly's Horner chain for the whole batch in line.
void ApplyPlanOverrides(pirg_server* S) {
  if (const char* env = std::getenv("PIRG_WIDE_MIN")) {
    const int v = std::atoi(env);
    if (v > 0) S->wide_min = v;
  }
  if (const char* env = std::getenv("PIRG_DIGIT_MAJOR")) S->digit_major = std::atoi(env) != 0;
  if (const char* env = std::getenv("PIRG_TALL")) S->tall = std::atoi(env) != 0;
  if (const char* env = std::getenv("PIRG_HORNER_SPLIT")) S->horner_split = std::atoi(env) != 0;
  if (const char* env = std::getenv("PIRG_FUSE_SCANS")) S->fuse_scans = std::atoi(env) != 0;
  if (const char* env = std::getenv("PIRG_PACK_STREAM")) {
    const int v = std::atoi(env);
    if (v >= 0 && v <= 2) S->pack_stream = v;
  }
}

// Matrix<uint64_t> pack blocks (PirServer<Db, uint64_t, uint64_t>,
// pir_server.cc:806-818) hold BuildPreprocessMatrix's values
// (lwes_to_rlwe.cc:210-222): NTT outputs over Z_{2^64}, minus 2^log_q when at
// least 2^(log_q-1), so an entry equals a small integer mod 2^log_q but not
// necessarily mod 2^64. Only the product mod 2^log_q is read (FinalizeMatrixPack
// reduces b + b_part), so each entry is reduced mod 2^log_q, centered, and
// narrowed to the int32_t form (whose condensed words are the row-major int32
// values); the preprocessing's bound |x| < 2^(pack_log_digit + log d - 1)
// keeps the centered values inside [-2^31, 2^31).
Error NarrowU64Pack(const pirg_server* S, const std::vector<const void*>& blocks,
                    std::vector<std::vector<int32_t>>* store, std::vector<const void*>* out) {
  const size_t n = static_cast<size_t>(S->d) * S->k * S->d;
  store->assign(blocks.size(), {});
  out->assign(blocks.size(), nullptr);
  for (size_t b = 0; b < blocks.size(); ++b) {
    if (blocks[b] == nullptr) return Invalid("pack block " + std::to_string(b) + " is null.");
    const uint64_t* src = static_cast<const uint64_t*>(blocks[b]);
    std::vector<int32_t>& dst = (*store)[b];
    dst.resize(n);
    const int lq = S->log_q;
    for (size_t x = 0; x < n; ++x) {
      int64_t v = static_cast<int64_t>(src[x]);
      if (lq < 64) {
        const u64 r = src[x] & S->qmask;
        v = r >= (1ULL << (lq - 1)) ? static_cast<int64_t>(r) - static_cast<int64_t>(1ULL << lq)
                                    : static_cast<int64_t>(r);
      }
      if (v < INT32_MIN || v > INT32_MAX) {
        return Invalid("pack matrix entry " + std::to_string(x) + " of block " +
                       std::to_string(b) +
                       " lies outside [-2^31, 2^31) mod q; the GPU path holds uint64_t "
                       "pack matrices as int32.");
      }
      dst[x] = static_cast<int32_t>(v);
    }
    (*out)[b] = dst.data();
  }
  return Ok();
}

Error CreateFromParts(const pirg_params* params, const pirg_preprocessed_view* v,
                      std::unique_ptr<pirg_server>* out) {
  std::unique_ptr<pirg_server> S(new pirg_server());
  Error er = DeriveServer(*params, S.get());
  if (!er.ok()) return er;
  ApplyPlanOverrides(S.get());
  er = ValidateView(S.get(), v);
  if (!er.ok()) return er;
  const int lt = S->ell * S->t;
  std::vector<const void*> pk(v->pack, v->pack + lt);
  std::vector<std::vector<int32_t>> narrowed;
  if (S->P.mat_bytes == 8) {
    std::vector<const void*> pk32;
    er = NarrowU64Pack(S.get(), pk, &narrowed, &pk32);
    if (!er.ok()) return er;
    pk = pk32;
    S->P.mat_bytes = 4;
  }
  er = CheckDevice();
  if (!er.ok()) return er;
  std::vector<const uint64_t*> db(v->db_cdata, v->db_cdata + S->ell);
  std::vector<const uint64_t*> at(v->a_tilde, v->a_tilde + lt);
  std::vector<const uint64_t*> th(v->t_vec_h, v->t_vec_h + lt);
  er = BuildResident(S.get(), db, v->db_cols[0], pk, at, th, v->second_dim_a);
  if (!er.ok()) return er;
  *out = std::move(S);
  return Ok();
}

}  // namespace

Error CreateFromView(const pirg_params* params, const pirg_preprocessed_view* v,
                     std::unique_ptr<pirg_server>* out) {
  return CreateFromParts(params, v, out);
}

Error CreateFromImage(const pirg_params* params, const uint64_t* db_cdata, int db_cols,
                      const PackedPack& pack, const uint64_t* a_tilde, const uint64_t* t_vec_h,
                      const uint64_t* second_dim_a, std::unique_ptr<pirg_server>* out) {
  std::unique_ptr<pirg_server> S(new pirg_server());
  Error er = DeriveServer(*params, S.get());
  if (!er.ok()) return er;
  ApplyPlanOverrides(S.get());
  if (S->P.mat_bytes == 8) {
    // BitUnpackSigned sign-extends each entry from pack.bits bits: int32
    // holds it whenever pack.bits <= 32.
    if (pack.bits > 32) {
      return Invalid("image pack entries are wider than 32 bits; the GPU path holds "
                     "uint64_t pack matrices as int32.");
    }
    S->P.mat_bytes = 4;
  }
  // The image holds the combined matrices: view them as ell shards.
  const int ell = S->ell, t = S->t, d = S->d, lt = ell * t;
  std::vector<int> rows(ell, t * d), cols(ell, db_cols), olen(ell, t);
  const size_t wpr = S->P.db_bytes == 2 ? db_cols / 4 : db_cols / 8;
  std::vector<const uint64_t*> db(ell);
  for (int s = 0; s < ell; ++s) db[s] = db_cdata + static_cast<size_t>(s) * t * d * wpr;
  pirg_preprocessed_view v{};
  v.num_db_matrices = ell;
  v.db_cdata = db.data();
  v.db_rows = rows.data();
  v.db_cols = cols.data();
  v.num_outputs = ell;
  v.outputs_len = olen.data();
  v.second_dim_a_len = t > 1 ? 2 * S->e : 0;
  er = ValidateView(S.get(), &v);
  if (!er.ok()) return er;
  er = CheckDevice();
  if (!er.ok()) return er;
  std::vector<const uint64_t*> at(lt), th(lt);
  for (int oi = 0; oi < lt; ++oi) {
    at[oi] = a_tilde + static_cast<size_t>(oi) * d;
    th[oi] = t_vec_h + static_cast<size_t>(oi) * S->k * d;
  }
  er = BuildResident(S.get(), db, db_cols, {}, at, th, second_dim_a, &pack);
  if (!er.ok()) return er;
  *out = std::move(S);
  return Ok();
}

Error CreateFromDevice(const pirg_params* params, const uint64_t* db_cdata, int db_cols,
                       const void* pack, const uint64_t* a_tilde, const uint64_t* t_vec_h,
                       const uint64_t* second_dim_a, std::unique_ptr<pirg_server>* out) {
  std::unique_ptr<pirg_server> S(new pirg_server());
  Error er = DeriveServer(*params, S.get());
  if (!er.ok()) return er;
  ApplyPlanOverrides(S.get());
  const int ell = S->ell, t = S->t, d = S->d, lt = ell * t;
  std::vector<int> rows(ell, t * d), cols(ell, db_cols), olen(ell, t);
  const size_t wpr = S->P.db_bytes == 2 ? db_cols / 4 : db_cols / 8;
  std::vector<const uint64_t*> db(ell);
  for (int s = 0; s < ell; ++s) db[s] = db_cdata + static_cast<size_t>(s) * t * d * wpr;
  pirg_preprocessed_view v{};
  v.num_db_matrices = ell;
  v.db_cdata = db.data();
  v.db_rows = rows.data();
  v.db_cols = cols.data();
  v.num_outputs = ell;
  v.outputs_len = olen.data();
  v.second_dim_a_len = t > 1 ? 2 * S->e : 0;
  er = ValidateView(S.get(), &v);
  if (!er.ok()) return er;
  er = CheckDevice();
  if (!er.ok()) return er;
  const size_t pk_rowbytes = static_cast<size_t>(S->k) * d * (S->P.mat_bytes == 4 ? 4 : 2);
  std::vector<const void*> pk(ell);
  for (int s = 0; s < ell; ++s) {
    pk[s] = static_cast<const char*>(pack) + static_cast<size_t>(s) * t * d * pk_rowbytes;
  }
  std::vector<const uint64_t*> at(lt), th(lt);
  for (int oi = 0; oi < lt; ++oi) {
    at[oi] = a_tilde + static_cast<size_t>(oi) * d;
    th[oi] = t_vec_h + static_cast<size_t>(oi) * S->k * d;
  }
  er = BuildResident(S.get(), db, db_cols, pk, at, th, second_dim_a, nullptr, true);
  if (!er.ok()) return er;
  *out = std::move(S);
  return Ok();
}

namespace {

thread_local std::string g_version;

}  // namespace
}  // namespace pirg

namespace pirg {
namespace {

Error MatrixMultiply(const uint64_t* cdata, int rows, int cols, const uint64_t* vec,
                     uint64_t* out, bool is_u16) {
  Error er = CheckDevice();
  if (!er.ok()) return er;
  PIRG_HIP(ConfigureScanKernels());
  CallStream cs;
  PIRG_HIP(cs.create());
  hipStream_t s = cs.s;
  const int planes = is_u16 ? 2 : 4;
  const u64 c_off = is_u16 ? 32896ULL : 128ULL * (1 + 256 + 65536);
  LimbLayout L = MakeLayout(rows, cols, planes, c_off);
  DevBuf tiles, src, v, bt, corr, partial;
  const size_t wpr = is_u16 ? cols / 4 : cols / 2;
  if ((is_u16 && cols % 4) || (!is_u16 && cols % 2)) return Invalid("Vector size mismatch.");
  if ((er = tiles.alloc(TileBufferBytes(L))).code) return er;
  if ((er = src.alloc(static_cast<size_t>(rows) * wpr * 8)).code) return er;
  if ((er = v.alloc(static_cast<size_t>(cols) * 8)).code) return er;
  const ScanPlan plan = MakeScanPlan(L, 1, 64, 1 << 30);
  const int ks = plan.ksplit;
  if ((er = bt.alloc(plan.btile_bytes)).code) return er;
  if ((er = corr.alloc((1 + VectorSumScratchWords(1)) * 8)).code) return er;
  if ((er = partial.alloc(static_cast<size_t>(ks) * rows * 8)).code) return er;
  PIRG_HIP(CopySync(src.p, cdata, static_cast<size_t>(rows) * wpr * 8, hipMemcpyHostToDevice, s));
  PIRG_HIP(CopySync(v.p, vec, static_cast<size_t>(cols) * 8, hipMemcpyHostToDevice, s));
  if (is_u16) {
    PIRG_HIP(LaunchDbToTiles(src.as<u64>(), 2, rows, cols, 0, rows, L, tiles.as<int8_t>(), s));
  } else {
    PIRG_HIP(LaunchPackToTiles(src.p, 4, rows, cols, 0, rows, L, tiles.as<int8_t>(), s));
  }
  PIRG_HIP(LaunchQueryDigits(v.as<u64>(), cols, cols, L, plan, bt.as<int8_t>(), s));
  PIRG_HIP(LaunchVectorSums(v.as<u64>(), 1, cols, cols, c_off, corr.as<u64>() + 1,
                            corr.as<u64>(), s));
  PIRG_HIP(LaunchScan(tiles.as<int8_t>(), L, plan, bt.as<int8_t>(), partial.as<u64>(), s));
  std::vector<u64> part(static_cast<size_t>(ks) * rows);
  u64 cc = 0;
  PIRG_HIP(CopySync(part.data(), partial.p, part.size() * 8, hipMemcpyDeviceToHost, s));
  PIRG_HIP(CopySync(&cc, corr.p, 8, hipMemcpyDeviceToHost, s));
  for (int i = 0; i < rows; ++i) {
    u64 acc = cc;
    for (int k = 0; k < ks; ++k) acc += part[static_cast<size_t>(k) * rows + i];
    out[i] = acc;
  }
  return Ok();
}

}  // namespace
}  // namespace pirg

namespace pirg {
namespace {

std::string DescribePlan(const ScanPlan& P) {
  char buf[128];
  std::snprintf(buf, sizeof(buf), "%s tiles=%d ksplit=%d",
                P.digit_major
                    ? (P.ns == 4 && P.tall != 0
                           ? (P.tall == 1 ? "wide-dm-stream-2g6" : "wide-dm-stream-2g8")
                           : (P.tall == 1 ? "wide-dm-tall" : (P.tall == 2 ? "wide-dm-stream" : "wide-dm")))
                    : (P.wide ? "wide" : "narrow"),
                P.ctg, P.ksplit);
  return buf;
}

// Times every candidate scan configuration for nq queries on the resident
// matrix and keeps the fastest.
struct TimedPlan {
  ScanPlan plan;
  double ms;
};

// Scan operands for timing: B tiles with the bit patterns of real digits
// (the MFMA clock depends on the data, and all-zero operands run faster than
// query digits), so they are filled with bytes of the resident matrix.
Error TuningOperands(pirg_server* S, const DevBuf& tiles, size_t bt_bytes, DevBuf* bt) {
  Error er = bt->alloc(bt_bytes);
  if (!er.ok()) return er;
  hipStream_t st = S->main.stream;
  const size_t fill = std::min(bt_bytes, tiles.bytes);
  PIRG_HIP(hipMemcpyAsync(bt->p, tiles.p, fill, hipMemcpyDeviceToDevice, st));
  if (fill < bt_bytes) {
    PIRG_HIP(hipMemsetAsync(static_cast<char*>(bt->p) + fill, 0x5b, bt_bytes - fill, st));
  }
  return Ok();
}

Error TuneOne(pirg_server* S, bool db, int nq, std::string* report,
              std::vector<TimedPlan>* timed) {
  const LimbLayout& L = db ? S->dbL : S->pkL;
  const DevBuf& tiles = db ? S->db_tiles : S->pk_tiles;
  const std::vector<ScanPlan> cands = ScanPlanCandidates(L, nq, S->log_q);
  size_t bt_bytes = 0, part_words = 0;
  for (const ScanPlan& c : cands) {
    bt_bytes = std::max(bt_bytes, c.btile_bytes);
    part_words = std::max(part_words, static_cast<size_t>(c.ksplit) * nq * L.rows);
  }
  DevBuf bt, part;
  Error er = TuningOperands(S, tiles, bt_bytes, &bt);
  if (!er.ok()) return er;
  er = part.alloc(part_words * 8);
  if (!er.ok()) return er;
  hipStream_t st = S->main.stream;
  double best = 1e30;
  ScanPlan best_plan = cands.front();
  char line[160];
  std::snprintf(line, sizeof(line), "%s nq=%d:", db ? "db" : "pack", nq);
  *report += line;
  for (const ScanPlan& c : cands) {
    PIRG_HIP(LaunchScan(tiles.as<int8_t>(), L, c, bt.as<int8_t>(), part.as<u64>(), st));
    PIRG_HIP(hipEventRecord(S->main.kev[0], st));
    const int reps = 3;
    for (int r = 0; r < reps; ++r) {
      PIRG_HIP(LaunchScan(tiles.as<int8_t>(), L, c, bt.as<int8_t>(), part.as<u64>(), st));
    }
    PIRG_HIP(hipEventRecord(S->main.kev[1], st));
    PIRG_HIP(hipEventSynchronize(S->main.kev[1]));
    float ms = 0;
    PIRG_HIP(hipEventElapsedTime(&ms, S->main.kev[0], S->main.kev[1]));
    ms /= reps;
    std::snprintf(line, sizeof(line), " [%s %.3f ms]", DescribePlan(c).c_str(), ms);
    *report += line;
    timed->push_back({c, ms});
    if (ms < best) {
      best = ms;
      best_plan = c;
    }
  }
  std::snprintf(line, sizeof(line), " -> %s\n", DescribePlan(best_plan).c_str());
  *report += line;
  (db ? S->tuned_db : S->tuned_pk)[nq] = best_plan;
  return Ok();
}

// After both products are tuned: times the fastest pair of candidates that
// LaunchScanPair fuses as one launch, and keeps the pair and fusion when that
// launch beats the fastest two separate ones (the pack product then fills
// the CUs the DB product's last round leaves idle).
Error TunePair(pirg_server* S, int nq, const std::vector<TimedPlan>& tdb,
               const std::vector<TimedPlan>& tpk, std::string* report) {
  S->tuned_fuse[nq] = false;
  double best_db = 1e30, best_pk = 1e30;
  for (const TimedPlan& x : tdb) best_db = std::min(best_db, x.ms);
  for (const TimedPlan& x : tpk) best_pk = std::min(best_pk, x.ms);
  const TimedPlan* fd = nullptr;
  const TimedPlan* fp = nullptr;
  for (const TimedPlan& x : tdb) {
    for (const TimedPlan& y : tpk) {
      if (!ScanPairFusable(S->dbL, x.plan, S->pkL, y.plan) ||
          ScanWorkgroups(S->dbL, x.plan) > 4LL * S->cus) {
        continue;
      }
      if (fd == nullptr || x.ms + y.ms < fd->ms + fp->ms) {
        fd = &x;
        fp = &y;
      }
    }
  }
  if (!S->fuse_scans || fd == nullptr) return Ok();
  DevBuf bt_db, bt_pk, part_db, part_pk;
  Error er = TuningOperands(S, S->db_tiles, fd->plan.btile_bytes, &bt_db);
  if (er.ok()) er = TuningOperands(S, S->pk_tiles, fp->plan.btile_bytes, &bt_pk);
  if (er.ok()) er = part_db.alloc(static_cast<size_t>(fd->plan.ksplit) * nq * S->dbL.rows * 8);
  if (er.ok()) er = part_pk.alloc(static_cast<size_t>(fp->plan.ksplit) * nq * S->pkL.rows * 8);
  if (!er.ok()) return er;
  hipStream_t st = S->main.stream;
  auto launch = [&]() {
    return LaunchScanPair(S->db_tiles.as<int8_t>(), S->dbL, fd->plan, bt_db.as<int8_t>(),
                          part_db.as<u64>(), S->pk_tiles.as<int8_t>(), S->pkL, fp->plan,
                          bt_pk.as<int8_t>(), part_pk.as<u64>(), st);
  };
  PIRG_HIP(launch());
  PIRG_HIP(hipEventRecord(S->main.kev[0], st));
  const int reps = 3;
  for (int r = 0; r < reps; ++r) PIRG_HIP(launch());
  PIRG_HIP(hipEventRecord(S->main.kev[1], st));
  PIRG_HIP(hipEventSynchronize(S->main.kev[1]));
  float ms = 0;
  PIRG_HIP(hipEventElapsedTime(&ms, S->main.kev[0], S->main.kev[1]));
  ms /= reps;
  const bool fuse = ms < best_db + best_pk;
  char line[200];
  std::snprintf(line, sizeof(line),
                "pair nq=%d: [one launch %s + %s %.3f ms] [two launches %.3f ms] -> %s\n", nq,
                DescribePlan(fd->plan).c_str(), DescribePlan(fp->plan).c_str(), ms,
                best_db + best_pk, fuse ? "one launch" : "two launches");
  *report += line;
  if (fuse) {
    S->tuned_db[nq] = fd->plan;
    S->tuned_pk[nq] = fp->plan;
    S->tuned_fuse[nq] = true;
  }
  return Ok();
}

}  // namespace
}  // namespace pirg

extern "C" {

int pirg_server_autotune(pirg_server* S, const int* widths, int nwidths, char* report,
                         size_t report_len) {
  if (S == nullptr || widths == nullptr || nwidths < 0) return PIRG_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lock(S->serve_mu);
    if (S->serve) return PIRG_FAILED_PRECONDITION;
  }
  std::string rep;
  for (int i = 0; i < nwidths; ++i) {
    if (widths[i] < 1) return PIRG_INVALID_ARGUMENT;
    std::vector<TimedPlan> tdb, tpk;
    Error er = TuneOne(S, true, widths[i], &rep, &tdb);
    if (er.ok()) er = TuneOne(S, false, widths[i], &rep, &tpk);
    if (er.ok()) er = TunePair(S, widths[i], tdb, tpk, &rep);
    if (!er.ok()) return Report(er, report, report_len);
  }
  // Plans changed: graphs and workspace sizes follow on the next batch.
  (void)hipStreamSynchronize(S->main.stream);
  S->main.DropGraphs();
  S->main.ws.cap = 0;
  if (report != nullptr && report_len > 0) std::snprintf(report, report_len, "%s", rep.c_str());
  return 0;
}

int pirg_server_create(const pirg_params* params,
                       const pirg_preprocessed_view* data, pirg_server** out,
                       char* err, size_t errlen) {
  if (params == nullptr || data == nullptr || out == nullptr) {
    return Report(Invalid("params, data and out must not be null."), err, errlen);
  }
  std::unique_ptr<pirg_server> S;
  Error er = CreateFromParts(params, data, &S);
  if (!er.ok()) return Report(er, err, errlen);
  *out = S.release();
  return 0;
}

int pirg_server_create_flat(const pirg_params* params, const uint64_t* db_cdata,
                            int db_cols, const void* pack,
                            const uint64_t* a_tilde, const uint64_t* t_vec_h,
                            const uint64_t* second_dim_a, pirg_server** out,
                            char* err, size_t errlen) {
  if (params == nullptr || out == nullptr || db_cdata == nullptr || pack == nullptr ||
      a_tilde == nullptr || t_vec_h == nullptr) {
    return Report(Invalid("params, out, db_cdata, pack, a_tilde and t_vec_h must not be null."),
                  err, errlen);
  }
  pirg_server probe;
  Error er = DeriveServer(*params, &probe);
  if (!er.ok()) return Report(er, err, errlen);
  const int ell = probe.ell, t = probe.t, d = probe.d, k = probe.k;
  const int lt = ell * t;
  const size_t wpr = params->db_bytes == 2 ? db_cols / 4 : db_cols / 8;
  const size_t pk_rowbytes = static_cast<size_t>(k) * d * params->mat_bytes;
  std::vector<const uint64_t*> dbp(ell);
  std::vector<int> rows(ell, t * d), cols(ell, db_cols), olen(ell, t);
  for (int s = 0; s < ell; ++s) dbp[s] = db_cdata + static_cast<size_t>(s) * t * d * wpr;
  std::vector<const void*> pkp(lt);
  std::vector<const uint64_t*> atp(lt), thp(lt);
  for (int oi = 0; oi < lt; ++oi) {
    pkp[oi] = static_cast<const char*>(pack) + static_cast<size_t>(oi) * d * pk_rowbytes;
    atp[oi] = a_tilde + static_cast<size_t>(oi) * d;
    thp[oi] = t_vec_h + static_cast<size_t>(oi) * k * d;
  }
  pirg_preprocessed_view v{};
  v.num_db_matrices = ell;
  v.db_cdata = dbp.data();
  v.db_rows = rows.data();
  v.db_cols = cols.data();
  v.num_outputs = ell;
  v.outputs_len = olen.data();
  v.pack = pkp.data();
  v.a_tilde = atp.data();
  v.t_vec_h = thp.data();
  v.second_dim_a_len = (t > 1) ? 2 * probe.e : 0;
  v.second_dim_a = second_dim_a;
  return pirg_server_create(params, &v, out, err, errlen);
}

void pirg_server_destroy(pirg_server* server) { delete server; }

int pirg_process(pirg_server* S, const uint64_t* first_dim, size_t first_dim_len,
                 const uint64_t* second_dim, size_t second_dim_polys,
                 const uint64_t* packing_key, size_t packing_key_polys,
                 uint64_t* out, char* err, size_t errlen) {
  pirg_request_view r{first_dim, first_dim_len, second_dim, second_dim_polys,
                      packing_key, packing_key_polys};
  int status = 0;
  uint64_t* outs[1] = {out};
  int rc = pirg_process_batch(S, 1, &r, outs, &status, err, errlen);
  return rc != 0 ? rc : status;
}

int pirg_process_batch(pirg_server* S, int nreq, const pirg_request_view* reqs,
                       uint64_t* const* outs, int* statuses, char* err,
                       size_t errlen) {
  if (S == nullptr || nreq < 0) return Report(Invalid("invalid server or batch size."), err, errlen);
  if (nreq > 0 && (reqs == nullptr || outs == nullptr || statuses == nullptr)) {
    return Report(Invalid("reqs, outs and statuses must hold nreq entries."), err, errlen);
  }
  std::vector<int> idx;
  Error first_bad;
  for (int i = 0; i < nreq; ++i) {
    Error er = ValidateRequest(S, reqs[i].first_dim_len, reqs[i].second_dim_polys,
                               reqs[i].packing_key_polys);
    statuses[i] = er.code;
    if (er.ok()) {
      idx.push_back(i);
    } else if (first_bad.ok()) {
      first_bad = er;
    }
  }
  if (!first_bad.ok()) Report(first_bad, err, errlen);
  if (idx.empty()) return 0;
  const int cap = static_cast<int>(idx.size());
  Error er = EnsureWorkspace(S, S->main, cap);
  if (!er.ok()) return Report(er, err, errlen);
  // A small batch moves its records in and its responses out with kernels
  // in line with its own: no copy-engine transfers to wait on either side.
  const bool host_io = cap <= kHostIoMax;
  if (host_io) {
    WriteRecords(S, S->main, reqs, idx);
  } else {
    er = StageRequests(S, S->main, reqs, idx);
    if (!er.ok()) return Report(er, err, errlen);
  }
  er = RunBatch(S, S->main, cap, S->use_graphs, host_io);
  if (!er.ok()) return Report(er, err, errlen);
  er = FetchOutputs(S, S->main, cap, outs, &idx, !host_io);
  if (!er.ok()) return Report(er, err, errlen);
  return 0;
}

int pirg_batch_stage(pirg_server* S, int nreq, const pirg_request_view* reqs,
                     char* err, size_t errlen) {
  if (S == nullptr || nreq <= 0 || reqs == nullptr) {
    return Report(Invalid("invalid server, batch size or requests."), err, errlen);
  }
  std::vector<int> idx(nreq);
  for (int i = 0; i < nreq; ++i) {
    Error er = ValidateRequest(S, reqs[i].first_dim_len, reqs[i].second_dim_polys,
                               reqs[i].packing_key_polys);
    if (!er.ok()) return Report(er, err, errlen);
    idx[i] = i;
  }
  Error er = EnsureWorkspace(S, S->main, nreq);
  if (!er.ok()) return Report(er, err, errlen);
  er = StageRequests(S, S->main, reqs, idx);
  if (!er.ok()) return Report(er, err, errlen);
  if (hipStreamSynchronize(S->main.stream) != hipSuccess) {
    return Report(Internal("stream synchronize failed"), err, errlen);
  }
  S->staged = nreq;
  return 0;
}

int pirg_batch_run(pirg_server* S, int nreq, char* err, size_t errlen) {
  if (S == nullptr) return Report(Invalid("invalid server."), err, errlen);
  if (nreq <= 0 || nreq > S->staged) {
    return Report(Invalid("batch size exceeds the staged requests."), err, errlen);
  }
  Error er = RunBatch(S, S->main, nreq, S->use_graphs);
  if (!er.ok()) return Report(er, err, errlen);
  if (hipStreamSynchronize(S->main.stream) != hipSuccess) {
    return Report(Internal("stream synchronize failed"), err, errlen);
  }
  CollectTiming(S, S->main);
  return 0;
}

int pirg_batch_fetch(pirg_server* S, int nreq, uint64_t* const* outs, char* err,
                     size_t errlen) {
  if (S == nullptr || outs == nullptr) return Report(Invalid("invalid server or outputs."), err, errlen);
  if (nreq <= 0 || nreq > S->staged) {
    return Report(Invalid("batch size exceeds the staged requests."), err, errlen);
  }
  Error er = FetchOutputs(S, S->main, nreq, outs, nullptr);
  if (!er.ok()) return Report(er, err, errlen);
  return 0;
}

int pirg_server_use_graphs(pirg_server* S, int enable) {
  if (S == nullptr) return PIRG_INVALID_ARGUMENT;
  S->use_graphs = enable != 0;
  return 0;
}

int pirg_server_time_graph_scans(pirg_server* S, int enable) {
  if (S == nullptr) return PIRG_INVALID_ARGUMENT;
  const bool on = enable != 0;
  if (on != S->main.graph_scan_events) {
    // Graphs captured in the other mode are rebuilt on their next replay.
    CollectTiming(S, S->main);
    if (hipStreamSynchronize(S->main.stream) != hipSuccess) return PIRG_INTERNAL;
    S->main.DropGraphs();
    S->main.graph_scan_events = on;
  }
  return 0;
}

int pirg_server_stage_times(pirg_server* S, double* db_mult_ms, double* packing_ms,
                            double* poly_eval_ms) {
  CollectTiming(S, S->main);
  std::lock_guard<std::mutex> lock(S->mu);
  if (db_mult_ms) *db_mult_ms = S->db_ms;
  if (packing_ms) *packing_ms = S->pack_ms;
  if (poly_eval_ms) *poly_eval_ms = S->eval_ms;
  return 0;
}

int pirg_server_scan_times(pirg_server* S, double* db_scan_ms, double* pack_scan_ms,
                           long long* launches, int reset) {
  CollectTiming(S, S->main);
  std::lock_guard<std::mutex> lock(S->mu);
  if (db_scan_ms) *db_scan_ms = S->db_scan_ms;
  if (pack_scan_ms) *pack_scan_ms = S->pk_scan_ms;
  if (launches) *launches = S->scans;
  if (reset) {
    S->db_scan_ms = S->pk_scan_ms = 0;
    S->scans = 0;
  }
  return 0;
}

int pirg_server_scans_fused(pirg_server* S, int batch, int* fused) {
  if (S == nullptr || fused == nullptr || batch < 1) return PIRG_INVALID_ARGUMENT;
  *fused = FuseScans(S, batch, PlanFor(S, true, batch), PlanFor(S, false, batch)) ? 1 : 0;
  return 0;
}

double pirg_server_fft_margin(pirg_server* S, int reset) {
  unsigned long long bits = 0;
  (void)hipStreamSynchronize(S->main.stream);
  (void)CopySync(&bits, S->margin.p, 8, hipMemcpyDeviceToHost, S->main.stream);
  if (reset) (void)FillSync(S->margin.p, 0, 8, S->main.stream);
  double v;
  std::memcpy(&v, &bits, 8);
  return v;
}


int pirg_matrix_multiply_u16(const uint64_t* cdata, int rows, int cols,
                             const uint64_t* vec, uint64_t* out, char* err,
                             size_t errlen) {
  Error er = MatrixMultiply(cdata, rows, cols, vec, out, true);
  return er.ok() ? 0 : Report(er, err, errlen);
}

int pirg_matrix_multiply_i32(const uint64_t* cdata, int rows, int cols,
                             const uint64_t* vec, uint64_t* out, char* err,
                             size_t errlen) {
  Error er = MatrixMultiply(cdata, rows, cols, vec, out, false);
  return er.ok() ? 0 : Report(er, err, errlen);
}

int pirg_inner_product_fft(const uint64_t* u, const uint64_t* v, int n, int d,
                           int u_bits, int v_bits, int chunk_bits,
                           int u_is_signed, uint64_t* out, char* err,
                           size_t errlen) {
  auto run = [&]() -> Error {
    if (n <= 0) return Invalid("Polynomial vectors cannot be empty.");
    if (d < 2 || d > 4096 || !IsPow2(static_cast<u64>(d))) {
      return Invalid("Polynomial lengths or context mismatch.");
    }
    ChunkRule ur{}, vr{};
    Error er = MakeRules(n, d, u_bits, v_bits, chunk_bits, u_is_signed != 0, &ur, &vr);
    if (!er.ok()) return er;
    er = CheckDevice();
    if (!er.ok()) return er;
    const int N = 2 * d;
    int log_d = 0;
    while ((1 << log_d) < d) ++log_d;
    FftDims F{d, log_d, N, log_d + 1};
    PIRG_HIP(ConfigureFftKernels(N));
    PairList pairs = MakePairs(ur, vr);
    std::vector<double2> fwd, bwd;
    Twiddles(N, &fwd, &bwd);
    DevBuf du, dv, us, vs, part, res, twf, twb, tab;
    if ((er = du.alloc(static_cast<size_t>(n) * d * 8)).code) return er;
    if ((er = dv.alloc(static_cast<size_t>(n) * d * 8)).code) return er;
    if ((er = us.alloc(static_cast<size_t>(n) * ur.num_chunks * N * sizeof(double2))).code) return er;
    if ((er = vs.alloc(static_cast<size_t>(n) * vr.num_chunks * N * sizeof(double2))).code) return er;
    if ((er = part.alloc(static_cast<size_t>(pairs.count) * d * 8)).code) return er;
    if ((er = res.alloc(static_cast<size_t>(d) * 8)).code) return er;
    if ((er = twf.alloc(fwd.size() * sizeof(double2))).code) return er;
    if ((er = twb.alloc(bwd.size() * sizeof(double2))).code) return er;
    if ((er = tab.alloc(3 * sizeof(void*))).code) return er;
    CallStream cs;
    PIRG_HIP(cs.create());
    hipStream_t s = cs.s;
    PIRG_HIP(CopySync(du.p, u, static_cast<size_t>(n) * d * 8, hipMemcpyHostToDevice, s));
    PIRG_HIP(CopySync(dv.p, v, static_cast<size_t>(n) * d * 8, hipMemcpyHostToDevice, s));
    PIRG_HIP(CopySync(twf.p, fwd.data(), fwd.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    PIRG_HIP(CopySync(twb.p, bwd.data(), bwd.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    const void* ptrs[3] = {us.p, vs.p, part.p};
    PIRG_HIP(CopySync(tab.p, ptrs, sizeof(ptrs), hipMemcpyHostToDevice, s));
    PIRG_HIP(LaunchSpectra(du.as<u64>(), n, 1, d, ur, F, twf.as<double2>(), us.as<double2>(), s));
    PIRG_HIP(LaunchSpectra(dv.as<u64>(), n, 1, d, vr, F, twf.as<double2>(), vs.as<double2>(), s));
    void* const* t = tab.as<void* const>();
    PIRG_HIP(LaunchIp(reinterpret_cast<const double2* const*>(t),
                      reinterpret_cast<const double2* const*>(t + 1),
                      reinterpret_cast<u64* const*>(t + 2), 1, 1, n,
                      ur.num_chunks, vr.num_chunks, pairs, F, twb.as<double2>(), nullptr, s));
    PIRG_HIP(LaunchSumPairs(part.as<u64>(), pairs.count, d, res.as<u64>(), s));
    PIRG_HIP(CopySync(out, res.p, static_cast<size_t>(d) * 8, hipMemcpyDeviceToHost, s));
    return Ok();
  };
  Error er = run();
  return er.ok() ? 0 : Report(er, err, errlen);
}

int pirg_fft_twiddles(int n, double* fwd, double* bwd) {
  if (n < 2 || n > (1 << 20) || (n & (n - 1)) != 0) return PIRG_INVALID_ARGUMENT;
  pirg_internal_twiddles(n, fwd, bwd);
  return 0;
}

const char* pirg_version(void) {
  g_version = "pirgpu 0.1 (gfx950 int8-limb MFMA + reference-exact FP64 FFT)";
  return g_version.c_str();
}

}  // extern "C"
