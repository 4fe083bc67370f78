// Host-side state shared by server.cpp (Create / ProcessResponse / batches)
// and serve.cpp (the serving loop).
#ifndef INSPIRE_AMD_SERVER_INTERNAL_H_
#define INSPIRE_AMD_SERVER_INTERNAL_H_

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pirgpu.h"
#include "pirg_internal.h"

namespace pirg {

struct Error {
  int code = 0;
  std::string msg;
  bool ok() const { return code == 0; }
};

inline Error Ok() { return Error{}; }
inline Error Invalid(const std::string& m) { return Error{PIRG_INVALID_ARGUMENT, m}; }
inline Error Internal(const std::string& m) { return Error{PIRG_INTERNAL, m}; }
inline Error HipErr(hipError_t e, const char* what) {
  return Error{PIRG_INTERNAL, std::string(what) + ": " + hipGetErrorString(e)};
}

int Report(const Error& e, char* err, size_t errlen);

#define PIRG_HIP(expr)                              \
  do {                                              \
    hipError_t _e = (expr);                         \
    if (_e != hipSuccess) return HipErr(_e, #expr); \
  } while (0)

// Blocking copies and fills ordered on one of the library's non-blocking
// streams. Nothing here uses the legacy default stream: HIP rejects a
// legacy-stream operation while any thread captures a graph, and the serving
// loop captures on its own thread while callers keep using the library.
inline hipError_t CopySync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                           hipStream_t s) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
  return e != hipSuccess ? e : hipStreamSynchronize(s);
}
inline hipError_t FillSync(void* dst, int value, size_t bytes, hipStream_t s) {
  const hipError_t e = hipMemsetAsync(dst, value, bytes, s);
  return e != hipSuccess ? e : hipStreamSynchronize(s);
}

// A non-blocking stream for one call of a standalone entry point.
struct CallStream {
  hipStream_t s = nullptr;
  CallStream() = default;
  CallStream(const CallStream&) = delete;
  CallStream& operator=(const CallStream&) = delete;
  ~CallStream() {
    if (s != nullptr) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
  hipError_t create() { return hipStreamCreateWithFlags(&s, hipStreamNonBlocking); }
};

// Device allocation released with its owner.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset();
  Error alloc(size_t n);
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Pinned host allocation released with its owner.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { reset(); }
  void reset();
  Error alloc(size_t n);
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Per-batch working memory, sized for `cap` requests.
struct Workspace {
  int cap = 0;
  DevBuf req;                    // staged requests, one record each (RequestWords)
  DevBuf bt_db, bt_pk;           // B-operand tiles
  DevBuf corr_db, corr_pk;
  DevBuf db_partial, pk_partial;
  DevBuf yh_spec, sdb_spec, d_spec;
  DevBuf fin_partial, ext_partial;
  DevBuf ct_b, st_a, st_b, out;
  DevBuf fin_tab, ext_tab;       // IpKernel pointer tables
  DevBuf mac_acc;                // accumulated spectra of small IpKernel launches
  HostBuf h_req, h_out;
};

// One execution context: a workspace, its stream, timing events and the
// hipGraphs captured for it (one per batch size). The server answers direct
// calls on its main slot; the serving loop owns several.
struct Slot {
  Workspace ws;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // stage boundaries
  hipEvent_t kev[4] = {nullptr, nullptr, nullptr, nullptr};  // DB / pack scans
  hipEvent_t done = nullptr;                                 // serving: batch finished
  hipEvent_t graph_done = nullptr;  // serving: the batch's kernels finished
  // The request-only kernels of a small batch (packing-key digits,
  // FinalizeMatrixPack's inner products, second-dimension spectra) run on
  // `side` beside the DB scan, forked from `stream`; the pack scan waits for
  // the digits (pk_join_ev), FinalizeCombine for the rest (join_ev).
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr, pk_join_ev = nullptr;
  // Larger batches run EvalPoly's Horner chain as two independent halves of
  // their requests, the second half on `side`.
  hipEvent_t hfork_ev = nullptr, hjoin_ev = nullptr;
  bool timing_pending = false;
  bool graph_run = false;  // the pending timing came from a graph launch
  // Graphs captured on this slot bracket the two scans with event-record
  // nodes (kev), so replays report scan-kernel times too.
  bool graph_scan_events = false;
  std::map<int, hipGraphExec_t> graphs;
  Slot() = default;
  Slot(const Slot&) = delete;
  Slot& operator=(const Slot&) = delete;
  ~Slot();
  Error Init();
  void DropGraphs();
};

class ServeLoop;

}  // namespace pirg

struct pirg_server {
  pirg_params P{};
  int device = 0;
  int cus = 0;  // compute units of the device
  int d = 0, log_d = 0, N = 0, log_N = 0, t = 0, ell = 0, k = 0, e = 0;
  int r = 0, K = 0, M = 0, log_q = 0;
  pirg::u64 qmask = 0;
  // CoeffType: width 64 (uint64_t) or 32 (uint32_t); cmask = 2^width - 1.
  int width = 64;
  pirg::u64 cmask = ~0ULL;
  pirg::FftDims F{};
  pirg::ChunkRule fin_u{}, fin_v{}, ext_u{}, ext_v{};
  pirg::PairList fin_pairs{}, ext_pairs{};
  pirg::Error deferred;  // error ProcessResponse must return (chunk validation)
  pirg::LimbLayout dbL{}, pkL{};
  // Batches of at least this many requests use the MFMA-bound WideScan.
  int wide_min = 10;
  // Device memory the serving loop's slots may take (0: 80% of the free
  // memory at start); under serve_mu.
  size_t serve_mem_limit = 0;
  // Small batches run their pack branch beside the DB scan.
  bool fork_branches = true;
  // Default plans may use the digit-major WideScan layout, and for two-plane
  // matrices its tall kernel.
  bool digit_major = true;
  bool tall = true;
  // Digit-major scans of four-plane (int32 pack) matrices: ScanPlan::tall,
  // or -1 for MakeScanPlan's default.
  int pack_stream = -1;
  // Batches with enough requests split EvalPoly's Horner chain into two
  // concurrent halves.
  bool horner_split = true;
  // MFMA-bound batches run the DB and pack scans as one launch when their
  // plans allow (LaunchScanPair).
  bool fuse_scans = true;
  // Scan configurations chosen by pirg_server_autotune, per batch size.
  std::map<int, pirg::ScanPlan> tuned_db, tuned_pk;
  // Per tuned batch size: whether the DB and pack scans run as one launch.
  std::map<int, bool> tuned_fuse;
  pirg::DevBuf db_tiles, pk_tiles, a_tilde, t_spec, sda_spec, tw_fwd, tw_bwd, margin;
  // t_vec_h ((ell*t) x k x d) and second_dim_a (2e x d, when t > 1) as given,
  // for SaveToFile (the kernels read their spectra).
  pirg::DevBuf t_raw, sda_raw;
  // The pack matrix type the caller passed (mat_bytes 8 is held as int32).
  int mat_src_bytes = 4;
  pirg::Slot main;
  int staged = 0;
  // pirg_process / pirg_process_batch / pirg_batch_run replay one hipGraph
  // per batch size instead of launching eagerly (the serving loop always
  // does). On by default; pirg_server_use_graphs(server, 0) turns it off.
  bool use_graphs = true;
  // FinalizeMatrixPack without the t_vec_h inner product (vPIR's local finalize).
  bool local_finalize = false;
  // Cumulative stage times (GetDbMultiplyTime / GetPackingTime /
  // GetPolyEvalTime) and scan-kernel times, guarded by `mu` (the serving
  // loop adds to them from its own thread).
  std::mutex mu;
  double db_ms = 0, pack_ms = 0, eval_ms = 0;
  double db_scan_ms = 0, pk_scan_ms = 0;
  long long scans = 0;
  // The serving loop (see serve.cpp for its lifetime rules); serve_mu guards
  // the pointer, not the loop.
  std::mutex serve_mu;
  std::shared_ptr<pirg::ServeLoop> serve;
  // Defined beside ServeLoop (serve.cpp), where its type is complete.
  pirg_server();
  ~pirg_server();
};

namespace pirg {

// A staged request's record: first_dim zero-padded to K words, then (when
// t > 1) the 2e second-dimension polynomials, then the 2k packing-key
// polynomials (y_vec_g, y_vec_h). Batches hold one record per request.
size_t RequestWords(const pirg_server* S);
size_t RecordSdOffset(const pirg_server* S);
size_t RecordPkOffset(const pirg_server* S);
// Writes a validated request's record to dst (RequestWords words).
void WriteRecord(const pirg_server* S, const pirg_request_view& r, u64* dst);
// Enqueues the host-to-device copies of nq records (host memory, pinned for
// an asynchronous copy) into the slot's workspace, record i at position i.
Error StageRecords(pirg_server* S, Slot& slot, const u64* const* records, int nq);
// pir_server.cc:161-194 request validation, in the reference's order.
Error ValidateRequest(const pirg_server* S, size_t fd_len, size_t sd_polys, size_t pk_polys);
// Bytes of device workspace a slot needs for `cap` requests.
size_t WorkspaceBytes(const pirg_server* S, int cap);
Error EnsureWorkspace(pirg_server* S, Slot& slot, int cap);
// Writes requests reqs[idx[i]] as records into the slot's pinned request
// buffer (several threads for large batches).
void WriteRecords(pirg_server* S, Slot& slot, const pirg_request_view* reqs,
                  const std::vector<int>& idx);
// WriteRecords, then the host-to-device copy on the slot's stream.
Error StageRequests(pirg_server* S, Slot& slot, const pirg_request_view* reqs,
                    const std::vector<int>& idx);
// Batches of at most this many requests on the direct path move their
// records in and responses out with kernels inside the batch (host_io).
constexpr int kHostIoMax = 16;
// Enqueues ProcessResponse for the nq staged requests on the slot's stream,
// eagerly (with per-stage timing) or as a replay of the slot's hipGraph for
// nq (captured on first use). host_io: the batch itself copies the records
// from the slot's pinned request buffer first and its responses into the
// pinned output buffer last (WriteRecords before, no EnqueueFetch after).
Error RunBatch(pirg_server* S, Slot& slot, int nq, bool use_graph, bool host_io = false);
// Captures and instantiates the slot's hipGraph for nq requests (and
// host_io), if it has none, without running it.
Error PrepareGraph(pirg_server* S, Slot& slot, int nq, bool host_io = false);
// Enqueues the device-to-host copy of nq responses into the slot's pinned
// output buffer.
Error EnqueueFetch(pirg_server* S, Slot& slot, int nq);
// Adds the slot's finished batch to the server's stage times.
void CollectTiming(pirg_server* S, Slot& slot);
size_t ResponseWords(const pirg_server* S);
// The scan configuration of the DB (db = true) or pack product for nq
// queries: the autotuned one when there is one, else the built-in choice.
ScanPlan PlanFor(const pirg_server* S, bool db, int nq);

Error CheckDevice();
// crypto/lwes_to_rlwe.cc:60-74 and crypto/external_product.cc:100-102.
// Sizes, chunk rules and checks of a server for params (pir_params.cc:29-68).
Error DeriveServer(const pirg_params& P, pirg_server* S);
int ChunkFor(int n_terms, int d);
// crypto/polynomial_fft.cc:90-107.
Error MakeRules(int n_terms, int d, int u_bits, int v_bits, int chunk, bool u_signed,
                ChunkRule* ur, ChunkRule* vr);
// Stage-ordered FFT twiddle tables (n - 1 complex values each).
void Twiddles(int n, std::vector<double2>* fwd, std::vector<double2>* bwd);
LimbLayout MakeLayout(int rows, int cols, int planes, u64 c_off);

// Bit-packed streams (codec.hip): value i of stream p is bits
// [bit0 + i*bits, +bits) of the little-endian words at words + p*pitch.
hipError_t LaunchBitUnpack(const u64* words, long long pitch, long long bit0, int nstreams,
                           long long len, int bits, bool sign, int out_bytes, void* out,
                           hipStream_t s);
hipError_t LaunchBitPack(const u64* vals, int nstreams, long long len, int bits, u64* words,
                         long long pitch, hipStream_t s);
long long StreamWords(long long len, int bits);

// The pack matrix of a server image: signed entries, row-major, bit-packed
// at `bits` bits (host memory).
struct PackedPack {
  const uint8_t* bytes;
  size_t nbytes;
  int bits;
};

// PirServer::Create on a view (polynomial pointers may be device pointers).
Error CreateFromView(const pirg_params* params, const pirg_preprocessed_view* v,
                     std::unique_ptr<pirg_server>* out);

// Server from image parts: the combined DB's condensed words (host), the
// bit-packed pack matrix, and the polynomials already in device memory
// (a_tilde: lt x d, t_vec_h: lt x k x d, second_dim_a: 2e x d or null).
Error CreateFromImage(const pirg_params* params, const uint64_t* db_cdata, int db_cols,
                      const PackedPack& pack, const uint64_t* a_tilde, const uint64_t* t_vec_h,
                      const uint64_t* second_dim_a, std::unique_ptr<pirg_server>* out);

// Server from material already in device memory (the combined matrices, as
// CreateFromImage takes them, with the pack matrix as int32 / int16 rows).
Error CreateFromDevice(const pirg_params* params, const uint64_t* db_cdata, int db_cols,
                       const void* pack, const uint64_t* a_tilde, const uint64_t* t_vec_h,
                       const uint64_t* second_dim_a, std::unique_ptr<pirg_server>* out);

// Uploads rows [0, rows) of a host matrix in chunks and converts them to
// tiles. `rowbytes` is the size of one host row.
template <typename Convert>
inline Error UploadTiles(const void* host, size_t rowbytes, int rows, bool host_is_device,
                  hipStream_t s, Convert conv) {
  if (host_is_device) {
    PIRG_HIP(conv(host, 0, rows));
    PIRG_HIP(hipStreamSynchronize(s));
    return Ok();
  }
  // Chunk size of the staged upload (PIRG_UPLOAD_CHUNK_BYTES overrides it so
  // tests can exercise the multi-chunk path on small matrices).
  size_t chunk_bytes = static_cast<size_t>(256) << 20;
  if (const char* env = std::getenv("PIRG_UPLOAD_CHUNK_BYTES")) {
    const long long v = std::atoll(env);
    if (v > 0) chunk_bytes = static_cast<size_t>(v);
  }
  int rows_per = static_cast<int>(std::max<size_t>(16, chunk_bytes / std::max<size_t>(rowbytes, 1)));
  rows_per = (rows_per / 16) * 16;
  if (rows_per < 16) rows_per = 16;
  DevBuf stage;
  Error er = stage.alloc(static_cast<size_t>(std::min(rows_per, rows)) * rowbytes);
  if (!er.ok()) return er;
  for (int r0 = 0; r0 < rows; r0 += rows_per) {
    const int nr = std::min(rows_per, rows - r0);
    PIRG_HIP(hipMemcpyAsync(stage.p, static_cast<const char*>(host) + static_cast<size_t>(r0) * rowbytes,
                            static_cast<size_t>(nr) * rowbytes, hipMemcpyHostToDevice, s));
    PIRG_HIP(conv(stage.p, r0, nr));
    PIRG_HIP(hipStreamSynchronize(s));
  }
  return Ok();
}


}  // namespace pirg

#endif  // INSPIRE_AMD_SERVER_INTERNAL_H_
