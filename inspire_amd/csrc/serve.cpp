// The serving loop: continuous batching of PirServer::ProcessResponse
// requests over a server whose DB and preprocessed material stay resident in
// HBM.
//
// Callers submit requests from any thread; a worker thread keeps several
// execution slots (workspace + stream + captured hipGraphs) busy. Whenever a
// slot is free, every request waiting at that moment (up to the slot's batch
// capacity) is staged and launched as one batch, so requests that arrive
// while other batches run on the GPU are admitted as soon as a slot frees,
// and the batches of different slots overlap on the device. Memory bounds
// admission twice: the number of slots and their batch capacity are chosen at
// start from the free HBM, and a submit is refused (RESOURCE_EXHAUSTED) while
// `max_outstanding` requests are submitted but not yet polled.
//
// Lifetime: the server holds the loop through a shared_ptr, and every
// submit / poll / stats call works on its own copy of it, taken under the
// server's serve_mu. pirg_serve_stop detaches the loop, stops it (waking any
// thread blocked in poll) and drops its reference; the loop is destroyed when
// the last call still inside it returns.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "server_internal.h"

namespace pirg {

using Clock = std::chrono::steady_clock;

namespace {

double MsSince(Clock::time_point t0, Clock::time_point t1) {
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

}  // namespace

// A queued request, as its record (RequestWords words: first dimension,
// second dimension, packing key): the caller's own record in pinned memory
// (ext), one of the loop's pinned record pool (rec >= 0) or, when the pool
// is exhausted, a heap buffer.
struct ServeRequest {
  uint64_t tag = 0;
  Clock::time_point submitted;
  const u64* ext = nullptr;
  int rec = -1;
  std::vector<u64> data;
};

// Pinned host blocks handed out by pirg_host_alloc, by start address.
std::mutex g_host_mu;
std::map<uintptr_t, size_t> g_host_blocks;

// Whether [p, p + bytes) lies inside one block from pirg_host_alloc.
bool InHostBlock(const void* p, size_t bytes) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  std::lock_guard<std::mutex> lock(g_host_mu);
  auto it = g_host_blocks.upper_bound(a);
  if (it == g_host_blocks.begin()) return false;
  --it;
  return a - it->first <= it->second && bytes <= it->second - (a - it->first);
}

struct ServeResult {
  uint64_t tag = 0;
  int status = 0;
  std::string msg;
  std::vector<u64> out;
  double latency_ms = 0;
};

struct InFlight {
  Slot* slot = nullptr;
  std::vector<ServeRequest> reqs;
  std::vector<int> recs;  // pool records to free when the batch retires
};

// Runs fn(lo, hi) over [0, n) on up to max_threads threads, one per 8 MB
// of the `bytes` the ranges copy (one thread's memcpy moves about 10 GB/s,
// slower than the link and the GPU work a large batch feeds).
template <typename Fn>
void CopyInParallel(size_t n, size_t bytes, size_t max_threads, Fn fn) {
  const size_t threads = std::min<size_t>(
      {std::max<size_t>(1, std::thread::hardware_concurrency()), max_threads, n,
       bytes / (size_t(8) << 20) + 1});
  if (threads <= 1) {
    fn(size_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  const size_t per = (n + threads - 1) / threads;
  for (size_t t = 1; t < threads; ++t) {
    const size_t lo = std::min(n, t * per), hi = std::min(n, lo + per);
    if (lo < hi) pool.emplace_back(fn, lo, hi);
  }
  fn(size_t(0), std::min(n, per));
  for (auto& th : pool) th.join();
}

class ServeLoop {
 public:
  ServeLoop(pirg_server* S, int max_batch, int nslots, int max_outstanding)
      : S_(S), max_batch_(max_batch), nslots_(nslots), max_outstanding_(max_outstanding) {}

  ~ServeLoop() { Stop(); }

  Error Start() {
    // Admission by free HBM: as many slots of max_batch requests as fit in
    // 80% of the free device memory, or in the server's serving memory
    // limit when that is lower (fewer slots, then smaller batches).
    size_t free_b = 0, total_b = 0;
    PIRG_HIP(hipMemGetInfo(&free_b, &total_b));
    double budget = 0.8 * static_cast<double>(free_b);
    if (S_->serve_mem_limit > 0) {
      budget = std::min(budget, static_cast<double>(S_->serve_mem_limit));
    }
    while (max_batch_ > 1 &&
           static_cast<double>(WorkspaceBytes(S_, max_batch_)) * nslots_ > budget) {
      if (nslots_ > 2) {
        --nslots_;
      } else {
        max_batch_ /= 2;
      }
    }
    if (static_cast<double>(WorkspaceBytes(S_, max_batch_)) * nslots_ > budget) {
      return Error{PIRG_RESOURCE_EXHAUSTED,
                   "not enough free device memory for a serving slot (" +
                       std::to_string(nslots_) + " x " +
                       std::to_string(WorkspaceBytes(S_, max_batch_)) + " bytes, budget " +
                       std::to_string(static_cast<unsigned long long>(budget)) + ")"};
    }
    // The pinned record pool: submits write requests here and batches copy
    // them to HBM from here, so a request crosses host memory once. A record
    // for every request that may be outstanding, within 4 GiB of pinned
    // memory (at least the batches in flight plus one being gathered);
    // requests beyond it wait in heap records.
    rw_ = RequestWords(S_);
    const size_t cap_records = std::max<size_t>(
        static_cast<size_t>(max_batch_) * (nslots_ + 1), (size_t(4) << 30) / (rw_ * 8));
    const size_t nrec = std::min(static_cast<size_t>(max_outstanding_), cap_records);
    {
      Error pe = records_.alloc(nrec * rw_ * 8);
      if (!pe.ok()) return pe;
      free_recs_.resize(nrec);
      for (size_t i = 0; i < nrec; ++i) free_recs_[i] = static_cast<int>(nrec - 1 - i);
    }
    for (int i = 0; i < nslots_; ++i) {
      slots_.emplace_back(new Slot());
      Error er = slots_.back()->Init();
      if (!er.ok()) return er;
      er = EnsureWorkspace(S_, *slots_.back(), max_batch_);
      if (!er.ok()) return er;
      // Every batch size the loop launches has its graph before the first
      // request arrives: a capture in the loop would stall every batch
      // behind it.
      for (int n = 1; n <= max_batch_; n = Bucket(n) + 1) {
        er = PrepareGraph(S_, *slots_.back(), Bucket(n));
        if (!er.ok()) return er;
      }
      free_.push_back(slots_.back().get());
    }
    started_ = Clock::now();
    worker_ = std::thread([this] { Worker(); });
    return Ok();
  }

  Error Submit(const pirg_request_view& r, uint64_t tag) {
    int status = 0;
    Error er = SubmitMany(1, &r, &tag, &status);
    if (!er.ok()) return er;
    if (status != 0) {
      return ValidateRequest(S_, r.first_dim_len, r.second_dim_polys, r.packing_key_polys);
    }
    return Ok();
  }

  // Validates every request (statuses[i], ProcessResponse's checks), copies
  // the valid ones (several threads for large calls) and queues them under
  // one lock. Queues none when they do not all fit under max_outstanding.
  Error SubmitMany(int n, const pirg_request_view* reqs, const uint64_t* tags, int* statuses) {
    std::vector<int> ok;
    ok.reserve(n);
    for (int i = 0; i < n; ++i) {
      const pirg_request_view& r = reqs[i];
      statuses[i] =
          ValidateRequest(S_, r.first_dim_len, r.second_dim_polys, r.packing_key_polys).code;
      if (statuses[i] == 0) ok.push_back(i);
    }
    if (ok.empty()) return Ok();
    Error re = Reserve(ok.size());
    if (!re.ok()) return re;
    std::vector<ServeRequest> qs(ok.size());
    {
      std::lock_guard<std::mutex> lock(pool_mu_);
      for (auto& q : qs) {
        if (free_recs_.empty()) break;
        q.rec = free_recs_.back();
        free_recs_.pop_back();
      }
    }
    auto copy = [&](size_t lo, size_t hi) {
      for (size_t j = lo; j < hi; ++j) {
        ServeRequest& q = qs[j];
        q.tag = tags[ok[j]];
        u64* dst;
        if (q.rec >= 0) {
          dst = records_.as<u64>() + static_cast<size_t>(q.rec) * rw_;
        } else {
          q.data = TakeBuffer(rw_);
          dst = q.data.data();
        }
        WriteRecord(S_, reqs[ok[j]], dst);
      }
    };
    CopyInParallel(qs.size(), qs.size() * rw_ * 8, 8, copy);
    Enqueue(qs);
    return Ok();
  }

  // Queues n records the caller wrote (pirg_request_record) into memory from
  // pirg_host_alloc: they are copied to HBM from where they lie, with no
  // host copy. All or none are queued.
  Error SubmitRecords(int n, const u64* const* records, const uint64_t* tags) {
    for (int i = 0; i < n; ++i) {
      if (records[i] == nullptr || !InHostBlock(records[i], rw_ * 8)) {
        return Invalid("record " + std::to_string(i) + " does not lie, whole (" +
                       std::to_string(rw_ * 8) +
                       " bytes), in memory from pirg_host_alloc.");
      }
    }
    if (!S_->deferred.ok()) return S_->deferred;
    if (n == 0) return Ok();
    Error re = Reserve(static_cast<size_t>(n));
    if (!re.ok()) return re;
    std::vector<ServeRequest> qs(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
      qs[i].tag = tags[i];
      qs[i].ext = records[i];
    }
    Enqueue(qs);
    return Ok();
  }

  // PIRG_OK when a result was delivered within timeout_ms, else
  // PIRG_DEADLINE_EXCEEDED, or PIRG_FAILED_PRECONDITION once the loop has
  // stopped with nothing left to deliver.
  int Poll(int timeout_ms, ServeResult* res) {
    std::unique_lock<std::mutex> lock(mu_);
    const auto until = Clock::now() + std::chrono::milliseconds(std::max(timeout_ms, 0));
    while (done_.empty()) {
      if (stopped_) return PIRG_FAILED_PRECONDITION;
      if (cv_done_.wait_until(lock, until) == std::cv_status::timeout && done_.empty()) {
        return stopped_ ? PIRG_FAILED_PRECONDITION : PIRG_DEADLINE_EXCEEDED;
      }
    }
    *res = std::move(done_.front());
    done_.pop_front();
    --outstanding_;
    return PIRG_OK;
  }

  // Up to max results: waits up to timeout_ms for the first, then takes
  // whatever else is ready. Return codes as Poll.
  int PollMany(int max, int timeout_ms, std::vector<ServeResult>* res) {
    std::unique_lock<std::mutex> lock(mu_);
    const auto until = Clock::now() + std::chrono::milliseconds(std::max(timeout_ms, 0));
    while (done_.empty()) {
      if (stopped_) return PIRG_FAILED_PRECONDITION;
      if (cv_done_.wait_until(lock, until) == std::cv_status::timeout && done_.empty()) {
        return stopped_ ? PIRG_FAILED_PRECONDITION : PIRG_DEADLINE_EXCEEDED;
      }
    }
    while (!done_.empty() && static_cast<int>(res->size()) < max) {
      res->push_back(std::move(done_.front()));
      done_.pop_front();
      --outstanding_;
    }
    return PIRG_OK;
  }

  // Answers everything submitted, joins the worker and wakes every poller.
  // Safe to call more than once and from several threads.
  void Stop() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stopping_ = true;
    }
    cv_work_.notify_all();
    {
      std::lock_guard<std::mutex> join_lock(join_mu_);
      if (worker_.joinable()) worker_.join();
    }
    {
      std::lock_guard<std::mutex> lock(mu_);
      stopped_ = true;
    }
    cv_done_.notify_all();
  }

  void Stats(pirg_serve_stats_t* st, bool reset) {
    std::vector<double> l;
    {
      std::lock_guard<std::mutex> lock(mu_);
      st->submitted = submitted_;
      st->answered = answered_;
      st->batches = batches_;
      st->mean_batch = batches_ > 0 ? static_cast<double>(batched_) / batches_ : 0.0;
      l = latencies_;
      st->slots = nslots_;
      st->max_batch = max_batch_;
      st->elapsed_ms = MsSince(started_, Clock::now());
      if (reset) {
        submitted_ = answered_ = batches_ = batched_ = 0;
        latencies_.clear();
        lat_next_ = 0;
        started_ = Clock::now();
      }
    }
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
      if (l.empty()) return 0.0;
      const size_t i = std::min(l.size() - 1, static_cast<size_t>(p * (l.size() - 1) + 0.5));
      return l[i];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = l.empty() ? 0.0 : l.back();
  }

 private:
  void Deliver(InFlight& f, int status, const std::string& msg) {
    const size_t per = ResponseWords(S_);
    const Clock::time_point now = Clock::now();
    // The responses leave the slot's pinned buffer before the lock is taken:
    // submitters and pollers are not held up by the copies.
    std::vector<ServeResult> rs(f.reqs.size());
    for (size_t i = 0; i < f.reqs.size(); ++i) {
      ServeResult& r = rs[i];
      r.tag = f.reqs[i].tag;
      r.status = status;
      r.msg = msg;
      if (status == 0) {
        const u64* src = f.slot->ws.h_out.as<u64>() + i * per;
        r.out.assign(src, src + per);
      }
      r.latency_ms = MsSince(f.reqs[i].submitted, now);
    }
    std::lock_guard<std::mutex> lock(mu_);
    for (auto& r : rs) {
      // The latest kLatencyWindow latencies, as a ring.
      if (latencies_.size() < kLatencyWindow) {
        latencies_.push_back(r.latency_ms);
      } else {
        latencies_[lat_next_] = r.latency_ms;
        lat_next_ = (lat_next_ + 1) % kLatencyWindow;
      }
      done_.push_back(std::move(r));
      ++answered_;
    }
    free_.push_back(f.slot);
    if (!f.recs.empty()) {
      std::lock_guard<std::mutex> plock(pool_mu_);
      // Reversed, so later takes pop them in ascending order and a batch's
      // records stay contiguous (one copy each run).
      free_recs_.insert(free_recs_.end(), f.recs.rbegin(), f.recs.rend());
      f.recs.clear();
    }
  }

  // Counts n requests as outstanding, or refuses them all.
  Error Reserve(size_t n) {
    std::lock_guard<std::mutex> lock(mu_);
    if (stopping_) return Error{PIRG_FAILED_PRECONDITION, "the serving loop is stopping"};
    if (outstanding_ + static_cast<long long>(n) > max_outstanding_) {
      return Error{PIRG_RESOURCE_EXHAUSTED,
                   "serving queue is full (" + std::to_string(outstanding_) +
                       " requests submitted and not yet polled)"};
    }
    // Reserved now, so concurrent submits cannot overshoot while copying.
    outstanding_ += static_cast<long long>(n);
    return Ok();
  }

  // Queues reserved requests for the worker.
  void Enqueue(std::vector<ServeRequest>& qs) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      const Clock::time_point now = Clock::now();
      for (auto& q : qs) {
        q.submitted = now;
        pending_.push_back(std::move(q));
      }
      submitted_ += static_cast<long long>(qs.size());
    }
    cv_work_.notify_one();
  }

  // The batch size n requests run at: powers of two up to 16, then
  // multiples of 16, at most max_batch. The padding reads zero or earlier
  // valid records of the slot and its answers are not fetched; a scan costs
  // little more for up to 15 more queries, and the loop needs one graph per
  // size (at most 20 for 256 requests).
  int Bucket(int n) const {
    int b = 1;
    while (b < n && b < 16) b <<= 1;
    if (n > 16) b = (n + 15) / 16 * 16;
    return std::min(b, max_batch_);
  }

  // A buffer of at least n words from the pool (or a new one).
  std::vector<u64> TakeBuffer(size_t n) {
    std::vector<u64> b;
    {
      std::lock_guard<std::mutex> lock(pool_mu_);
      if (!pool_.empty()) {
        b = std::move(pool_.back());
        pool_.pop_back();
      }
    }
    b.resize(n);
    return b;
  }

  void GiveBuffer(std::vector<u64>&& b) {
    std::lock_guard<std::mutex> lock(pool_mu_);
    if (pool_.size() < static_cast<size_t>(max_batch_) * (nslots_ + 1)) {
      pool_.push_back(std::move(b));
    }
  }

  Error Launch(Slot& slot, InFlight& f) {
    std::vector<ServeRequest>& reqs = f.reqs;
    const int n = static_cast<int>(reqs.size());
    // Pool records go to HBM as they are; heap records (pool exhausted) are
    // moved into the slot's pinned buffer at their batch position first.
    std::vector<const u64*> recs(reqs.size());
    size_t heap = 0;
    for (size_t i = 0; i < reqs.size(); ++i) {
      ServeRequest& q = reqs[i];
      if (q.ext != nullptr) {
        recs[i] = q.ext;
      } else if (q.rec >= 0) {
        recs[i] = records_.as<u64>() + static_cast<size_t>(q.rec) * rw_;
        f.recs.push_back(q.rec);
      } else {
        recs[i] = slot.ws.h_req.as<u64>() + i * rw_;
        ++heap;
      }
    }
    auto on_heap = [](const ServeRequest& q) { return q.ext == nullptr && q.rec < 0; };
    if (heap > 0) {
      CopyInParallel(reqs.size(), heap * rw_ * 8, 16, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
          if (on_heap(reqs[i])) {
            std::memcpy(const_cast<u64*>(recs[i]), reqs[i].data.data(), rw_ * 8);
          }
        }
      });
      for (auto& q : reqs) {
        if (on_heap(q)) GiveBuffer(std::move(q.data));
      }
    }
    Error er = StageRecords(S_, slot, recs.data(), n);
    if (!er.ok()) return er;
    // A batch whose DB scan is MFMA-bound (a WideScan plan) gains nothing
    // from sharing the GPU: it runs its kernels after the previous batch's,
    // in launch order, while its own copies go ahead. Two such batches
    // running together would finish together, and both slots would then
    // stage their next batches with the GPU idle. Smaller batches, whose
    // HBM-bound scans and latency-bound tails do gain, overlap freely.
    const int nb = Bucket(n);
    if (last_ != nullptr && last_ != &slot && PlanFor(S_, true, nb).wide) {
      PIRG_HIP(hipStreamWaitEvent(slot.stream, last_->graph_done, 0));
    }
    er = RunBatch(S_, slot, nb, true);
    if (!er.ok()) return er;
    PIRG_HIP(hipEventRecord(slot.graph_done, slot.stream));
    last_ = &slot;
    er = EnqueueFetch(S_, slot, n);
    if (!er.ok()) return er;
    PIRG_HIP(hipEventRecord(slot.done, slot.stream));
    return Ok();
  }

  void Worker() {
    (void)hipSetDevice(S_->device);
    std::deque<InFlight> inflight;
    for (;;) {
      bool progressed = false;
      // Retire finished batches, oldest first.
      for (auto it = inflight.begin(); it != inflight.end();) {
        const hipError_t q = hipEventQuery(it->slot->done);
        if (q == hipErrorNotReady) {
          ++it;
          continue;
        }
        CollectTiming(S_, *it->slot);
        if (q == hipSuccess) {
          Deliver(*it, 0, "");
        } else {
          Deliver(*it, PIRG_INTERNAL, std::string("batch failed: ") + hipGetErrorString(q));
        }
        it = inflight.erase(it);
        progressed = true;
      }
      if (progressed) cv_done_.notify_all();
      // Launch whatever is waiting on every free slot.
      for (;;) {
        InFlight f;
        {
          std::lock_guard<std::mutex> lock(mu_);
          if (pending_.empty() || free_.empty()) break;
          f.slot = free_.front();
          free_.pop_front();
          const size_t n = std::min(pending_.size(), static_cast<size_t>(max_batch_));
          for (size_t i = 0; i < n; ++i) {
            f.reqs.push_back(std::move(pending_.front()));
            pending_.pop_front();
          }
          ++batches_;
          batched_ += static_cast<long long>(n);
        }
        Error er = Launch(*f.slot, f);
        if (!er.ok()) {
          (void)hipStreamSynchronize(f.slot->stream);
          Deliver(f, er.code, er.msg);
          cv_done_.notify_all();
        } else {
          inflight.push_back(std::move(f));
        }
        progressed = true;
      }
      std::unique_lock<std::mutex> lock(mu_);
      if (stopping_ && pending_.empty() && inflight.empty()) break;
      if (progressed) continue;
      if (!inflight.empty()) {
        // Batches on the GPU: check again shortly (or at the next submit).
        cv_work_.wait_for(lock, std::chrono::microseconds(50));
      } else {
        cv_work_.wait(lock, [this] { return stopping_ || !pending_.empty(); });
      }
    }
  }

  pirg_server* S_;
  int max_batch_;
  int nslots_;
  long long max_outstanding_;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::deque<Slot*> free_;  // guarded by mu_
  Slot* last_ = nullptr;    // the slot of the latest launched batch (worker only)
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<ServeRequest> pending_;
  std::deque<ServeResult> done_;
  long long outstanding_ = 0;
  bool stopping_ = false;
  bool stopped_ = false;  // the worker has finished; guarded by mu_
  long long submitted_ = 0, answered_ = 0, batches_ = 0, batched_ = 0;
  // Percentiles cover the latest kLatencyWindow answers since start or reset.
  static constexpr size_t kLatencyWindow = size_t(1) << 20;
  std::vector<double> latencies_;
  size_t lat_next_ = 0;
  Clock::time_point started_;
  std::mutex join_mu_;
  std::thread worker_;
  // The pinned record pool and its free records, and heap record buffers
  // for reuse when the pool is exhausted (a large request would otherwise
  // cost a fresh allocation and its page faults on every submit); all under
  // pool_mu_.
  size_t rw_ = 0;
  HostBuf records_;
  std::mutex pool_mu_;
  std::vector<int> free_recs_;
  std::vector<std::vector<u64>> pool_;
};

// The server's loop, if serving, as a reference the caller keeps for the call.
std::shared_ptr<ServeLoop> LoopOf(pirg_server* S) {
  std::lock_guard<std::mutex> lock(S->serve_mu);
  return S->serve;
}

}  // namespace pirg

using namespace pirg;

pirg_server::pirg_server() = default;

pirg_server::~pirg_server() {
  std::shared_ptr<ServeLoop> loop;
  {
    std::lock_guard<std::mutex> lock(serve_mu);
    loop.swap(serve);
  }
  if (loop) loop->Stop();
  loop.reset();
  if (main.stream != nullptr) (void)hipStreamSynchronize(main.stream);
}

extern "C" {

int pirg_serve_start(pirg_server* S, int max_batch, int slots, int max_outstanding, char* err,
                     size_t errlen) {
  if (S == nullptr) return Report(Invalid("invalid server."), err, errlen);
  if (max_batch < 1 || slots < 1 || max_outstanding < 1) {
    return Report(Invalid("max_batch, slots and max_outstanding must be positive."), err, errlen);
  }
  if (!S->deferred.ok()) return Report(S->deferred, err, errlen);
  if (LoopOf(S)) {
    return Report(Error{PIRG_FAILED_PRECONDITION, "the server is already serving"}, err, errlen);
  }
  std::shared_ptr<ServeLoop> loop = std::make_shared<ServeLoop>(S, max_batch, slots,
                                                                max_outstanding);
  Error er = loop->Start();
  if (!er.ok()) return Report(er, err, errlen);
  std::lock_guard<std::mutex> lock(S->serve_mu);
  S->serve = std::move(loop);
  return 0;
}

int pirg_serve_submit(pirg_server* S, const pirg_request_view* req, uint64_t tag, char* err,
                      size_t errlen) {
  if (S == nullptr || req == nullptr) return Report(Invalid("invalid server or request."), err, errlen);
  const std::shared_ptr<ServeLoop> loop = LoopOf(S);
  if (!loop) return Report(Error{PIRG_FAILED_PRECONDITION, "the server is not serving"}, err, errlen);
  Error er = loop->Submit(*req, tag);
  return er.ok() ? 0 : Report(er, err, errlen);
}

int pirg_serve_submit_batch(pirg_server* S, int n, const pirg_request_view* reqs,
                            const uint64_t* tags, int* statuses, char* err, size_t errlen) {
  if (S == nullptr || n < 0 || (n > 0 && (reqs == nullptr || tags == nullptr ||
                                          statuses == nullptr))) {
    return Report(Invalid("invalid server, count, requests, tags or statuses."), err, errlen);
  }
  const std::shared_ptr<ServeLoop> loop = LoopOf(S);
  if (!loop) return Report(Error{PIRG_FAILED_PRECONDITION, "the server is not serving"}, err, errlen);
  Error er = loop->SubmitMany(n, reqs, tags, statuses);
  if (!er.ok()) return Report(er, err, errlen);
  for (int i = 0; i < n; ++i) {
    if (statuses[i] != 0) {
      (void)Report(ValidateRequest(S, reqs[i].first_dim_len, reqs[i].second_dim_polys,
                                   reqs[i].packing_key_polys),
                   err, errlen);
      break;
    }
  }
  return 0;
}

int pirg_serve_set_memory_limit(pirg_server* S, size_t bytes) {
  if (S == nullptr) return PIRG_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(S->serve_mu);
  S->serve_mem_limit = bytes;
  return 0;
}

size_t pirg_serve_workspace_bytes(const pirg_server* S, int batch) {
  if (S == nullptr || batch < 1) return 0;
  return WorkspaceBytes(S, batch);
}

int pirg_host_alloc(size_t bytes, void** out, char* err, size_t errlen) {
  if (out == nullptr || bytes == 0) return Report(Invalid("bytes must be positive and out non-null."), err, errlen);
  *out = nullptr;
  void* p = nullptr;
  const hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
  if (e != hipSuccess || p == nullptr) {
    return Report(Error{PIRG_RESOURCE_EXHAUSTED, std::string("hipHostMalloc of ") +
                                                     std::to_string(bytes) + " bytes: " +
                                                     hipGetErrorString(e)},
                  err, errlen);
  }
  {
    std::lock_guard<std::mutex> lock(g_host_mu);
    g_host_blocks[reinterpret_cast<uintptr_t>(p)] = bytes;
  }
  *out = p;
  return 0;
}

void pirg_host_free(void* p) {
  if (p == nullptr) return;
  {
    std::lock_guard<std::mutex> lock(g_host_mu);
    auto it = g_host_blocks.find(reinterpret_cast<uintptr_t>(p));
    if (it == g_host_blocks.end()) return;
    g_host_blocks.erase(it);
  }
  (void)hipHostFree(p);
}

size_t pirg_request_record_words(const pirg_server* S) {
  return S == nullptr ? 0 : RequestWords(S);
}

int pirg_request_record(const pirg_server* S, const pirg_request_view* req, uint64_t* record,
                        char* err, size_t errlen) {
  if (S == nullptr || req == nullptr || record == nullptr) {
    return Report(Invalid("server, request and record must not be null."), err, errlen);
  }
  Error er = ValidateRequest(S, req->first_dim_len, req->second_dim_polys,
                             req->packing_key_polys);
  if (!er.ok()) return Report(er, err, errlen);
  WriteRecord(S, *req, record);
  return 0;
}

int pirg_serve_submit_records(pirg_server* S, int n, const uint64_t* const* records,
                              const uint64_t* tags, char* err, size_t errlen) {
  if (S == nullptr || n < 0 || (n > 0 && (records == nullptr || tags == nullptr))) {
    return Report(Invalid("invalid server, count, records or tags."), err, errlen);
  }
  const std::shared_ptr<ServeLoop> loop = LoopOf(S);
  if (!loop) return Report(Error{PIRG_FAILED_PRECONDITION, "the server is not serving"}, err, errlen);
  Error er = loop->SubmitRecords(n, records, tags);
  if (!er.ok()) return Report(er, err, errlen);
  return 0;
}

int pirg_serve_poll_batch(pirg_server* S, int max, int timeout_ms, uint64_t* tags,
                          uint64_t* const* outs, int* statuses, double* latency_ms, int* count,
                          char* err, size_t errlen) {
  if (S == nullptr || max < 1 || tags == nullptr || outs == nullptr || statuses == nullptr ||
      count == nullptr) {
    return Report(Invalid("invalid server, count or output arrays."), err, errlen);
  }
  *count = 0;
  const std::shared_ptr<ServeLoop> loop = LoopOf(S);
  if (!loop) return Report(Error{PIRG_FAILED_PRECONDITION, "the server is not serving"}, err, errlen);
  std::vector<ServeResult> rs;
  const int rc = loop->PollMany(max, timeout_ms, &rs);
  if (rc == PIRG_FAILED_PRECONDITION) {
    return Report(Error{PIRG_FAILED_PRECONDITION, "the serving loop stopped"}, err, errlen);
  }
  if (rc != PIRG_OK) return rc;
  bool reported = false;
  for (size_t i = 0; i < rs.size(); ++i) {
    tags[i] = rs[i].tag;
    statuses[i] = rs[i].status;
    if (latency_ms) latency_ms[i] = rs[i].latency_ms;
    if (rs[i].status == 0) {
      if (outs[i]) std::memcpy(outs[i], rs[i].out.data(), rs[i].out.size() * 8);
    } else if (!reported && err != nullptr && errlen > 0) {
      std::snprintf(err, errlen, "%s", rs[i].msg.c_str());
      reported = true;
    }
  }
  *count = static_cast<int>(rs.size());
  return 0;
}

int pirg_serve_poll(pirg_server* S, int timeout_ms, uint64_t* tag, uint64_t* out, int* status,
                    double* latency_ms, char* err, size_t errlen) {
  if (S == nullptr) return Report(Invalid("invalid server."), err, errlen);
  const std::shared_ptr<ServeLoop> loop = LoopOf(S);
  if (!loop) return Report(Error{PIRG_FAILED_PRECONDITION, "the server is not serving"}, err, errlen);
  ServeResult r;
  const int rc = loop->Poll(timeout_ms, &r);
  if (rc == PIRG_FAILED_PRECONDITION) {
    return Report(Error{PIRG_FAILED_PRECONDITION, "the serving loop stopped"}, err, errlen);
  }
  if (rc != PIRG_OK) return rc;
  if (tag) *tag = r.tag;
  if (status) *status = r.status;
  if (latency_ms) *latency_ms = r.latency_ms;
  if (r.status == 0) {
    if (out) std::memcpy(out, r.out.data(), r.out.size() * 8);
  } else if (err != nullptr && errlen > 0) {
    std::snprintf(err, errlen, "%s", r.msg.c_str());
  }
  return 0;
}

int pirg_serve_stats(pirg_server* S, pirg_serve_stats_t* st, int reset) {
  if (S == nullptr || st == nullptr) return PIRG_INVALID_ARGUMENT;
  const std::shared_ptr<ServeLoop> loop = LoopOf(S);
  if (!loop) return PIRG_FAILED_PRECONDITION;
  loop->Stats(st, reset != 0);
  return 0;
}

int pirg_serve_stop(pirg_server* S) {
  if (S == nullptr) return PIRG_INVALID_ARGUMENT;
  std::shared_ptr<ServeLoop> loop;
  {
    std::lock_guard<std::mutex> lock(S->serve_mu);
    loop.swap(S->serve);
  }
  if (loop) loop->Stop();
  return 0;
}

}  // extern "C"
