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
#include <thread>

#include "server_internal.h"

namespace pirg {

using Clock = std::chrono::steady_clock;

namespace {

double MsSince(Clock::time_point t0, Clock::time_point t1) {
  return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

}  // namespace

struct ServeRequest {
  uint64_t tag = 0;
  Clock::time_point submitted;
  std::vector<u64> fd, sd, pk;
};

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
};

class ServeLoop {
 public:
  ServeLoop(pirg_server* S, int max_batch, int nslots, int max_outstanding)
      : S_(S), max_batch_(max_batch), nslots_(nslots), max_outstanding_(max_outstanding) {}

  ~ServeLoop() { Stop(); }

  Error Start() {
    // Admission by free HBM: as many slots of max_batch requests as fit in
    // 80% of the free device memory (fewer slots, then smaller batches).
    size_t free_b = 0, total_b = 0;
    PIRG_HIP(hipMemGetInfo(&free_b, &total_b));
    const double budget = 0.8 * static_cast<double>(free_b);
    while (max_batch_ > 1 &&
           static_cast<double>(WorkspaceBytes(S_, max_batch_)) * nslots_ > budget) {
      if (nslots_ > 2) {
        --nslots_;
      } else {
        max_batch_ /= 2;
      }
    }
    if (static_cast<double>(WorkspaceBytes(S_, max_batch_)) * nslots_ > budget) {
      return Error{PIRG_RESOURCE_EXHAUSTED, "not enough free device memory for a serving slot"};
    }
    for (int i = 0; i < nslots_; ++i) {
      slots_.emplace_back(new Slot());
      Error er = slots_.back()->Init();
      if (!er.ok()) return er;
      er = EnsureWorkspace(S_, *slots_.back(), max_batch_);
      if (!er.ok()) return er;
      free_.push_back(slots_.back().get());
    }
    started_ = Clock::now();
    worker_ = std::thread([this] { Worker(); });
    return Ok();
  }

  Error Submit(const pirg_request_view& r, uint64_t tag) {
    Error er = ValidateRequest(S_, r.first_dim_len, r.second_dim_polys, r.packing_key_polys);
    if (!er.ok()) return er;
    ServeRequest q;
    q.tag = tag;
    q.fd.assign(r.first_dim, r.first_dim + r.first_dim_len);
    if (S_->t > 1) q.sd.assign(r.second_dim, r.second_dim + r.second_dim_polys * S_->d);
    q.pk.assign(r.packing_key, r.packing_key + r.packing_key_polys * S_->d);
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (stopping_) return Error{PIRG_FAILED_PRECONDITION, "the serving loop is stopping"};
      if (outstanding_ >= max_outstanding_) {
        return Error{PIRG_RESOURCE_EXHAUSTED,
                     "serving queue is full (" + std::to_string(outstanding_) +
                         " requests submitted and not yet polled)"};
      }
      ++outstanding_;
      ++submitted_;
      q.submitted = Clock::now();
      pending_.push_back(std::move(q));
    }
    cv_work_.notify_one();
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
    std::lock_guard<std::mutex> lock(mu_);
    for (size_t i = 0; i < f.reqs.size(); ++i) {
      ServeResult r;
      r.tag = f.reqs[i].tag;
      r.status = status;
      r.msg = msg;
      if (status == 0) {
        const u64* src = f.slot->ws.h_out.as<u64>() + i * per;
        r.out.assign(src, src + per);
      }
      r.latency_ms = MsSince(f.reqs[i].submitted, now);
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
  }

  Error Launch(Slot& slot, std::vector<ServeRequest>& reqs) {
    std::vector<pirg_request_view> views(reqs.size());
    std::vector<int> idx(reqs.size());
    for (size_t i = 0; i < reqs.size(); ++i) {
      const ServeRequest& q = reqs[i];
      views[i] = pirg_request_view{q.fd.data(), q.fd.size(), q.sd.data(),
                                   q.sd.size() / static_cast<size_t>(S_->d), q.pk.data(),
                                   q.pk.size() / static_cast<size_t>(S_->d)};
      idx[i] = static_cast<int>(i);
    }
    const int n = static_cast<int>(reqs.size());
    Error er = StageRequests(S_, slot, views.data(), idx);
    if (!er.ok()) return er;
    er = RunBatch(S_, slot, n, true);
    if (!er.ok()) return er;
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
        Error er = Launch(*f.slot, f.reqs);
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
