// server_sched.h - the serving form's scheduler: submission queue, lanes, admission by predicted
// bytes, completion and shutdown.  It holds no device code (standard C++ threads only), so the
// same scheduler the server runs is exercised on the CPU under ThreadSanitizer and
// AddressSanitizer with a stub executor.
//
// Workers 0..small_lanes-1 form the small lane, the others the large lane (with no lanes every
// worker is a large-lane worker).  Every worker takes submitted jobs in arrival order; a
// small-lane worker that finds a job large hands it to the large lane (hand_over), whose workers
// take handed-over jobs first.  A job is admitted when its predicted bytes fit the budget next
// to the admitted jobs, or when no other job is admitted (alone it always runs).  A job can also
// ask to be admitted alone (admit_alone: after it ran out of device memory next to others); no
// other job is admitted while it waits or runs.
//
// Shutdown (stop) lets the workers finish every submitted job.  A small-lane worker exits once
// no job is waiting; a large-lane worker only once, in addition, no small-lane worker is busy,
// because a busy small-lane worker may still hand a job over.
//
// Cancellation (cancel): a job no worker has taken yet (queued, or handed over and waiting for the
// large lane) finishes at once with the given status; a taken job is flagged (Job::cancelled),
// stops waiting for admission, and its executor ends it at its next check.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace smcm_sched {

struct Job {
  int64_t ticket = 0;
  bool done = false;
  bool large = false;          // handed to the large lane
  int status = 0;              // set by the executor (0 = ok)
  std::string error;
  double t_submit = 0, t_start = 0, t_done = 0;  // seconds since the scheduler's creation
  uint64_t admitted_bytes = 0;
  bool admitted = false;
  std::atomic<bool> cancelled{false};  // set by Scheduler::cancel; read by the executor
  virtual ~Job() = default;
};

class Scheduler {
 public:
  // Runs one job on worker `worker`: returns true when the job is finished (status and error
  // set by the executor), false after hand_over().  It must not throw.
  using RunFn = std::function<bool(Job&, int worker, bool small_lane)>;

  Scheduler(int slots, int large_slots, uint64_t budget)
      : slots_(slots),
        small_lanes_((large_slots > 0 && large_slots < slots) ? slots - large_slots : 0),
        budget_(budget),
        t0_(clock_s()) {}
  ~Scheduler() { stop(); }
  Scheduler(const Scheduler&) = delete;
  Scheduler& operator=(const Scheduler&) = delete;

  void start(RunFn run) {
    run_ = std::move(run);
    for (int i = 0; i < slots_; ++i) workers_.emplace_back([this, i] { worker(i); });
  }

  // finishes every submitted job, then joins the workers (idempotent)
  void stop() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : workers_)
      if (t.joinable()) t.join();
    workers_.clear();
  }

  int64_t submit(std::unique_ptr<Job> j) {
    std::lock_guard<std::mutex> lk(mu_);
    j->ticket = next_ticket_++;
    j->t_submit = now();
    Job* raw = j.get();
    pending_.push_back(raw);
    all_.emplace(raw->ticket, std::move(j));
    cv_work_.notify_all();
    return raw->ticket;
  }

  // 0: done, the job is moved to *out and its ticket released; 1: still running (wait == false);
  // -1: unknown ticket
  int poll(int64_t ticket, bool wait, std::unique_ptr<Job>* out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(ticket);
    if (it == all_.end()) return -1;
    if (!it->second->done) {
      if (!wait) return 1;
      // looked up again on every wake-up: a concurrent poll of the same ticket may release it
      cv_done_.wait(lk, [&] {
        it = all_.find(ticket);
        return it == all_.end() || it->second->done;
      });
      if (it == all_.end()) return -1;
    }
    if (out) *out = std::move(it->second);
    all_.erase(it);
    return 0;
  }

  // blocks until `bytes` fit the budget next to the admitted jobs (or nothing else is admitted);
  // every job counts at least one byte, so a budget of b bytes admits at most b jobs at once.
  // No job is admitted while a job waits for admit_alone().  Returns false, admitting nothing,
  // when the job is cancelled before or while it waits.
  bool admit(Job& j, uint64_t bytes) {
    bytes = bytes > 0 ? bytes : 1;
    std::unique_lock<std::mutex> lk(mu_);
    cv_budget_.wait(lk, [&] {
      return j.cancelled.load() || (alone_waiting_ == 0 && (admitted_ == 0 || inflight_ + bytes <= budget_));
    });
    if (j.cancelled.load()) return false;
    take(j, bytes);
    return true;
  }

  // blocks until no other job is admitted, then admits j with the whole budget, so nothing else
  // is admitted until j is released: a job that ran out of device memory next to others runs
  // again by itself.  Jobs waiting here are admitted one at a time, before any admit() waiter.
  // Returns false, admitting nothing, when the job is cancelled.
  bool admit_alone(Job& j) {
    std::unique_lock<std::mutex> lk(mu_);
    alone_waiting_++;
    cv_budget_.wait(lk, [&] { return j.cancelled.load() || admitted_ == 0; });
    alone_waiting_--;
    if (j.cancelled.load()) {
      lk.unlock();
      cv_budget_.notify_all();  // admit() waiters held back by this one may go on
      return false;
    }
    take(j, budget_ > 0 ? budget_ : 1);
    alone_runs_++;
    return true;
  }

  // 0: cancelled (a job no worker has taken finishes now with `status` and `msg`; a taken one is
  // flagged for its executor, which may still finish it normally if it is past its last check);
  // 1: the job had already finished; -1: unknown ticket (or already polled)
  int cancel(int64_t ticket, int status, const std::string& msg) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = all_.find(ticket);
      if (it == all_.end()) return -1;
      Job* j = it->second.get();
      if (j->done) return 1;
      j->cancelled = true;
      auto drop = [&](std::deque<Job*>& dq) {
        auto q = std::find(dq.begin(), dq.end(), j);
        if (q == dq.end()) return false;
        dq.erase(q);
        return true;
      };
      if (drop(pending_) || drop(large_ready_)) {
        j->status = status;
        j->error = msg;
        j->t_done = now();
        if (j->t_start == 0) j->t_start = j->t_done;
        j->done = true;
        completed_++;
        cancelled_++;
      }
    }
    cv_done_.notify_all();
    cv_budget_.notify_all();  // a flagged job waiting for admission gives up
    cv_work_.notify_all();    // large-lane workers waiting to exit re-check the queues
    return 0;
  }

  // the job's device work is done: its budget is freed
  void release(Job& j) {
    if (!j.admitted) return;
    {
      std::lock_guard<std::mutex> lk(mu_);
      j.t_done = now();
      inflight_ -= j.admitted_bytes;
      admitted_--;
      j.admitted = false;
    }
    cv_budget_.notify_all();
  }

  // a small-lane worker passes a planned job to the large lane
  void hand_over(Job& j) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      j.large = true;
      large_ready_.push_back(&j);
    }
    cv_work_.notify_all();
  }

  double now() const { return clock_s() - t0_; }
  double t0() const { return t0_; }
  int small_lanes() const { return small_lanes_; }

  struct Counters {
    uint64_t inflight;
    int admitted, running, small_busy, alone_waiting;
    int64_t completed, over_budget, alone_runs, cancelled_queued;
    size_t pending, large_ready, tickets;
  };
  Counters counters() {
    std::lock_guard<std::mutex> lk(mu_);
    return Counters{inflight_, admitted_, running_, small_busy_, alone_waiting_, completed_, over_budget_,
                    alone_runs_, cancelled_, pending_.size(), large_ready_.size(), all_.size()};
  }

 private:
  static double clock_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }

  // caller holds mu_
  void take(Job& j, uint64_t bytes) {
    inflight_ += bytes;
    admitted_++;
    if (admitted_ > 1 && inflight_ > budget_) over_budget_++;  // never: checked by the stress test
    j.admitted_bytes = bytes;
    j.admitted = true;
    j.t_start = now();
  }

  void worker(int wi) {
    const bool small = wi < small_lanes_;
    for (;;) {
      Job* q = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] {
          return !pending_.empty() || (!small && !large_ready_.empty()) || (stop_ && (small || small_busy_ == 0));
        });
        if (!small && !large_ready_.empty()) {
          q = large_ready_.front();
          large_ready_.pop_front();
        } else if (!pending_.empty()) {
          q = pending_.front();
          pending_.pop_front();
        } else {
          return;  // stopping, nothing left for this worker
        }
        running_++;
        if (small) small_busy_++;
      }
      const bool finished = run_(*q, wi, small && !q->large);
      {
        std::lock_guard<std::mutex> lk(mu_);
        running_--;
        if (small) small_busy_--;
        if (finished) {
          if (q->t_done == 0) q->t_done = now();
          if (q->t_start == 0) q->t_start = q->t_done;
          q->done = true;
          completed_++;
        }
      }
      cv_done_.notify_all();
      // large-lane workers waiting to exit re-check small_busy_
      cv_work_.notify_all();
    }
  }

  const int slots_, small_lanes_;
  const uint64_t budget_;
  const double t0_;
  RunFn run_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_, cv_budget_;
  std::deque<Job*> pending_;      // submitted, not yet taken
  std::deque<Job*> large_ready_;  // handed over, waiting for a large-lane worker
  std::map<int64_t, std::unique_ptr<Job>> all_;
  int64_t next_ticket_ = 1;
  uint64_t inflight_ = 0;  // predicted bytes of the admitted jobs
  int admitted_ = 0;
  int running_ = 0;
  int small_busy_ = 0;
  int alone_waiting_ = 0;  // jobs blocked in admit_alone()
  int64_t completed_ = 0, over_budget_ = 0, alone_runs_ = 0;
  int64_t cancelled_ = 0;  // jobs cancelled before a worker took them
  bool stop_ = false;
};

}  // namespace smcm_sched
