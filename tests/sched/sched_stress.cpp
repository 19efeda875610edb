// Stress test of the serving form's scheduler (server_sched.h) on the CPU with a stub executor,
// built with ThreadSanitizer or AddressSanitizer + UndefinedBehaviorSanitizer by
// tests/test_server_sched.py.
//
// Scenarios (each repeated with fresh schedulers in one process, as the server is created and
// destroyed repeatedly by one program):
//   drain : an 8-slot server (3 large-lane workers), then a 4-slot one (2 large-lane workers),
//           jobs submitted with random gaps and polled, some of them busy-polled;
//   stop  : jobs that small-lane workers hand over to the large lane, stop() right after the
//           submissions: every job must still finish (no large-lane worker may exit while a
//           small-lane worker can still hand a job over); the stop_slow_plan variant plans much
//           longer than it executes, which makes that window wide;
//   tight : a budget that admits about one job at a time;
//   nolane: no lanes (every worker takes everything);
//   cancel: jobs cancelled at random moments (queued, handed over, planning, waiting for
//           admission, running, done): a cancelled job finishes once, with the cancel status or,
//           if its executor was past its last check, normally; it never holds budget afterwards.
// In every scenario some jobs "run out of memory" once they are admitted and run again through
// admit_alone(), as the server reruns a query that ran out of device memory next to others.
// Checks: every job finishes exactly once with its status, no admission beyond the budget
// (unless alone), nothing admitted next to a job admitted alone, handed-over jobs run on
// large-lane workers, nothing left queued.  A watchdog
// ends the process with exit status 3 and the scheduler's counters if a scenario hangs.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <thread>

#include "server_sched.h"

using smcm_sched::Job;
using smcm_sched::Scheduler;

namespace {

struct StubJob : Job {
  int plan_us = 0, exec_us = 0;
  uint64_t bytes = 0;
  bool big = false, fail = false, oom = false;
  bool alone_ran = false;
  std::atomic<int> runs{0};
  int ran_on = -1;
};

std::mutex g_mu;
Scheduler* g_current = nullptr;
std::string g_scenario;
std::atomic<bool> g_finished{false};

void sleep_us(int us) {
  if (us > 0) std::this_thread::sleep_for(std::chrono::microseconds(us));
}

int failures = 0;
// cancellations over all scenarios: of jobs no worker had taken, and of jobs a worker was running
std::atomic<int64_t> g_cancel_queued{0}, g_cancel_taken{0};
#define CHECK(cond, ...)                                 \
  do {                                                   \
    if (!(cond)) {                                       \
      std::fprintf(stderr, "FAIL %s: ", g_scenario.c_str()); \
      std::fprintf(stderr, __VA_ARGS__);                 \
      std::fprintf(stderr, "\n");                        \
      ++failures;                                        \
    }                                                    \
  } while (0)

constexpr int kCancelStatus = -125;

void run_scenario(const char* name, int slots, int large_slots, uint64_t budget, int njobs, double p_big,
                  bool stop_early, uint64_t seed, int plan_max_us = 300, int exec_max_us = 600,
                  double p_cancel = 0.0) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_scenario = std::string(name) + " slots=" + std::to_string(slots) + " large=" + std::to_string(large_slots) +
                 " seed=" + std::to_string(seed);
  }
  Scheduler sc(slots, large_slots, budget);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_current = &sc;
  }
  const int small_lanes = sc.small_lanes();
  std::atomic<int64_t> over{0}, not_alone{0};
  sc.start([&](Job& j, int wi, bool small) {
    auto& q = static_cast<StubJob&>(j);
    q.runs++;
    // planning; on the large lane a handed-over job's plan is shorter, so large-lane workers can
    // run dry while small-lane workers still plan
    sleep_us(small ? q.plan_us : q.plan_us / 20);
    auto cancelled = [&] {
      if (!q.cancelled.load()) return false;
      q.status = kCancelStatus;
      q.error = "cancelled";
      return true;
    };
    if (cancelled()) return true;
    if (small && q.big) {
      sc.hand_over(q);
      return false;
    }
    q.ran_on = wi;
    if (q.fail) {
      q.status = -5;
      q.error = "stub failure";
      return true;
    }
    if (!sc.admit(q, q.bytes)) {
      cancelled();
      return true;
    }
    const auto c = sc.counters();
    if (c.admitted > 1 && c.inflight > budget) over++;
    if (budget == 1 && c.admitted > 1) over++;  // a one-byte budget admits one job at a time
    sleep_us(q.exec_us / 2);  // device work, with a check half way
    if (cancelled()) {
      sc.release(q);
      return true;
    }
    sleep_us(q.exec_us - q.exec_us / 2);
    sc.release(q);
    if (q.oom) {  // ran out of memory: once more, by itself
      if (!sc.admit_alone(q)) {
        cancelled();
        return true;
      }
      q.alone_ran = true;
      if (sc.counters().admitted != 1) not_alone++;
      sleep_us(q.exec_us);
      if (sc.counters().admitted != 1) not_alone++;
      sc.release(q);
    }
    q.status = 0;
    return true;
  });
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<int64_t> tickets;
  std::vector<bool> want_fail;
  std::map<int64_t, int> cancel_rc;  // ticket -> cancel()'s return
  int64_t n_oom = 0;
  std::map<int64_t, std::unique_ptr<Job>> early;  // finished jobs taken by a non-blocking poll
  for (int i = 0; i < njobs; ++i) {
    auto j = std::make_unique<StubJob>();
    j->plan_us = (int)(u(rng) * plan_max_us);
    j->exec_us = (int)(u(rng) * exec_max_us);
    j->big = u(rng) < p_big;
    j->bytes = u(rng) < 0.1 ? 0 : (uint64_t)(u(rng) * (double)budget * (j->big ? 0.9 : 0.3));
    j->fail = u(rng) < 0.03;
    j->oom = !j->fail && u(rng) < 0.08;
    n_oom += j->oom ? 1 : 0;
    want_fail.push_back(j->fail);
    tickets.push_back(sc.submit(std::move(j)));
    if (u(rng) < p_cancel) {  // cancel a random earlier ticket (it may be done, or polled)
      const int64_t t = tickets[(size_t)(u(rng) * (double)tickets.size())];
      if (!cancel_rc.count(t)) {
        const int rc = sc.cancel(t, kCancelStatus, "cancelled");
        CHECK(rc == 0 || rc == 1 || (rc == -1 && early.count(t)), "cancel of %lld: rc %d", (long long)t, rc);
        cancel_rc[t] = rc;
      }
    }
    if (!stop_early) {
      sleep_us((int)(u(rng) * 150));
      if (i % 7 == 3) {  // non-blocking poll of an earlier ticket
        const int64_t t = tickets[(size_t)(u(rng) * (double)tickets.size())];
        if (!early.count(t)) {
          std::unique_ptr<Job> out;
          const int rc = sc.poll(t, false, &out);
          CHECK(rc == 0 || rc == 1, "non-blocking poll of %lld: rc %d", (long long)t, rc);
          if (rc == 0) early[t] = std::move(out);
        }
      }
    }
  }
  if (stop_early) sc.stop();
  int done = 0;
  for (size_t i = 0; i < tickets.size(); ++i) {
    std::unique_ptr<Job> out;
    int rc = 0;
    if (early.count(tickets[i])) out = std::move(early[tickets[i]]);
    else rc = sc.poll(tickets[i], !stop_early, &out);
    if (stop_early) {
      CHECK(rc == 0, "ticket %lld not finished after stop (rc %d)", (long long)tickets[i], rc);
      if (rc != 0) rc = sc.poll(tickets[i], true, &out);
    }
    CHECK(rc == 0 && out, "ticket %lld: poll rc %d", (long long)tickets[i], rc);
    if (!out) continue;
    auto& q = static_cast<StubJob&>(*out);
    ++done;
    CHECK(q.done, "ticket %lld polled before done", (long long)q.ticket);
    const bool was_cancelled = cancel_rc.count(tickets[i]) && cancel_rc[tickets[i]] == 0;
    if (was_cancelled) {
      CHECK(q.status == kCancelStatus || q.status == (want_fail[i] ? -5 : 0), "cancelled ticket %lld status %d",
            (long long)q.ticket, q.status);
      if (q.status == kCancelStatus) {
        (q.runs == 0 || (q.large && q.ran_on < 0) ? g_cancel_queued : g_cancel_taken)++;
        if (q.oom && !q.alone_ran) n_oom--;  // cancelled before its alone run
        CHECK(q.runs <= 2, "cancelled ticket %lld ran %d times", (long long)q.ticket, q.runs.load());
      }
    } else {
      CHECK(q.status == (want_fail[i] ? -5 : 0), "ticket %lld status %d", (long long)q.ticket, q.status);
      CHECK(q.runs == (q.big && small_lanes > 0 && q.large ? 2 : 1) || (q.runs == 1 && !q.large),
            "ticket %lld ran %d times (large %d)", (long long)q.ticket, q.runs.load(), (int)q.large);
    }
    if (q.large && q.ran_on >= 0)
      CHECK(q.ran_on >= small_lanes, "handed-over ticket %lld ran on small-lane worker %d", (long long)q.ticket,
            q.ran_on);
    CHECK(q.t_done >= q.t_start && q.t_start >= q.t_submit, "ticket %lld times", (long long)q.ticket);
    CHECK(!q.admitted, "ticket %lld still holds budget", (long long)q.ticket);
    CHECK(sc.poll(tickets[i], false, nullptr) == -1, "ticket %lld polled twice", (long long)q.ticket);
  }
  CHECK(done == njobs, "%d of %d jobs finished", done, njobs);
  sc.stop();
  const auto c = sc.counters();
  CHECK(c.over_budget == 0 && over == 0, "admitted beyond the budget (%lld, %lld)", (long long)c.over_budget,
        (long long)over.load());
  CHECK(not_alone == 0 && c.alone_runs == n_oom && c.alone_waiting == 0,
        "alone admissions: %lld shared, %lld of %lld run, %d waiting", (long long)not_alone.load(),
        (long long)c.alone_runs, (long long)n_oom, c.alone_waiting);
  CHECK(c.inflight == 0 && c.admitted == 0 && c.running == 0 && c.small_busy == 0,
        "left: inflight %llu admitted %d running %d small_busy %d", (unsigned long long)c.inflight, c.admitted,
        c.running, c.small_busy);
  CHECK(c.pending == 0 && c.large_ready == 0 && c.tickets == 0 && c.completed == njobs,
        "left: pending %zu large_ready %zu tickets %zu completed %lld", c.pending, c.large_ready, c.tickets,
        (long long)c.completed);
  std::lock_guard<std::mutex> lk(g_mu);
  g_current = nullptr;
}

}  // namespace

int main(int argc, char** argv) {
  const int reps = argc > 1 ? std::atoi(argv[1]) : 10;
  const int timeout_s = argc > 2 ? std::atoi(argv[2]) : 240;
  std::thread([timeout_s] {
    for (int s = 0; s < timeout_s * 10 && !g_finished; ++s) std::this_thread::sleep_for(std::chrono::milliseconds(100));
    if (g_finished) return;
    std::lock_guard<std::mutex> lk(g_mu);
    std::fprintf(stderr, "HANG in %s", g_scenario.c_str());
    if (g_current) {
      const auto c = g_current->counters();
      std::fprintf(stderr, ": inflight %llu admitted %d running %d small_busy %d pending %zu large_ready %zu "
                   "tickets %zu completed %lld", (unsigned long long)c.inflight, c.admitted, c.running,
                   c.small_busy, c.pending, c.large_ready, c.tickets, (long long)c.completed);
    }
    std::fprintf(stderr, "\n");
    std::fflush(stderr);
    std::_Exit(3);
  }).detach();
  for (int r = 0; r < reps; ++r) {
    const uint64_t seed = 1000 + (uint64_t)r;
    // the shape that hung on the GPU: an 8-slot server, then a 4-slot server, in one process
    run_scenario("drain", 8, 3, 200, 120, 0.3, false, seed);
    run_scenario("drain", 4, 2, 200, 120, 0.3, false, seed + 7);
    run_scenario("stop", 8, 3, 200, 60, 0.8, true, seed + 13);
    run_scenario("stop", 4, 2, 200, 60, 1.0, true, seed + 17);
    // planning much longer than execution: large-lane workers run dry while small-lane workers
    // are still planning jobs they will hand over
    run_scenario("stop_slow_plan", 8, 3, 200, 40, 1.0, true, seed + 29, 3000, 50);
    run_scenario("stop_slow_plan", 4, 2, 200, 40, 1.0, true, seed + 31, 3000, 50);
    run_scenario("tight", 4, 2, 10, 60, 0.3, false, seed + 19);
    run_scenario("one_byte", 4, 2, 1, 40, 0.3, false, seed + 37);  // one job at a time
    run_scenario("nolane", 3, 0, 200, 60, 0.3, false, seed + 23);
    run_scenario("cancel", 8, 3, 200, 120, 0.3, false, seed + 41, 300, 600, 0.4);
    run_scenario("cancel_tight", 4, 2, 10, 80, 0.3, false, seed + 43, 300, 600, 0.4);
    run_scenario("cancel_stop", 4, 2, 200, 60, 1.0, true, seed + 47, 3000, 50, 0.5);
  }
  g_finished = true;
  CHECK(g_cancel_queued > 0 && g_cancel_taken > 0, "cancellations: %lld queued, %lld taken",
        (long long)g_cancel_queued.load(), (long long)g_cancel_taken.load());
  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("sched stress ok: %d rounds, %lld jobs cancelled before a worker took them, %lld while taken\n", reps,
              (long long)g_cancel_queued.load(), (long long)g_cancel_taken.load());
  return 0;
}
