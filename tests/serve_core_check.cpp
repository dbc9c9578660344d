// CPU checks of the serving scheduler (guanaco_amd/csrc/gx_serve_core.hpp)
// with a fake executor: no device, jobs finish after a set number of polls.
// Prints one line per check; exit status 0 when all pass.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "gx_serve_core.hpp"

using gx::serve::Arena;
using gx::serve::Job;
using gx::serve::Scheduler;
using gx::serve::State;

static int failures = 0;
#define CHECK(cond)                                                      \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);         \
      ++failures;                                                        \
    }                                                                    \
  } while (0)

struct FakeReq {
  int polls = 1;           // polls until done
  bool fail_start = false; // start() fails (an allocation failure inside)
  bool fail_run = false;   // the device reports an error
  int seen = 0;
};

struct FakeExec {
  int n;
  std::vector<uint64_t> started;
  int finished = 0, dropped = 0;
  explicit FakeExec(int slots) : n(slots) {}
  int slots() { return n; }
  bool start(Job &j, int) {
    auto *r = static_cast<FakeReq *>(j.payload.get());
    started.push_back(j.id);
    if (r->fail_start) {
      j.error = "injected allocation failure";
      return false;
    }
    return true;
  }
  int poll(Job &j, int) {
    auto *r = static_cast<FakeReq *>(j.payload.get());
    if (++r->seen < r->polls) return 0;
    if (r->fail_run) {
      j.error = "injected device error";
      return -1;
    }
    return 1;
  }
  void finish(Job &) { ++finished; }
  void drop(Job &) { ++dropped; }
};

static std::shared_ptr<void> req(int polls, bool fail_start = false,
                                 bool fail_run = false) {
  auto r = std::make_shared<FakeReq>();
  r->polls = polls;
  r->fail_start = fail_start;
  r->fail_run = fail_run;
  return r;
}

static uint64_t submit(Scheduler<FakeExec> &s, size_t work, size_t keep,
                       std::shared_ptr<void> p) {
  uint64_t id = 0;
  std::string err;
  CHECK(s.submit(work, keep, 1.0, std::move(p), &id, &err) == 0);
  return id;
}

static State state(Scheduler<FakeExec> &s, uint64_t id) {
  return s.info(id)->state;
}

static void arena_checks() {
  Arena a(1024, 256);
  size_t o1, o2, o3, o4;
  CHECK(a.alloc(100, &o1) && o1 == 0);
  CHECK(a.alloc(256, &o2) && o2 == 256);
  CHECK(a.alloc(512, &o3) && o3 == 512);
  CHECK(!a.alloc(1, &o4));  // full
  CHECK(a.used() == 1024);
  CHECK(a.release(o2));
  CHECK(!a.release(o2));    // not live any more
  CHECK(a.largest_free() == 256);
  CHECK(!a.alloc(300, &o4));
  CHECK(a.release(o1));     // coalesces with the hole after it
  CHECK(a.largest_free() == 512);
  CHECK(a.alloc(300, &o4) && o4 == 0);
  CHECK(a.release(o3));
  CHECK(a.release(o4));
  CHECK(a.used() == 0 && a.largest_free() == 1024 && a.peak() == 1024);
  std::printf("OK arena\n");
}

static void fifo_and_backfill() {
  for (int backfill = 0; backfill < 2; ++backfill) {
    FakeExec ex(4);
    Scheduler<FakeExec> s(&ex, 1024, backfill != 0);
    const uint64_t a = submit(s, 768, 0, req(3));
    const uint64_t b = submit(s, 768, 0, req(1));
    const uint64_t c = submit(s, 256, 0, req(1));
    s.step();
    CHECK(state(s, a) == State::kRunning);
    CHECK(state(s, b) == State::kQueued);
    CHECK(state(s, c) == (backfill ? State::kRunning : State::kQueued));
    for (int i = 0; i < 10; ++i) s.step();
    CHECK(state(s, a) == State::kDone && state(s, b) == State::kDone &&
          state(s, c) == State::kDone);
    if (backfill) {
      CHECK(ex.started == (std::vector<uint64_t>{a, c, b}));
    } else {
      CHECK(ex.started == (std::vector<uint64_t>{a, b, c}));
    }
    CHECK(s.stats().used_bytes == 0);
  }
  std::printf("OK fifo and backfill admission\n");
}

static void bounded_bypass() {
  // a large request at the head is overtaken by small ones at most
  // kMaxBypass times, then waits for the room it needs
  FakeExec ex(64);
  Scheduler<FakeExec> s(&ex, 4096, true);
  const uint64_t small0 = submit(s, 2048, 0, req(1000000));  // holds half
  const uint64_t big = submit(s, 4096, 0, req(1));
  std::vector<uint64_t> smalls;
  for (int i = 0; i < 20; ++i) smalls.push_back(submit(s, 256, 0, req(1)));
  for (int i = 0; i < 40; ++i) s.step();
  CHECK(state(s, small0) == State::kRunning);
  CHECK(state(s, big) == State::kQueued);
  int done = 0;
  for (uint64_t id : smalls) done += state(s, id) == State::kDone;
  CHECK(done == gx::serve::kMaxBypass);
  CHECK(s.info(big)->bypassed == gx::serve::kMaxBypass);
  std::printf("OK bounded backfill\n");
}

static void slots_limit() {
  FakeExec ex(2);
  Scheduler<FakeExec> s(&ex, 1 << 20, false);
  std::vector<uint64_t> ids;
  for (int i = 0; i < 5; ++i) ids.push_back(submit(s, 1024, 0, req(2)));
  s.step();
  CHECK(s.stats().running == 2 && s.stats().queued == 3);
  for (int i = 0; i < 20; ++i) s.step();
  CHECK(s.stats().completed == 5);
  std::printf("OK one request per slot\n");
}

static void refusal() {
  FakeExec ex(4);
  Scheduler<FakeExec> s(&ex, 1024, false);
  uint64_t id = 0;
  std::string err;
  CHECK(s.submit(900, 200, 1.0, req(1), &id, &err) == 1);
  CHECK(err.find("budget") != std::string::npos);
  CHECK(s.stats().submitted == 0);
  std::printf("OK refusal above the budget\n");
}

static void cancel_checks() {
  FakeExec ex(1);
  Scheduler<FakeExec> s(&ex, 1024, false);
  const uint64_t a = submit(s, 500, 100, req(3));
  const uint64_t b = submit(s, 500, 100, req(1));
  s.step();
  CHECK(state(s, a) == State::kRunning && state(s, b) == State::kQueued);
  CHECK(s.cancel(b));  // queued: dropped at once, never started
  CHECK(state(s, b) == State::kCancelled);
  CHECK(s.cancel(a));  // running: reported cancelled when it ends
  for (int i = 0; i < 5; ++i) s.step();
  CHECK(state(s, a) == State::kCancelled);
  CHECK(ex.started == std::vector<uint64_t>{a});
  CHECK(s.stats().used_bytes == 0);  // its result memory was dropped too
  CHECK(s.stats().cancelled == 2 && s.stats().completed == 0);
  std::printf("OK cancellation\n");
}

static void failure_checks() {
  FakeExec ex(2);
  Scheduler<FakeExec> s(&ex, 2048, false);
  const uint64_t a = submit(s, 400, 100, req(1, true));
  const uint64_t b = submit(s, 400, 100, req(2, false, true));
  const uint64_t c = submit(s, 400, 100, req(1));
  for (int i = 0; i < 10; ++i) s.step();
  CHECK(state(s, a) == State::kFailed);
  CHECK(s.info(a)->error == "injected allocation failure");
  CHECK(state(s, b) == State::kFailed);
  CHECK(s.info(b)->error == "injected device error");
  CHECK(state(s, c) == State::kDone);
  const auto st = s.stats();
  CHECK(st.failed == 2 && st.completed == 1);
  CHECK(st.used_bytes == 256);  // c's result (100 bytes, one 256-byte unit)
  CHECK(s.release(c));
  CHECK(s.stats().used_bytes == 0);
  CHECK(!s.info(c));  // released requests are forgotten
  std::printf("OK failures free their memory\n");
}

static void keep_until_release() {
  FakeExec ex(4);
  Scheduler<FakeExec> s(&ex, 1024, false);
  const uint64_t a = submit(s, 256, 512, req(1));
  const uint64_t b = submit(s, 256, 512, req(1));
  for (int i = 0; i < 5; ++i) s.step();
  CHECK(state(s, a) == State::kDone);
  // a's volume holds 512 bytes: b (768) does not fit until a is released
  CHECK(state(s, b) == State::kQueued);
  CHECK(!s.release(b));  // not finished
  CHECK(s.release(a));
  for (int i = 0; i < 5; ++i) s.step();
  CHECK(state(s, b) == State::kDone);
  std::printf("OK results held until released\n");
}

static void stats_checks() {
  std::vector<double> v = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
  CHECK(gx::serve::percentile(v, 50) == 5.5);
  CHECK(gx::serve::percentile(v, 0) == 1 && gx::serve::percentile(v, 100) == 10);
  CHECK(std::abs(gx::serve::percentile(v, 99) - 9.91) < 1e-12);
  CHECK(gx::serve::percentile({}, 50) == 0);
  std::printf("OK percentiles\n");
}

static void threaded() {
  FakeExec ex(3);
  Scheduler<FakeExec> s(&ex, 1 << 16, true);
  s.start_thread();
  std::vector<uint64_t> ids;
  for (int i = 0; i < 40; ++i)
    ids.push_back(submit(s, 1024 + 512 * (i % 5), 256, req(1 + i % 4)));
  for (uint64_t id : ids) {
    auto j = s.wait(id, -1);
    CHECK(j && j->state == State::kDone);
    CHECK(j->end_s >= j->start_s && j->start_s >= j->submit_s);
    CHECK(s.release(id));
  }
  s.stop();
  const auto st = s.stats();
  CHECK(st.completed == 40 && st.used_bytes == 0);
  CHECK(st.p99_ms >= st.p50_ms && st.p50_ms > 0);
  std::printf("OK threaded loop\n");
}

int main() {
  arena_checks();
  fifo_and_backfill();
  bounded_bypass();
  slots_limit();
  refusal();
  cancel_checks();
  failure_checks();
  keep_until_release();
  stats_checks();
  threaded();
  if (failures) {
    std::printf("%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("ALL OK\n");
  return 0;
}
