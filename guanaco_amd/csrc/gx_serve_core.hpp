// Request scheduling of the serving loop, independent of the device.
//
// - Arena: a first-fit offset allocator over a fixed budget [0, capacity),
//   coalescing freed ranges; the executor adds the offsets to the base of
//   one device allocation.
// - Scheduler<Exec>: FIFO admission by footprint. A request holds `work`
//   bytes while it runs (plan, input copy) and `keep` bytes until the client
//   releases its result (the volume). The head of the queue starts as soon as
//   a slot (stream) is free and its footprint fits; with backfill, later
//   requests that fit may start while the head waits, but no request is
//   overtaken more than kMaxBypass times (then nothing may start before it,
//   so large requests cannot starve). Requests larger than
//   the budget are refused at submission. Completions are found by polling
//   the executor; several requests run at once, one per slot.
// - latency statistics (submission to completion) with percentiles.
//
// Exec provides, all called from the scheduler's thread:
//   int slots();                       concurrent requests
//   bool start(Job &, int slot);       launch; false: failed (Job::error set)
//   int poll(Job &, int slot);         0 running, 1 done, -1 failed
//   void finish(Job &);                release the executor's per-run state
//   void drop(Job &);                  release what start() left for `keep`
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace gx {
namespace serve {

class Arena {
 public:
  explicit Arena(size_t capacity, size_t align = 256)
      : cap_(capacity / align * align), align_(align) {
    if (cap_) free_[0] = cap_;
  }
  // false when no free range fits (the budget is full or fragmented)
  bool alloc(size_t bytes, size_t *offset) {
    const size_t need = round(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      const size_t off = it->first, rest = it->second - need;
      free_.erase(it);
      if (rest) free_[off + need] = rest;
      live_[off] = need;
      used_ += need;
      peak_ = std::max(peak_, used_);
      *offset = off;
      return true;
    }
    return false;
  }
  // frees a live allocation; false when `offset` is not one
  bool release(size_t offset) {
    auto it = live_.find(offset);
    if (it == live_.end()) return false;
    size_t off = offset, size = it->second;
    live_.erase(it);
    used_ -= size;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + size == next->first) {
      size += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += size;
        return true;
      }
    }
    free_[off] = size;
    return true;
  }
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t largest_free() const {
    size_t m = 0;
    for (auto &f : free_) m = std::max(m, f.second);
    return m;
  }
  size_t round(size_t b) const { return (b + align_ - 1) / align_ * align_; }

 private:
  size_t cap_, align_, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  std::map<size_t, size_t> live_;  // offset -> size
};

enum class State { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3,
                   kCancelled = 4 };

constexpr int kMaxBypass = 8;

struct Job {
  uint64_t id = 0;
  int bypassed = 0;  // later requests admitted before this one
  size_t work_bytes = 0, keep_bytes = 0;
  size_t work_off = 0, keep_off = 0;
  bool has_work = false, has_keep = false;
  State state = State::kQueued;
  bool cancel_requested = false;
  bool released = false;
  int slot = -1;
  double submit_s = 0, start_s = 0, end_s = 0;  // scheduler clock
  double updates = 0;
  std::string error;
  std::shared_ptr<void> payload;  // the executor's request data
};

inline double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double k = (v.size() - 1) * q / 100.0;
  const size_t f = (size_t)k, c = std::min(f + 1, v.size() - 1);
  return v[f] + (v[c] - v[f]) * (k - (double)f);
}

struct Stats {
  size_t submitted = 0, completed = 0, failed = 0, cancelled = 0;
  size_t queued = 0, running = 0;
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
  double updates_per_s = 0, span_s = 0;
  size_t budget_bytes = 0, used_bytes = 0, peak_bytes = 0;
};

template <class Exec>
class Scheduler {
 public:
  Scheduler(Exec *exec, size_t budget, bool backfill)
      : exec_(exec), arena_(budget), backfill_(backfill),
        t0_(std::chrono::steady_clock::now()),
        running_(exec->slots(), nullptr) {}

  ~Scheduler() { stop(); }

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                         t0_).count();
  }

  // Starts the scheduling thread (without it, step() drives the loop).
  void start_thread() {
    std::lock_guard<std::mutex> g(mu_);
    if (thread_.joinable()) return;
    stop_ = false;
    thread_ = std::thread([this] { loop(); });
  }
  // Waits for every admitted request, then stops the thread.
  void stop() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
  }

  // 0 and the id, or 1 with `error` when the footprint exceeds the budget.
  int submit(size_t work, size_t keep, double updates,
             std::shared_ptr<void> payload, uint64_t *id, std::string *error) {
    std::lock_guard<std::mutex> g(mu_);
    const size_t need = arena_.round(work) + arena_.round(keep);
    if (need > arena_.capacity()) {
      *error = "request needs " + std::to_string(need) +
               " bytes of device memory; the server's budget is " +
               std::to_string(arena_.capacity());
      return 1;
    }
    auto j = std::make_shared<Job>();
    j->id = ++next_id_;
    j->work_bytes = work;
    j->keep_bytes = keep;
    j->updates = updates;
    j->payload = std::move(payload);
    j->submit_s = now();
    jobs_[j->id] = j;
    queue_.push_back(j);
    ++submitted_;
    *id = j->id;
    cv_.notify_all();
    return 0;
  }

  // A queued request is dropped at once; a running one finishes and is then
  // reported cancelled (kernels in flight are not interrupted).
  bool cancel(uint64_t id) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return false;
    Job &j = *it->second;
    if (j.state == State::kQueued) {
      queue_.erase(std::find(queue_.begin(), queue_.end(), it->second));
      j.state = State::kCancelled;
      j.end_s = now();
      ++cancelled_;
      cv_.notify_all();
    } else if (j.state == State::kRunning) {
      j.cancel_requested = true;
    }
    return true;
  }

  // Frees a finished request's result memory and forgets the request.
  bool release(uint64_t id) {
    std::shared_ptr<Job> j;
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = jobs_.find(id);
      if (it == jobs_.end()) return false;
      j = it->second;
      if (j->state == State::kQueued || j->state == State::kRunning)
        return false;
      if (j->has_keep) arena_.release(j->keep_off);
      j->has_keep = false;
      j->released = true;
      jobs_.erase(it);
    }
    exec_->drop(*j);
    cv_.notify_all();
    return true;
  }

  // Blocks until the request leaves the queued/running states or the timeout
  // (seconds, < 0: none) passes; returns a copy of the job, or nullptr for an
  // unknown id.
  std::shared_ptr<Job> wait(uint64_t id, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return nullptr;
    auto j = it->second;
    auto pending = [&] {
      return j->state == State::kQueued || j->state == State::kRunning;
    };
    if (timeout_s < 0) {
      cv_.wait(lk, [&] { return !pending(); });
    } else {
      cv_.wait_for(lk, std::chrono::duration<double>(timeout_s),
                   [&] { return !pending(); });
    }
    return std::make_shared<Job>(*j);
  }

  std::shared_ptr<Job> info(uint64_t id) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) return nullptr;
    return std::make_shared<Job>(*it->second);
  }

  Stats stats() {
    std::lock_guard<std::mutex> g(mu_);
    Stats s;
    s.submitted = submitted_;
    s.completed = lat_.size();
    s.failed = failed_;
    s.cancelled = cancelled_;
    s.queued = queue_.size();
    for (auto &r : running_) s.running += r != nullptr;
    s.p50_ms = 1e3 * percentile(lat_, 50);
    s.p99_ms = 1e3 * percentile(lat_, 99);
    double sum = 0, mx = 0;
    for (double v : lat_) {
      sum += v;
      mx = std::max(mx, v);
    }
    s.mean_ms = lat_.empty() ? 0 : 1e3 * sum / lat_.size();
    s.max_ms = 1e3 * mx;
    s.span_s = last_end_ - first_submit_done_;
    s.updates_per_s = s.span_s > 0 ? updates_done_ / s.span_s : 0;
    s.budget_bytes = arena_.capacity();
    s.used_bytes = arena_.used();
    s.peak_bytes = arena_.peak();
    return s;
  }

  // One scheduling pass: finish completed requests, admit what fits.
  // Returns true when anything is running or queued.
  bool step() {
    poll_running();
    admit();
    std::lock_guard<std::mutex> g(mu_);
    if (!queue_.empty()) return true;
    for (auto &r : running_)
      if (r) return true;
    return false;
  }

 private:
  void loop() {
    for (;;) {
      const bool busy = step();
      std::unique_lock<std::mutex> lk(mu_);
      bool any_running = false;
      for (auto &r : running_) any_running |= r != nullptr;
      if (stop_ && !any_running) {
        // what is still queued cannot start (unreleased results hold the
        // budget): report it cancelled rather than wait forever
        for (auto &q : queue_) {
          q->state = State::kCancelled;
          q->end_s = now();
          ++cancelled_;
        }
        queue_.clear();
        cv_.notify_all();
        break;
      }
      if (!busy) {
        cv_.wait_for(lk, std::chrono::milliseconds(50));
      } else if (any_running) {
        // completions are polled; a short sleep keeps latency low
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(100));
      }
    }
  }

  void poll_running() {
    for (size_t slot = 0; slot < running_.size(); ++slot) {
      std::shared_ptr<Job> j;
      {
        std::lock_guard<std::mutex> g(mu_);
        j = running_[slot];
      }
      if (!j) continue;
      const int r = exec_->poll(*j, (int)slot);
      if (r == 0) continue;
      exec_->finish(*j);
      std::lock_guard<std::mutex> g(mu_);
      running_[slot] = nullptr;
      j->end_s = now();
      if (j->has_work) arena_.release(j->work_off);
      j->has_work = false;
      if (r < 0) {
        j->state = State::kFailed;
        ++failed_;
        drop_keep(*j);
      } else if (j->cancel_requested) {
        j->state = State::kCancelled;
        ++cancelled_;
        drop_keep(*j);
      } else {
        j->state = State::kDone;
        lat_.push_back(j->end_s - j->submit_s);
        updates_done_ += j->updates;
        if (lat_.size() == 1) first_submit_done_ = j->submit_s;
        first_submit_done_ = std::min(first_submit_done_, j->submit_s);
        last_end_ = std::max(last_end_, j->end_s);
      }
      cv_.notify_all();
    }
  }

  void drop_keep(Job &j) {
    if (j.has_keep) arena_.release(j.keep_off);
    j.has_keep = false;
  }

  void admit() {
    for (;;) {
      std::shared_ptr<Job> j;
      int slot = -1;
      {
        std::lock_guard<std::mutex> g(mu_);
        for (size_t s = 0; s < running_.size(); ++s)
          if (!running_[s]) {
            slot = (int)s;
            break;
          }
        if (slot < 0 || queue_.empty()) return;
        for (auto it = queue_.begin(); it != queue_.end(); ++it) {
          Job &c = **it;
          size_t wo = 0, ko = 0;
          const bool w = c.work_bytes > 0, k = c.keep_bytes > 0;
          if (!w || arena_.alloc(c.work_bytes, &wo)) {
            if (!k || arena_.alloc(c.keep_bytes, &ko)) {
              c.work_off = wo;
              c.keep_off = ko;
              c.has_work = w;
              c.has_keep = k;
              j = *it;
              queue_.erase(it);
              break;
            }
            if (w) arena_.release(wo);
          }
          if (!backfill_ || c.bypassed >= kMaxBypass) break;
        }
        if (!j) return;
        for (auto &q : queue_) {  // every request still ahead of j
          if (q->id > j->id) break;
          ++q->bypassed;
        }
        j->state = State::kRunning;
        j->slot = slot;
        j->start_s = now();
        running_[slot] = j;
      }
      if (!exec_->start(*j, slot)) {
        exec_->finish(*j);
        std::lock_guard<std::mutex> g(mu_);
        running_[slot] = nullptr;
        j->state = State::kFailed;
        j->end_s = now();
        if (j->has_work) arena_.release(j->work_off);
        j->has_work = false;
        drop_keep(*j);
        ++failed_;
        cv_.notify_all();
      }
    }
  }

  Exec *exec_;
  Arena arena_;
  bool backfill_;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::thread thread_;
  bool stop_ = false;
  uint64_t next_id_ = 0;
  std::deque<std::shared_ptr<Job>> queue_;
  std::vector<std::shared_ptr<Job>> running_;
  std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs_;
  std::vector<double> lat_;
  size_t submitted_ = 0, failed_ = 0, cancelled_ = 0;
  double updates_done_ = 0, first_submit_done_ = 0, last_end_ = 0;
};

}  // namespace serve
}  // namespace gx
