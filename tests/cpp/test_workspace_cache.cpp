// WorkspaceCache (workspace_cache.h) with fake workspaces on the CPU: reuse
// across threads and devices, failed creation, clear() of idle and leased
// workspaces, destruction outside the cache lock, clear() racing a creation,
// and a 4-thread stress run.  The behaviours are those the reference's
// tests/cpp/test_workspace_cache.cpp:36-131 checks for its cache.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <functional>
#include <future>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <vector>

#include "workspace_cache.h"

namespace {

struct FakeWorkspace {
  int device, serial;
};

constexpr int kDevices = 2;
std::atomic<int> alive[kDevices];
std::atomic<int> serial{0};
std::function<void()> before_create, after_destroy;  // fault and race hooks

FakeWorkspace* make(int device) {
  if (device >= kDevices) throw std::invalid_argument("no such device");
  if (before_create) before_create();
  ++alive[device];
  return new FakeWorkspace{device, ++serial};
}

void unmake(FakeWorkspace* w) {
  if (!w) return;
  --alive[w->device];
  delete w;
  if (after_destroy) after_destroy();
}

using Cache = harvest::WorkspaceCache<FakeWorkspace>;
using Lease = Cache::Lease;
int failures = 0;

void require(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

bool none_alive() { return alive[0] == 0 && alive[1] == 0; }

void reuse_and_clear() {
  Cache cache(make, unmake);
  int first = 0;
  {
    Lease l(cache, 0);
    first = l.get()->serial;
  }
  // Another thread, and a switch between devices, reuse idle workspaces.
  std::thread([&] {
    { Lease l(cache, 0); require(l.get()->serial == first, "reused from another thread"); }
    { Lease l(cache, 1); require(l.get()->device == 1, "second device gets its own"); }
    { Lease l(cache, 0); require(l.get()->serial == first, "first device unchanged"); }
  }).join();
  require(alive[0] == 1 && alive[1] == 1 && cache.size() == 2, "one workspace per device");
  bool threw = false;
  try {
    Lease l(cache, 5);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  require(threw && cache.size() == 2, "a failed creation propagates and caches nothing");
  threw = false;
  try {
    Lease l(cache, -1);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  require(threw, "negative device rejected");
  // clear() from any thread releases the idle workspaces.
  std::thread([&] { cache.clear(); }).join();
  require(none_alive() && cache.size() == 0, "clear releases idle workspaces");
  cache.clear();  // clearing an empty cache is harmless
  {
    Lease l(cache, 0);
    require(l.get()->serial != first, "a fresh workspace after clear");
  }
}

void busy_device_and_leased_clear() {
  Cache cache(make, unmake);
  int first = 0;
  std::future<int> waiter;
  {
    Lease holder(cache, 0);
    first = holder.get()->serial;
    std::promise<void> started;
    auto started_f = started.get_future();
    waiter = std::async(std::launch::async, [&] {
      started.set_value();
      Lease l(cache, 0);  // waits for `holder`
      return l.get()->serial;
    });
    started_f.wait();
    require(waiter.wait_for(std::chrono::milliseconds(50)) == std::future_status::timeout,
            "a second call on a busy device waits");
    { Lease other(cache, 1); require(other.get()->device == 1, "other devices do not wait"); }
    require(alive[0] == 1, "the waiter creates nothing");
  }
  require(waiter.get() == first && alive[0] == 1, "the waiter reuses the workspace");
  {
    Lease in_use(cache, 0);
    std::thread([&] { cache.clear(); }).join();
    require(alive[0] == 1 && in_use.get()->device == 0, "a leased workspace survives clear");
    {
      Lease newer(cache, 0);
      require(newer.get() != in_use.get() && alive[0] == 2, "a new lease gets a new workspace");
    }
  }
  require(alive[0] == 1 && cache.size() == 1,
          "the cleared workspace goes when its lease ends; it is not cached again");
  cache.clear();
  require(alive[0] == 0, "cleared");
  try {
    Lease l(cache, 0);
    throw std::runtime_error("analysis failed");
  } catch (const std::runtime_error&) {
  }
  require(alive[0] == 1 && cache.size() == 1, "a failing call returns its workspace");
  // Destruction happens outside the cache lock: a destroy hook that calls
  // clear() would deadlock otherwise.
  after_destroy = [&] { cache.clear(); };
  cache.clear();
  after_destroy = {};
  require(alive[0] == 0, "destroyed without holding the lock");
}

void clear_during_creation() {
  Cache cache(make, unmake);
  std::mutex m;
  std::condition_variable cv;
  bool creating = false, resume = false;
  before_create = [&] {
    std::unique_lock<std::mutex> lk(m);
    creating = true;
    cv.notify_all();
    cv.wait(lk, [&] { return resume; });
  };
  std::thread worker([&] {
    Lease l(cache, 0);
    require(l.get() != nullptr, "the racing creation is still usable once");
  });
  {
    std::unique_lock<std::mutex> lk(m);
    cv.wait(lk, [&] { return creating; });
  }
  cache.clear();
  {
    std::lock_guard<std::mutex> lk(m);
    resume = true;
  }
  cv.notify_all();
  worker.join();
  before_create = {};
  require(alive[0] == 0 && cache.size() == 0, "a creation racing clear() is never cached");
}

void stress() {
  Cache cache(make, unmake);
  std::vector<std::thread> workers;
  std::atomic<int> wrong{0};
  for (int t = 0; t < 4; ++t)
    workers.emplace_back([&, t] {
      for (int call = 0; call < 500; ++call) {
        Lease l(cache, t % 2);
        if (l.get()->device != t % 2) ++wrong;
        if (call % 17 == 0) cache.clear();
        std::this_thread::yield();
      }
    });
  for (auto& w : workers) w.join();
  require(wrong == 0, "every lease gets its own device's workspace");
  require(alive[0] <= 1 && alive[1] <= 1, "at most one workspace per device is cached");
}

}  // namespace

int main() {
  reuse_and_clear();
  require(none_alive(), "cache destruction releases everything (1)");
  busy_device_and_leased_clear();
  require(none_alive(), "cache destruction releases everything (2)");
  clear_during_creation();
  stress();
  require(none_alive(), "cache destruction releases everything (3)");
  if (failures) return 1;
  std::puts("workspace cache tests passed");
  return 0;
}
