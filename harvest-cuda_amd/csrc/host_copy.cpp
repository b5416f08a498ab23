#include "host_copy.h"

#include <emmintrin.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

namespace hv {

bool copy_finite(double* dst, const double* src, long long n) {
  bool finite = true;
  long long i = 0;
  // One scalar sample first if dst is not 16-byte aligned (streaming stores
  // need it); every sample of a double array is 8-byte aligned.
  if (n > 0 && (reinterpret_cast<uintptr_t>(dst) & 15) != 0) {
    dst[0] = src[0];
    finite = std::fabs(src[0]) <= DBL_MAX;
    i = 1;
  }
  // |v| <= DBL_MAX is false exactly for infinities and NaNs.
  const __m128d abs_mask = _mm_castsi128_pd(_mm_set1_epi64x(0x7fffffffffffffffLL));
  const __m128d limit = _mm_set1_pd(DBL_MAX);
  __m128d ok0 = _mm_cmpeq_pd(limit, limit), ok1 = ok0;
  for (; i + 4 <= n; i += 4) {
    const __m128d a = _mm_loadu_pd(src + i), b = _mm_loadu_pd(src + i + 2);
    _mm_stream_pd(dst + i, a);
    _mm_stream_pd(dst + i + 2, b);
    ok0 = _mm_and_pd(ok0, _mm_cmple_pd(_mm_and_pd(a, abs_mask), limit));
    ok1 = _mm_and_pd(ok1, _mm_cmple_pd(_mm_and_pd(b, abs_mask), limit));
  }
  for (; i < n; ++i) {
    dst[i] = src[i];
    finite = finite && std::fabs(src[i]) <= DBL_MAX;
  }
  _mm_sfence();
  return finite && _mm_movemask_pd(_mm_and_pd(ok0, ok1)) == 3;
}

}  // namespace hv

// ---------------------------------------------------------------------------
// Parallel staging copy.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace hv {
namespace {

// The parts of one copy_finite_parallel call.  Threads take part indices
// from `next` until none is left; `running` counts pool threads inside
// work() (the caller waits for it to drop to zero before returning).
struct CopyJob {
  std::vector<CopyPiece> parts;
  std::atomic<size_t> next{0};
  std::atomic<bool> finite{true};
  std::mutex m;
  std::condition_variable cv;
  int running = 0;

  void work() {
    bool ok = true;
    for (size_t k = next.fetch_add(1); k < parts.size(); k = next.fetch_add(1))
      ok &= copy_finite(parts[k].dst, parts[k].src, parts[k].n);
    if (!ok) finite.store(false);
  }
};

// Threads created on first use; each runs jobs from the queue until the pool
// is destroyed at exit.
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool pool;
    return pool;
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lock(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  // Offers `job` to up to `helpers` pool threads.
  void offer(const std::shared_ptr<CopyJob>& job, int helpers) {
    {
      std::lock_guard<std::mutex> lock(m_);
      if (threads_.empty()) {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int n = (int)std::min<unsigned>(kCopyThreads, hw);
        for (int i = 0; i < n; ++i) threads_.emplace_back([this] { loop(); });
      }
      for (int i = 0; i < helpers; ++i) queue_.push_back(job);
    }
    cv_.notify_all();
  }

 private:
  CopyPool() = default;
  void loop() {
    for (;;) {
      std::shared_ptr<CopyJob> job;
      {
        std::unique_lock<std::mutex> lock(m_);
        cv_.wait(lock, [this] { return stop_ || !queue_.empty(); });
        if (stop_) return;
        job = std::move(queue_.front());
        queue_.pop_front();
        std::lock_guard<std::mutex> jl(job->m);
        ++job->running;
      }
      job->work();
      {
        std::lock_guard<std::mutex> jl(job->m);
        --job->running;
      }
      job->cv.notify_all();
    }
  }
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<CopyJob>> queue_;
  std::vector<std::thread> threads_;
  bool stop_ = false;
};

}  // namespace

bool copy_finite_parallel(const CopyPiece* pieces, int count) {
  auto job = std::make_shared<CopyJob>();
  for (int i = 0; i < count; ++i)
    for (long long a = 0; a < pieces[i].n; a += kCopyPart) {
      const long long n = std::min(kCopyPart, pieces[i].n - a);
      job->parts.push_back(CopyPiece{pieces[i].dst + a, pieces[i].src + a, n});
    }
  if (job->parts.size() > 1)
    CopyPool::get().offer(job, (int)std::min<size_t>(kCopyThreads, job->parts.size() - 1));
  job->work();
  // Every part has been taken; wait until the pool threads that took some
  // have finished them.  Pool threads that pick the job up later find no
  // part left and only drop their reference.
  std::unique_lock<std::mutex> lock(job->m);
  job->cv.wait(lock, [&] { return job->running == 0; });
  return job->finite.load();
}

}  // namespace hv
