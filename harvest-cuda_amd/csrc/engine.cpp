// Host engine: WORLD's parameter derivation, the backend call, and the
// contour stages in O(F) (voiced sections on several threads for long inputs).  Reference counterpart: src/engine.cpp
// (analyze :240-333, contour adapters :55-214).  This file is built with
// -ffp-contract=off; every floating-point expression below keeps WORLD's
// evaluation order because each one feeds a discrete decision.
#include "engine.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <thread>

#include <emmintrin.h>
#include <unistd.h>

namespace harvest {
namespace {

constexpr double kLog2 = 0.69314718055994529;  // constantnumbers.h:24

int mround(double x) { return x > 0 ? (int)(x + 0.5) : (int)(x - 0.5); }

// Contours shorter than this stay on the calling thread: spawning threads
// costs more than the per-section work they would share.
constexpr int kParallelFrames = 16384;
// Threads per utterance for its sections; more gain nothing on ~100 sections.
constexpr int kSectionThreads = 8;

// A persistent pool of host threads for the per-section work of long
// contours (thread creation per call would cost as much as the work).  One
// job runs at a time: a caller that finds the pool busy (another call, or a
// job nested inside a job) runs its loop on its own thread instead, so no
// call ever waits for workers that are waiting for it.
class HostPool {
 public:
  static HostPool& instance() {
    static HostPool pool;
    return pool;
  }

  // Runs body(i) for i in [0, n) on up to `threads` threads, the caller
  // included.  The first exception thrown by body is rethrown once every
  // participant has finished.
  void run(int n, int threads, const std::function<void(int)>& body) {
    threads = std::max(1, std::min(threads, n));
    std::unique_lock<std::mutex> job_lock(job_mutex_, std::try_to_lock);
    if (threads == 1 || !job_lock.owns_lock()) {
      for (int i = 0; i < n; ++i) body(i);
      return;
    }
    {
      std::unique_lock<std::mutex> lk(mutex_);
      forget_after_fork();
      while ((int)workers_.size() < threads - 1) {
        try {
          workers_.emplace_back([this] { work(); });
        } catch (...) {
          break;  // the system refused a thread: fewer participants
        }
      }
      body_ = &body;
      n_ = n;
      next_ = 0;
      error_ = nullptr;
      wanted_ = std::min(threads - 1, (int)workers_.size());
      running_ = wanted_;
      ++generation_;
    }
    wake_.notify_all();
    participate();
    std::unique_lock<std::mutex> lk(mutex_);
    done_.wait(lk, [this] { return running_ == 0; });
    body_ = nullptr;
    if (error_) std::rethrow_exception(error_);
  }

  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mutex_);
      forget_after_fork();
      stop_ = true;
    }
    wake_.notify_all();
    for (auto& t : workers_) t.join();
  }

 private:
  // A forked child inherits the thread objects but not the threads: they are
  // released without joining, and the child starts its own.
  void forget_after_fork() {
    if (owner_ == getpid()) return;
    if (!workers_.empty()) new std::vector<std::thread>(std::move(workers_));  // never joined
    workers_.clear();
    owner_ = getpid();
  }

  void participate() {
    try {
      for (int i = next_++; i < n_; i = next_++) (*body_)(i);
    } catch (...) {
      std::lock_guard<std::mutex> lk(mutex_);
      if (!error_) error_ = std::current_exception();
      next_ = n_;
    }
  }

  void work() {
    unsigned long long seen = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mutex_);
        wake_.wait(lk, [&] { return stop_ || (generation_ != seen && wanted_ > 0); });
        if (stop_) return;
        seen = generation_;
        --wanted_;
      }
      participate();
      {
        std::lock_guard<std::mutex> lk(mutex_);
        --running_;
      }
      done_.notify_all();
    }
  }

  std::mutex job_mutex_;  // held by the caller for the whole job
  std::mutex mutex_;
  std::condition_variable wake_, done_;
  std::vector<std::thread> workers_;
  const std::function<void(int)>* body_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, wanted_ = 0, running_ = 0;
  unsigned long long generation_ = 0;
  std::exception_ptr error_;
  bool stop_ = false;
  pid_t owner_ = getpid();
};

template <class F> void parallel_for(int n, int threads, F&& f) {
  if (n <= 0) return;
  if (threads <= 1 || n == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  const std::function<void(int)> body = [&](int i) { f(i); };
  HostPool::instance().run(n, threads, body);
}

// Sections of a contour as GetBoundaryList (harvest.cpp:727-743) reports
// them: entries alternate start (first voiced frame) and end (last voiced
// frame); the first and last frames count as unvoiced.
void boundary_list(const double* f0, int frames, std::vector<int>& list) {
  list.clear();
  int prev = 0;
  for (int i = 1; i < frames; ++i) {
    const int v = (i == frames - 1) ? 0 : (f0[i] > 0 ? 1 : 0);
    if (v != prev) list.push_back(i - (int)(list.size() % 2));
    prev = v;
  }
}

// FixStep1 (harvest.cpp:710-722, allowed range 0.008) and FixStep2
// (harvest.cpp:748-762) in one pass: FixStep1's value of each frame is
// written as it is found, and each voiced section of those values, as
// boundary_list delimits it (frames 0 and frames - 1 count as unvoiced), is
// zeroed once it closes if it spans fewer than 6 frame steps.
void fix_step12(const double* base, int frames, double* out) {
  int start = -1;  // first frame of the open section
  for (int i = 0; i < frames; ++i) {
    double v = 0.0;
    if (i >= 2 && base[i] != 0.0) {
      const double reference = base[i - 1] * 2 - base[i - 2];
      const bool jump = std::fabs((base[i] - reference) / reference) > 0.008 &&
                        std::fabs((base[i] - base[i - 1])) / base[i - 1] > 0.008;
      v = jump ? 0.0 : base[i];
    }
    out[i] = v;
    const bool voiced = i >= 1 && i < frames - 1 && v > 0;
    if (voiced) {
      if (start < 0) start = i;
    } else if (start >= 0) {
      if (i - 1 - start < 6) std::fill(out + start, out + i, 0.0);
      start = -1;
    }
  }
}

// One voiced section of FixStep3 together with everything ExtendF0 can write:
// at most 101 frames either side (harvest.cpp:865-873).
struct Section {
  int offset = 0, start = 0, end = 0;
  std::vector<double> values;
  double at(int frame) const {
    const int i = frame - offset;
    return i >= 0 && i < (int)values.size() ? values[i] : 0.0;
  }
};

// out[j] = s.at(j) for j in [a, b]: the section's stored values, 0 outside them.
void copy_section(const Section& s, int a, int b, double* out) {
  if (a > b) return;
  const int lo = std::max(a, s.offset), hi = std::min(b, s.offset + (int)s.values.size() - 1);
  if (lo > hi) {
    std::fill(out + a, out + b + 1, 0.0);
    return;
  }
  std::fill(out + a, out + lo, 0.0);
  std::memcpy(out + lo, s.values.data() + (lo - s.offset), (hi - lo + 1) * sizeof(double));
  std::fill(out + hi + 1, out + b + 1, 0.0);
}

// SelectBestF0 (harvest.cpp:636-650) over the nonzero candidates of a frame;
// zero candidates have error 1 and never beat the 0.18 allowance.  A
// candidate with |reference - c| > error * reference * (1 + 1e-14) has a
// rounded quotient above `error` (the margin exceeds every rounding involved),
// so it is rejected without the division; all others take WORLD's expression.
double select_best(double reference, const CandidateRows& rows, int frame) {
  double best = 0.0, error = 0.18;
  double bound = error * reference * (1.0 + 1e-14);
  for (int k = rows.offsets[frame]; k < rows.offsets[frame + 1]; ++k) {
    const double d = std::fabs(reference - rows.f0[k]);
    if (d > bound) continue;
    const double e = d / reference;
    if (e > error) continue;
    bound = e * reference * (1.0 + 1e-14);
    best = rows.f0[k];
    error = e;
  }
  return best;
}

// ExtendF0 (harvest.cpp:791-820): follow the closest candidate frame by
// frame, stop after 4 misses, return the last frame that found one.
int extend(Section& s, int origin, int last, int shift, const CandidateRows& rows) {
  double reference = s.at(origin);
  int reached = origin, misses = 0;
  const int distance = std::abs(last - origin);
  for (int i = 0; i <= distance; ++i) {
    const int frame = origin + shift * i + shift;
    const double v = select_best(reference, rows, frame);
    s.values[frame - s.offset] = v;
    if (v == 0.0) {
      ++misses;
    } else {
      reference = v;
      misses = 0;
      reached = frame;
    }
    if (misses == 4) break;
  }
  return reached;
}

// SearchScore (harvest.cpp:901-907): the largest score of a candidate equal
// to f0; zero candidates carry zero score and cannot raise it.
double search_score(double f0, const CandidateRows& rows, int frame) {
  double score = 0.0;
  for (int k = rows.offsets[frame]; k < rows.offsets[frame + 1]; ++k)
    if (f0 == rows.f0[k] && score < rows.scores[k]) score = rows.scores[k];
  return score;
}

// search_score of two values in one pass over the frame's candidates.
// The running maximum (a score replaces it only if larger, so NaN scores
// never do) is kept over the row's even and odd positions in two SSE2 lanes
// and the lanes are then combined: the maximum of the matches' non-NaN
// scores, or 0, is the same in any order.
void search_scores(double f1, double f2, const CandidateRows& rows, int frame, double& s1,
                   double& s2) {
  int k = rows.offsets[frame];
  const int end = rows.offsets[frame + 1];
  const __m128d v1 = _mm_set1_pd(f1), v2 = _mm_set1_pd(f2);
  __m128d m1 = _mm_setzero_pd(), m2 = _mm_setzero_pd();
  for (; k + 2 <= end; k += 2) {
    const __m128d f = _mm_loadu_pd(rows.f0 + k), sc = _mm_loadu_pd(rows.scores + k);
    const __m128d u1 = _mm_and_pd(_mm_cmpeq_pd(f, v1), _mm_cmpgt_pd(sc, m1));
    const __m128d u2 = _mm_and_pd(_mm_cmpeq_pd(f, v2), _mm_cmpgt_pd(sc, m2));
    m1 = _mm_or_pd(_mm_and_pd(u1, sc), _mm_andnot_pd(u1, m1));
    m2 = _mm_or_pd(_mm_and_pd(u2, sc), _mm_andnot_pd(u2, m2));
  }
  double a[2], b[2];
  _mm_storeu_pd(a, m1);
  _mm_storeu_pd(b, m2);
  s1 = a[1] > a[0] ? a[1] : a[0];
  s2 = b[1] > b[0] ? b[1] : b[0];
  for (; k < end; ++k) {
    const double f = rows.f0[k];
    if (f == f1 && s1 < rows.scores[k]) s1 = rows.scores[k];
    if (f == f2 && s2 < rows.scores[k]) s2 = rows.scores[k];
  }
}

// MakeSortedOrder (harvest.cpp:883-896): WORLD's exchange pass, kept as is.
void sorted_order(const std::vector<int>& list, int n, std::vector<int>& order) {
  order.resize(n);
  std::iota(order.begin(), order.end(), 0);
  for (int i = 1; i < n; ++i)
    for (int j = i - 1; j >= 0; --j) {
      if (list[order[j] * 2] > list[order[i] * 2]) std::swap(order[i], order[j]);
      else break;
    }
}

// FixStep3 (harvest.cpp:968-995) in O(F): each section keeps only its own
// neighbourhood instead of a full-length channel.
void fix_step3(const double* in, int frames, const CandidateRows& rows, double* out,
               std::vector<int>& list, int threads) {
  boundary_list(in, frames, list);
  const int count = (int)list.size() / 2;
  std::vector<Section> sections(count);
  if (rows.pending && count > 0) throw_status(hv_wait_candidates(rows.pending, 0));
  // Each section's ExtendF0 reads only the input and the candidates.
  parallel_for(count, frames >= kParallelFrames ? threads : 1, [&](int i) {
    Section& s = sections[i];
    const int st = list[2 * i], ed = list[2 * i + 1];
    s.offset = std::max(0, st - 101);
    s.values.assign(std::min(frames, ed + 102) - s.offset, 0.0);
    std::copy(in + st, in + ed + 1, s.values.begin() + (st - s.offset));
    s.end = extend(s, ed, std::min(frames - 2, ed + 100), 1, rows);
    s.start = extend(s, st, std::max(1, st - 100), -1, rows);
  });
  // ExtendSub (harvest.cpp:840-856): the running mean is carried across
  // sections, and selected sections move to the front in order.
  int selected = 0;
  double mean = 0.0;
  for (int i = 0; i < count; ++i) {
    Section& s = sections[i];
    for (int j = s.start; j < s.end; ++j) mean += s.at(j);
    mean /= s.end - s.start;
    if (2200.0 / mean < s.end - s.start) std::swap(sections[selected++], sections[i]);
  }
  if (selected == 0) {  // nothing merged: the input stands
    std::copy(in, in + frames, out);
    return;
  }
  // MergeF0 (harvest.cpp:937-963).  Boundary pair 0 doubles as the merge
  // cursor and is mutated in place, also when section 0 is revisited.
  std::vector<int> bl(2 * selected);
  for (int k = 0; k < selected; ++k) {
    bl[2 * k] = sections[k].start;
    bl[2 * k + 1] = sections[k].end;
  }
  std::vector<int> order;
  sorted_order(bl, selected, order);
  bool scores_ready = rows.pending == nullptr;
  copy_section(sections[0], 0, frames - 1, out);
  for (int i = 1; i < selected; ++i) {
    const int k = order[i];
    const Section& s = sections[k];
    if (bl[2 * k] - bl[1] > 0) {
      copy_section(s, bl[2 * k], bl[2 * k + 1], out);
      bl[0] = bl[2 * k];
      bl[1] = bl[2 * k + 1];
      continue;
    }
    // MergeF0Sub (harvest.cpp:912-932).
    const int st1 = bl[0], ed1 = bl[1], st2 = bl[2 * k], ed2 = bl[2 * k + 1];
    if (st1 <= st2 && ed1 >= ed2) {
      bl[1] = ed1;
      continue;
    }
    double score1 = 0.0, score2 = 0.0;
    if (!scores_ready) {
      throw_status(hv_wait_candidates(rows.pending, 1));
      scores_ready = true;
    }
    for (int j = st2; j <= ed1; ++j) {  // one pass over the row serves both searches
      double a1, a2;
      search_scores(out[j], s.at(j), rows, j, a1, a2);
      score1 += a1;
      score2 += a2;
    }
    const int from = score1 > score2 ? ed1 : st2;
    copy_section(s, from, ed2, out);
    bl[1] = ed2;
  }
}

// FixStep4 (harvest.cpp:1000-1022): bridge gaps shorter than 9 frames.
void fix_step4(const double* in, int frames, double* out, std::vector<int>& list) {
  std::copy(in, in + frames, out);
  boundary_list(in, frames, list);
  const int sections = (int)list.size() / 2;
  for (int i = 0; i < sections - 1; ++i) {
    const int distance = list[(i + 1) * 2] - list[i * 2 + 1] - 1;
    if (distance >= 9) continue;
    const double tmp0 = in[list[i * 2 + 1]] + 1;
    const double tmp1 = in[list[(i + 1) * 2]] - 1;
    const double coefficient = (tmp1 - tmp0) / (distance + 1.0);
    int count = 1;
    for (int j = list[i * 2 + 1] + 1; j <= list[(i + 1) * 2] - 1; ++j)
      out[j] = tmp0 + coefficient * count++;
  }
}

// The zero-lag smoother's second-order recursion (FilteringF0,
// harvest.cpp:1049-1074).  State (w0, w1); one step consumes one input.
struct Smoother {
  static constexpr double b0 = 0.0078202080334971724, b1 = 0.015640416066994345;
  static constexpr double a0 = 1.7347257688092754, a1 = -0.76600660094326412;
  double w0 = 0.0, w1 = 0.0;
  double step(double x) {
    const double wt = x + a0 * w0 + a1 * w1;
    const double y = b0 * wt + b1 * w0 + b0 * w1;
    w1 = w0;
    w0 = wt;
    return y;
  }
  bool same(const Smoother& o) const {
    return std::memcmp(&w0, &o.w0, sizeof(double)) == 0 &&
           std::memcmp(&w1, &o.w1, sizeof(double)) == 0;
  }
};

// Orbit of a deterministic map F on filter states (one step, or one period of
// a periodic input).  In floating point the orbit from any state is
// eventually periodic: it settles on a fixed point or, for some inputs, on a
// short limit cycle.  find_cycle returns (mu, lam): the state after mu
// iterations lies on a cycle of length lam.  The last kRing states are kept
// so that the usual short cycles are found in one pass; longer ones fall
// back to Brent's algorithm.  `limit` bounds the search.
struct Cycle {
  long mu = -1, lam = 0;  // mu < 0: none found within the limit
};

template <class F> Cycle find_cycle(const Smoother& start, long limit, F&& f) {
  constexpr int kRing = 8;
  Smoother ring[kRing];
  Smoother s = start;
  ring[0] = s;
  for (long k = 1; k <= limit; ++k) {
    f(s);
    // The ring is searched every kRing steps (and at the limit): a cycle of
    // length d <= kRing is then seen at most kRing - 1 steps later, with the
    // same d (the smallest match is the cycle length) and the same entry.
    const bool scan = (k % kRing) == 0 || k == limit;
    for (int d = 1; scan && d <= kRing && d <= k; ++d)
      if (ring[(k - d) % kRing].same(s)) {
        // State k equals state k - d: cycle of length d entered at or
        // before k - d; locate its first entry.
        Cycle c;
        c.lam = d;
        Smoother a = start, b = start;
        for (int q = 0; q < d; ++q) f(b);
        long mu = 0;
        while (!a.same(b)) {
          f(a);
          f(b);
          ++mu;
        }
        c.mu = mu;
        return c;
      }
    ring[k % kRing] = s;
  }
  // Brent's algorithm for longer cycles.
  long power = 1, lam = 1;
  Smoother tortoise = start, hare = start;
  f(hare);
  while (!tortoise.same(hare)) {
    if (power == lam) {
      tortoise = hare;
      power *= 2;
      lam = 0;
    }
    f(hare);
    ++lam;
    if (power > limit) return Cycle{};
  }
  Smoother a = start, b = start;
  for (long q = 0; q < lam; ++q) f(b);
  long mu = 0;
  while (!a.same(b)) {
    f(a);
    f(b);
    ++mu;
  }
  Cycle c;
  c.mu = mu;
  c.lam = lam;
  return c;
}

// F applied `count` times to s, exactly, skipping whole cycles.
template <class F> Smoother iterate(Smoother s, long count, F&& f) {
  if (count > 512) {
    const Cycle c = find_cycle(s, count, f);
    if (c.mu >= 0 && count > c.mu) count = c.mu + (count - c.mu) % c.lam;
  }
  for (long k = 0; k < count; ++k) f(s);
  return s;
}

}  // namespace

void throw_status(hv_status status) {
  if (status == HV_OK) return;
  std::string message = hv_last_error();
  if (message.empty()) message = "Harvest backend error " + std::to_string((int)status);
  if (status == HV_EINVAL) throw std::invalid_argument(message);
  throw std::runtime_error(message);
}

Params derive_params(int64_t length, int fs, double f0_floor, double f0_ceil,
                     double frame_period) {
  Params p{};
  const double af = f0_floor * 0.9, ac = f0_ceil * 1.1;
  p.bands = 1 + (int)(std::log(ac / af) / kLog2 * 40.0);
  p.ratio = std::max(std::min(mround(fs / 8000.0), 12), 1);
  p.y_length = (int)std::ceil((double)length / p.ratio);
  p.fs_work = (double)fs / p.ratio;
  const double b0 = af * std::pow(2.0, 1 / 40.0);
  const int sample = p.y_length + 5 + 2 * (int)(2.0 * p.fs_work / b0);
  p.fft_size = (int)std::pow(2.0, (int)(std::log((double)sample) / kLog2) + 1.0);
  p.frames = (int)(1000.0 * (double)length / fs / 1.0) + 1;
  p.out_frames = (int)(1000.0 * (double)length / fs / frame_period) + 1;
  p.capacity = mround(p.bands / 10.0);
  return p;
}

void validate_params(int64_t length, int fs, double f0_floor, double f0_ceil,
                     double frame_period) {
  if (length <= 0 || length > INT_MAX)
    throw std::invalid_argument("x must be nonempty and contain at most INT_MAX samples");
  if (fs < 8000 || fs > 192000)
    throw std::invalid_argument("fs must be between 8000 and 192000 Hz");
  if (!std::isfinite(f0_floor) || !std::isfinite(f0_ceil) || f0_floor <= 0 || f0_ceil <= f0_floor)
    throw std::invalid_argument("require finite 0 < f0_floor < f0_ceil");
  const int ratio = std::min(12, std::max(1, (int)(fs / 8000.0 + 0.5)));
  if (length < 2LL * ratio)
    throw std::invalid_argument(
        "x is too short: require at least two samples after internal decimation");
  if (f0_ceil >= (double)fs / ratio / 2)
    throw std::invalid_argument("f0_ceil must be below the internal Nyquist frequency");
  if (!std::isfinite(frame_period) || frame_period <= 0)
    throw std::invalid_argument("frame_period must be finite and positive (milliseconds)");
  if (1000.0 * length / fs / frame_period >= INT_MAX || 1000.0 * length / fs >= INT_MAX ||
      (double)length / ratio + 10.0 * fs / ratio / f0_floor >= (1 << 28))
    throw std::invalid_argument("input or analysis settings exceed supported allocation sizes");
}

void validate_samples(const double* x, int64_t length) {
  bool finite = true;
  for (int64_t i = 0; i < length; ++i) finite &= std::isfinite(x[i]);
  if (!finite) throw std::invalid_argument("x must contain only finite samples");
}

void validate(const double* x, int64_t length, int fs, double f0_floor, double f0_ceil,
              double frame_period) {
  validate_params(length, fs, f0_floor, f0_ceil, frame_period);
  validate_samples(x, length);
}

void fix_contour(const CandidateRows& rows, int frames, const double* base, double* out,
                 int threads) {
  // Every element of both is written before it is read.
  std::unique_ptr<double[]> step2(new double[frames]), step3(new double[frames]);
  std::vector<int> list;
  list.reserve(64);
  fix_step12(base, frames, step2.get());
  fix_step3(step2.get(), frames, rows, step3.get(), list, threads);
  fix_step4(step3.get(), frames, out, list);
}

// SmoothF0Contour (harvest.cpp:1079-1113).  WORLD pads the contour with 300
// zeros on each side and, for every voiced section, runs the zero-phase
// filter over the whole padded length with the section's end values held
// constant outside it.  Outside the section the filter sees constant (or,
// in the backward pass, eventually periodic) input, so those stretches are
// advanced exactly with iterate() instead of sample by sample: every output
// stays bit-identical to WORLD's, in O(section + transient) per section.
void smooth_contour(const double* f0, int frames, double* out, int threads) {
  constexpr int lag = 300;
  const long n = (long)frames + 2 * lag;
  auto padded = [&](long i) { return i >= lag && i < lag + frames ? f0[i - lag] : 0.0; };
  for (int i = 0; i < frames; ++i) out[i] = 0.0;
  // Voiced sections [st, ed] of the padded contour; each writes only its own
  // outputs, so they run independently.
  std::vector<std::pair<long, long>> sections;
  for (long i = 1; i < n - 1;) {
    if (!(padded(i) > 0)) {
      ++i;
      continue;
    }
    long ed = i;
    while (ed + 1 < n - 1 && padded(ed + 1) > 0) ++ed;
    sections.emplace_back(i, ed);
    i = ed + 1;
  }
  parallel_for((int)sections.size(), frames >= kParallelFrames ? threads : 1, [&](int si) {
    const long st = sections[si].first, ed = sections[si].second;
    std::vector<double> section_out, tail;
    // Forward pass: st steps of the constant x[st] from rest, then the section.
    const double head = padded(st);
    Smoother fwd = iterate(Smoother{}, st, [&](Smoother& s) { s.step(head); });
    section_out.resize(ed - st + 1);
    for (long k = st; k <= ed; ++k) section_out[k - st] = fwd.step(padded(k));
    // Forward suffix: `steps` steps of the constant x[ed].  Its outputs are a
    // function of the state, so once the state orbit enters its cycle (after
    // mu steps, period lam) the outputs repeat with period lam.
    const long steps = n - 1 - ed;
    const double held = padded(ed);
    auto fstep = [&](Smoother& s) { s.step(held); };
    long mu = steps, lam = 1;
    if (steps > 512) {
      const Cycle c = find_cycle(fwd, steps, fstep);
      if (c.mu >= 0) {
        mu = std::min(steps, c.mu);
        lam = c.lam;
      }
    }
    const long explicit_steps = std::min(steps, mu + lam);
    tail.resize(explicit_steps);
    {
      Smoother s = fwd;
      for (long k = 0; k < explicit_steps; ++k) tail[k] = s.step(held);
    }
    auto suffix_out = [&](long k) {  // forward output at padded index ed + 1 + k
      return k < mu ? tail[k] : tail[mu + (k - mu) % lam];
    };
    // Backward pass from the padded end (suffix outputs in reverse order).
    Smoother bwd;
    long k = steps - 1;
    if (steps > explicit_steps) {
      // Inputs k = steps-1 down to mu are periodic: advance whole periods.
      const long periodic = steps - mu;
      const long periods = periodic / lam;
      std::vector<double> period(lam);
      for (long q = 0; q < lam; ++q) period[q] = suffix_out(steps - 1 - q);
      bwd = iterate(bwd, periods, [&](Smoother& s) {
        for (long q = 0; q < lam; ++q) s.step(period[q]);
      });
      k = steps - 1 - periods * lam;
    }
    for (; k >= 0; --k) bwd.step(suffix_out(k));
    for (long q = ed; q >= st; --q) out[q - lag] = bwd.step(section_out[q - st]);
  });
}

void fix_contour_split(const CandidateRows& rows, int frames, const double* base, double* step3,
                       double* out, int threads) {
  std::unique_ptr<double[]> step2(new double[frames]);
  std::vector<int> list;
  fix_step12(base, frames, step2.get());
  fix_step3(step2.get(), frames, rows, step3, list, threads);
  fix_step4(step3, frames, out, list);
}

void fix_step4_apply(const double* step3, int frames, double* out) {
  std::vector<int> list;
  fix_step4(step3, frames, out, list);
}

// FixStep4 is the identity except on bridged gaps, where inner frame j = a +
// c (c = 1..d, d = b - a - 1) gets (in[a] + 1) + ((in[b] - 1) - (in[a] + 1))
// c / (d + 1): affine in in[a] and in[b].
void fix_step4_adjoint(const double* step3, int frames, const double* g_out, double* g_in) {
  std::vector<int> list;
  boundary_list(step3, frames, list);
  std::copy(g_out, g_out + frames, g_in);
  const int sections = (int)list.size() / 2;
  for (int i = 0; i < sections - 1; ++i) {
    const int a = list[i * 2 + 1], b = list[(i + 1) * 2];
    const int distance = b - a - 1;
    if (distance >= 9) continue;
    for (int j = a + 1, c = 1; j <= b - 1; ++j, ++c) {
      const double t = c / (distance + 1.0);
      g_in[a] += g_out[j] * (1.0 - t);
      g_in[b] += g_out[j] * t;
      g_in[j] = 0.0;
    }
  }
}

// Transpose of SmoothF0Contour.  Per voiced section [st, ed] of the padded
// contour the forward map is Sel R F R F E: E extends the section by its end
// values over the padded length, F is the recursion, R reverses, Sel keeps
// [st, ed].  Since F^T = R F R, the transpose is E^T R F R F Sel^T: scatter the
// section's output gradient, run the same two passes, and fold the extension
// back onto the end values.  Zero inputs leave a zero state at rest, and a
// decaying state reaches exact zero, so each pass runs only while nonzero.
void smooth_adjoint(const double* f0, int frames, const double* g_out, double* g_in) {
  constexpr int lag = 300;
  const long n = (long)frames + 2 * lag;
  auto padded = [&](long i) { return i >= lag && i < lag + frames ? f0[i - lag] : 0.0; };
  for (int i = 0; i < frames; ++i) g_in[i] = 0.0;
  std::vector<double> u(n);
  // One causal pass of the recursion over u[lo, n) (input zero below lo),
  // in place; returns the index after the last nonzero output.
  auto pass = [&](long lo, long hi_input) {
    Smoother s;
    long last = lo;
    for (long k = lo; k < n; ++k) {
      if (k >= hi_input && s.w0 == 0.0 && s.w1 == 0.0) break;
      const double in = k < hi_input ? u[k] : 0.0;
      u[k] = s.step(in);
      if (u[k] != 0.0) last = k + 1;
    }
    for (long k = last; k < n; ++k) u[k] = 0.0;
    return last;
  };
  long i = 1;
  while (i < n - 1) {
    if (!(padded(i) > 0)) {
      ++i;
      continue;
    }
    const long st = i;
    long ed = i;
    while (ed + 1 < n - 1 && padded(ed + 1) > 0) ++ed;
    i = ed + 1;
    std::fill(u.begin(), u.end(), 0.0);
    for (long k = st; k <= ed; ++k) u[k] = g_out[k - lag];
    const long end1 = pass(st, ed + 1);           // F over [st, ...)
    std::reverse(u.begin(), u.end());            // R
    const long end2 = pass(n - end1, n - st);    // F; nonzero inputs end at n - 1 - st
    std::reverse(u.begin(), u.end());            // R
    (void)end2;
    // E^T: the extension copies x[st] to [0, st) and x[ed] to (ed, n).
    double head = 0.0, tail = 0.0;
    for (long k = 0; k <= st; ++k) head += u[k];
    for (long k = ed; k < n; ++k) tail += u[k];
    if (st == ed) {
      g_in[st - lag] += head + tail - u[st];
    } else {
      g_in[st - lag] += head;
      g_in[ed - lag] += tail;
      for (long k = st + 1; k < ed; ++k) g_in[k - lag] += u[k];
    }
  }
}

void resample(const double* basic, int frames, double frame_period, int out_frames, double* f0,
              double* times) {
  for (int i = 0; i < out_frames; ++i) {
    times[i] = i * frame_period / 1000.0;
    f0[i] = basic[std::min(frames - 1, mround(times[i] * 1000.0))];
  }
}

void resample_indices(int frames, double frame_period, int out_frames, int* index) {
  for (int i = 0; i < out_frames; ++i)
    index[i] = std::min(frames - 1, mround(i * frame_period / 1000.0 * 1000.0));
}

namespace {
using Clock = std::chrono::steady_clock;

// Device spans and kernel sums (slot meaning in include/harvest_hip.h).
void record_device(const double* ms, Timings& t) {
  static const char* names[HV_TIMING_SLOTS] = {
      "preprocess",          "filter_events",  "refine",        "kernel_decimate",
      "kernel_filterbank",   "kernel_interp",  "kernel_refine", "kernel_prune_compact",
      "kernel_refine_dft"};
  for (int k = 0; k < HV_TIMING_SLOTS; ++k) t[names[k]] = ms[k];
}

void record_stats(hv_workspace* ws, Timings& t) {
  int64_t st[HV_STAT_SLOTS];
  if (hv_last_stats(ws, st) != HV_OK) return;
  t["work_window_samples"] = (double)st[0];
  t["work_bin_samples"] = (double)st[1];
  t["work_bins"] = (double)st[2];
  t["work_table_pairs"] = (double)st[3];
}

double ms_since(Clock::time_point& mark) {
  auto now = Clock::now();
  double ms = std::chrono::duration<double, std::milli>(now - mark).count();
  mark = now;
  return ms;
}

void finish_utterance(const CandidateRows& rows, int frames, const double* best,
                      double frame_period, int out_frames, Result& r, HostArena* arena,
                      Timings* timings, Clock::time_point* mark, int threads) {
  HostArray<double> contour(arena), basic(arena);
  contour.allocate(frames);
  basic.allocate(frames);
  fix_contour(rows, frames, best, contour.data(), threads);
  if (timings) (*timings)["contour"] += ms_since(*mark);
  smooth_contour(contour.data(), frames, basic.data(), threads);
  if (timings) (*timings)["smooth"] += ms_since(*mark);
  r.f0.resize(out_frames);
  r.times.resize(out_frames);
  resample(basic.data(), frames, frame_period, out_frames, r.f0.data(), r.times.data());
  if (timings) (*timings)["resample"] += ms_since(*mark);
}

// Results of the last analysis on `ws`, read in place from pinned memory.
struct View {
  const int32_t* offsets;
  const double* best;
  const double* f0;
  const double* scores;
  int64_t nnz;
};

View result_view(hv_workspace* ws) {
  View v{};
  throw_status(hv_result_view(ws, &v.offsets, &v.best, &v.f0, &v.scores, &v.nnz));
  return v;
}
}  // namespace

Result analyze(hv_workspace* ws, const double* x, int64_t length, int fs, double f0_floor,
               double f0_ceil, double frame_period, int host_threads, Timings* timings) {
  Clock::time_point mark = Clock::now();
  if (timings) timings->clear();
  const Params p = derive_params(length, fs, f0_floor, f0_ceil, frame_period);
  HostCache::Lease lease;
  if (timings) (*timings)["initialize"] = ms_since(mark);
  int64_t nnz = 0;
  double device_ms[HV_TIMING_SLOTS];
  const int64_t len = length;
  // The candidate download overlaps FixStep1-2 and the f0-only part of
  // FixStep3 (the contour waits for each part where it first needs it).
  throw_status(hv_analyze_batch_async(ws, 1, &x, &len, fs, f0_floor, f0_ceil, nullptr, nullptr,
                                      &nnz, timings ? device_ms : nullptr));
  const View v = result_view(ws);
  if (timings) {
    ms_since(mark);
    record_device(device_ms, *timings);
    record_stats(ws, *timings);
  }
  Result r;
  CandidateRows rows{v.offsets, v.f0, v.scores, ws};
  finish_utterance(rows, p.frames, v.best, frame_period, p.out_frames, r, lease.arena(), timings,
                   &mark, std::min(host_threads, kSectionThreads));
  return r;
}

namespace {
template <class Call>
std::vector<Result> batch_common(hv_workspace* ws, Call&& call,
                                 const std::vector<int64_t>& lengths, int fs, double f0_floor,
                                 double f0_ceil, double frame_period, int host_threads,
                                 Timings* timings) {
  Clock::time_point mark = Clock::now();
  if (timings) timings->clear();
  const int count = (int)lengths.size();
  std::vector<Params> params(count);
  std::vector<int64_t> frame_off(count + 1, 0);
  for (int u = 0; u < count; ++u) {
    params[u] = derive_params(lengths[u], fs, f0_floor, f0_ceil, frame_period);
    frame_off[u + 1] = frame_off[u] + params[u].frames;
  }
  int64_t nnz = 0;
  double device_ms[HV_TIMING_SLOTS];
  throw_status(call(nullptr, nullptr, &nnz, timings ? device_ms : nullptr));
  const View v = result_view(ws);
  if (timings) {
    (*timings)["device_call"] = ms_since(mark);
    record_device(device_ms, *timings);
    record_stats(ws, *timings);
  }
  std::vector<Result> results(count);
  // Longest utterance first over a small pool of host threads.
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return params[a].frames > params[b].frames; });
  // Threads left over when there are fewer utterances than threads work
  // inside each utterance (per voiced section).
  const int inner = std::max(1, std::min(kSectionThreads, host_threads / std::max(1, count)));
  parallel_for(count, host_threads, [&](int k) {
    const int u = order[k];
    CandidateRows rows{v.offsets + frame_off[u], v.f0, v.scores, ws};
    finish_utterance(rows, params[u].frames, v.best + frame_off[u], frame_period,
                     params[u].out_frames, results[u], nullptr, nullptr, nullptr, inner);
  });
  if (timings) (*timings)["contour"] = ms_since(mark);
  return results;
}
}  // namespace

namespace {
std::mutex g_tuning_mutex;
BatchTuning g_tuning;
}  // namespace

void set_batch_tuning(const BatchTuning& t) {
  if (t.max_groups < 1 || t.min_group_frames < 1 || t.split_frames < 0 || !(t.first_weight > 0) ||
      !(t.last_weight > 0) || !std::isfinite(t.first_weight) || !std::isfinite(t.last_weight) ||
      (t.serialize != 0 && t.serialize != 1) || t.depth < 2 || t.depth > 8)
    throw std::invalid_argument("invalid batch tuning");
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  g_tuning = t;
}

BatchTuning batch_tuning() {
  std::lock_guard<std::mutex> lock(g_tuning_mutex);
  return g_tuning;
}

int pipeline_groups(const std::vector<int64_t>& frames) {
  const BatchTuning t = batch_tuning();
  int64_t total = 0;
  for (int64_t f : frames) total += f;
  if (frames.size() < 4 || total < t.split_frames) return 1;
  return (int)std::max<int64_t>(
      1, std::min<int64_t>({(int64_t)t.max_groups, (int64_t)frames.size(),
                            total / t.min_group_frames}));
}

namespace {
// Length buckets of a batch: `order` lists the utterances longest first (or
// in the given order for back-to-back device input), and group g holds
// order[bounds[g] .. bounds[g + 1]), a run of about total / groups frames.
struct Buckets {
  std::vector<int> order, bounds;
};

Buckets make_buckets(const std::vector<int64_t>& frames, int groups, bool by_length) {
  Buckets b;
  const int count = (int)frames.size();
  b.order.resize(count);
  std::iota(b.order.begin(), b.order.end(), 0);
  if (by_length)
    std::stable_sort(b.order.begin(), b.order.end(),
                     [&](int x, int y) { return frames[x] > frames[y]; });
  int64_t total = 0;
  for (int64_t f : frames) total += f;
  // Cumulative frame shares of the group boundaries.
  const BatchTuning t = batch_tuning();
  std::vector<double> weight(groups, 1.0);
  if (groups > 1) {
    weight[0] = t.first_weight;
    weight[groups - 1] = t.last_weight;
  }
  double sum = 0.0;
  for (double w : weight) sum += w;
  std::vector<double> edge(groups, 0.0);  // edge[g]: share before group g
  for (int g = 1; g < groups; ++g) edge[g] = edge[g - 1] + weight[g - 1] / sum;
  b.bounds = {0};
  int64_t acc = 0;
  for (int k = 0; k < count; ++k) {
    acc += frames[b.order[k]];
    const int g = (int)b.bounds.size();
    if (g < groups && (double)acc >= edge[g] * (double)total && k + 1 < count)
      b.bounds.push_back(k + 1);
  }
  b.bounds.push_back(count);
  return b;
}

// Sizes each workspace of a pipelined batch for the largest of the groups it
// will run (group g runs on ws[g % depth]) before any is submitted, so that
// the submissions neither allocate nor free.
void reserve_groups(const std::vector<hv_workspace*>& ws, const Buckets& bk,
                    const std::vector<int64_t>& lengths, int fs, double f0_floor,
                    double f0_ceil, bool host_input) {
  const int groups = (int)bk.bounds.size() - 1, depth = (int)ws.size();
  for (int w = 0; w < depth && w < groups; ++w) {
    std::vector<int> counts;
    std::vector<int64_t> lens;
    for (int g = w; g < groups; g += depth) {
      counts.push_back(bk.bounds[g + 1] - bk.bounds[g]);
      for (int k = bk.bounds[g]; k < bk.bounds[g + 1]; ++k) lens.push_back(lengths[bk.order[k]]);
    }
    throw_status(hv_reserve_batches(ws[w], (int)counts.size(), counts.data(), lens.data(), fs,
                                    f0_floor, f0_ceil, host_input ? 1 : 0));
  }
}

// Adds one group's device spans, kernel sums and work counters to `t`.
void accumulate_device(hv_workspace* ws, const double* ms, Timings& t) {
  Timings one;
  record_device(ms, one);
  record_stats(ws, one);
  for (const auto& kv : one) t[kv.first] += kv.second;
}

// `submit(w, idx, n, timed)` enqueues utterances idx[0 .. n) on workspace w.
// Groups alternate between the two workspaces: while one group's kernels
// run, the host stages the next group's input and runs the previous group's
// contours.
template <class Submit>
std::vector<Result> batch_pipelined(const std::vector<hv_workspace*>& ws, Submit&& submit,
                                    const std::vector<int64_t>& lengths, const Buckets& bk,
                                    int fs, double f0_floor, double f0_ceil, double frame_period,
                                    int host_threads, Timings* timings) {
  Clock::time_point mark = Clock::now();
  if (timings) timings->clear();
  const int count = (int)lengths.size();
  std::vector<Params> params(count);
  for (int u = 0; u < count; ++u)
    params[u] = derive_params(lengths[u], fs, f0_floor, f0_ceil, frame_period);
  const int groups = (int)bk.bounds.size() - 1;
  const int depth = (int)ws.size();
  std::vector<char> pending(depth, 0);
  std::vector<Result> results(count);
  // Host time of each phase (with timings): staging and enqueueing a group,
  // waiting for its device work, and its contours.
  double host_submit = 0.0, host_wait = 0.0, host_contour = 0.0, worst_wait = 0.0;
  Clock::time_point phase = Clock::now();
  const bool serialize = batch_tuning().serialize != 0;
  auto enqueue = [&](int g) {
    if (timings) ms_since(phase);
    if (serialize && g > 0) throw_status(hv_order_after(ws[g % depth], ws[(g - 1) % depth]));
    throw_status(submit(ws[g % depth], bk.order.data() + bk.bounds[g],
                        bk.bounds[g + 1] - bk.bounds[g], timings != nullptr));
    pending[g % depth] = 1;
    if (timings) host_submit += ms_since(phase);
  };
  try {
    for (int g = 0; g < std::min(depth, groups); ++g) enqueue(g);
    for (int g = 0; g < groups; ++g) {
      hv_workspace* w = ws[g % depth];
      int64_t nnz = 0;
      double ms[HV_TIMING_SLOTS];
      pending[g % depth] = 0;
      if (timings) ms_since(phase);
      throw_status(hv_analyze_batch_collect(w, nullptr, nullptr, &nnz, timings ? ms : nullptr));
      if (timings) {
        const double waited = ms_since(phase);
        host_wait += waited;
        worst_wait = std::max(worst_wait, waited);
        accumulate_device(w, ms, *timings);
        ms_since(phase);
      }
      const View v = result_view(w);
      const int* idx = bk.order.data() + bk.bounds[g];
      const int n = bk.bounds[g + 1] - bk.bounds[g];
      std::vector<int64_t> frame_off(n + 1, 0);
      for (int k = 0; k < n; ++k) frame_off[k + 1] = frame_off[k] + params[idx[k]].frames;
      // Longest utterance first over the host pool.
      std::vector<int> by_len(n);
      std::iota(by_len.begin(), by_len.end(), 0);
      std::stable_sort(by_len.begin(), by_len.end(), [&](int a, int b) {
        return params[idx[a]].frames > params[idx[b]].frames;
      });
      const int inner = std::max(1, std::min(kSectionThreads, host_threads / std::max(1, n)));
      parallel_for(n, host_threads, [&](int k) {
        const int local = by_len[k];
        const int u = idx[local];
        CandidateRows rows{v.offsets + frame_off[local], v.f0, v.scores, w};
        finish_utterance(rows, params[u].frames, v.best + frame_off[local], frame_period,
                         params[u].out_frames, results[u], nullptr, nullptr, nullptr, inner);
      });
      if (timings) host_contour += ms_since(phase);
      // The workspace's pinned results are read: it takes the next group.
      if (g + depth < groups) enqueue(g + depth);
    }
  } catch (...) {
    // Leave every workspace usable: collect (and drop) what is still enqueued.
    for (int k = 0; k < depth; ++k) {
      int64_t nnz = 0;
      if (pending[k]) (void)hv_analyze_batch_collect(ws[k], nullptr, nullptr, &nnz, nullptr);
    }
    throw;
  }
  if (timings) {
    (*timings)["groups"] = groups;
    (*timings)["host_submit"] = host_submit;
    (*timings)["host_wait"] = host_wait;
    (*timings)["host_wait_max"] = worst_wait;
    (*timings)["host_contour"] = host_contour;
    (*timings)["call"] = ms_since(mark);
  }
  return results;
}

std::vector<int64_t> frames_of(const std::vector<int64_t>& lengths, int fs, double f0_floor,
                               double f0_ceil, double frame_period) {
  std::vector<int64_t> frames(lengths.size());
  for (size_t u = 0; u < lengths.size(); ++u)
    frames[u] = derive_params(lengths[u], fs, f0_floor, f0_ceil, frame_period).frames;
  return frames;
}
}  // namespace

std::vector<Result> analyze_batch(hv_workspace* ws, const std::vector<const double*>& xs,
                                  const std::vector<int64_t>& lengths, int fs,
                                  double f0_floor, double f0_ceil, double frame_period,
                                  int host_threads, Timings* timings,
                                  const std::vector<hv_workspace*>& more) {
  const std::vector<int64_t> frames = frames_of(lengths, fs, f0_floor, f0_ceil, frame_period);
  const int groups = pipeline_groups(frames);
  if (!more.empty() && groups > 1) {
    // Host input: any utterance order can be staged, so groups are length
    // buckets (longest first).
    const Buckets bk = make_buckets(frames, groups, true);
    auto submit = [&](hv_workspace* w, const int* idx, int n, bool timed) {
      std::vector<const double*> p(n);
      std::vector<int64_t> l(n);
      for (int k = 0; k < n; ++k) {
        p[k] = xs[idx[k]];
        l[k] = lengths[idx[k]];
      }
      return hv_analyze_batch_submit(w, n, p.data(), l.data(), fs, f0_floor, f0_ceil,
                                     timed ? 1 : 0);
    };
    std::vector<hv_workspace*> all{ws};
    all.insert(all.end(), more.begin(), more.end());
    reserve_groups(all, bk, lengths, fs, f0_floor, f0_ceil, true);
    return batch_pipelined(all, submit, lengths, bk, fs, f0_floor, f0_ceil, frame_period,
                           host_threads, timings);
  }
  auto call = [&](int32_t* offsets, double* best, int64_t* nnz, double* ms) {
    return hv_analyze_batch_async(ws, (int)xs.size(), xs.data(), lengths.data(), fs, f0_floor,
                                  f0_ceil, offsets, best, nnz, ms);
  };
  return batch_common(ws, call, lengths, fs, f0_floor, f0_ceil, frame_period, host_threads,
                      timings);
}

std::vector<Result> analyze_batch_device(hv_workspace* ws, const double* x_device,
                                         const std::vector<int64_t>& lengths, int fs,
                                         double f0_floor, double f0_ceil, double frame_period,
                                         int host_threads, Timings* timings,
                                         const std::vector<hv_workspace*>& more) {
  const std::vector<int64_t> frames = frames_of(lengths, fs, f0_floor, f0_ceil, frame_period);
  std::vector<int64_t> x_off(lengths.size() + 1, 0);
  for (size_t u = 0; u < lengths.size(); ++u) x_off[u + 1] = x_off[u] + lengths[u];
  const int groups = pipeline_groups(frames);
  if (!more.empty() && groups > 1) {
    // Back-to-back device input: groups are runs of consecutive utterances.
    const Buckets bk = make_buckets(frames, groups, false);
    auto submit = [&](hv_workspace* w, const int* idx, int n, bool timed) {
      std::vector<int64_t> l(lengths.begin() + idx[0], lengths.begin() + idx[0] + n);
      return hv_analyze_batch_device_submit(w, n, x_device + x_off[idx[0]], l.data(), fs,
                                            f0_floor, f0_ceil, timed ? 1 : 0);
    };
    std::vector<hv_workspace*> all{ws};
    all.insert(all.end(), more.begin(), more.end());
    reserve_groups(all, bk, lengths, fs, f0_floor, f0_ceil, false);
    return batch_pipelined(all, submit, lengths, bk, fs, f0_floor, f0_ceil, frame_period,
                           host_threads, timings);
  }
  auto call = [&](int32_t* offsets, double* best, int64_t* nnz, double* ms) {
    return hv_analyze_batch_device_async(ws, (int)lengths.size(), x_device, lengths.data(), fs,
                                         f0_floor, f0_ceil, offsets, best, nnz, ms);
  };
  return batch_common(ws, call, lengths, fs, f0_floor, f0_ceil, frame_period, host_threads,
                      timings);
}

}  // namespace harvest
