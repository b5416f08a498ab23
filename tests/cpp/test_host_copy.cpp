// Host test of the backend's parallel staging copy (host_copy.cpp): several
// callers at once, pieces of every size around the part length, results and
// finiteness flags checked; run plainly and under the thread and address
// sanitizers by tests/test_host_native.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
#include <thread>
#include <vector>

#include "host_copy.h"

using hv::CopyPiece;

static int failures = 0;
static void require(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}

// One caller: `rounds` copies of random pieces (some with a non-finite
// sample), each checked against the source.
static void caller(unsigned seed, int rounds, int* errors) {
  std::mt19937_64 rng(seed);
  for (int r = 0; r < rounds; ++r) {
    const int count = 1 + (int)(rng() % 5);
    std::vector<std::vector<double>> src(count), dst(count);
    std::vector<CopyPiece> pieces;
    bool want_finite = true;
    for (int i = 0; i < count; ++i) {
      const long long sizes[] = {0, 1, 3, hv::kCopyPart - 1, hv::kCopyPart, hv::kCopyPart + 1,
                                 3 * hv::kCopyPart + 7, (long long)(rng() % 300000)};
      const long long n = sizes[rng() % 8];
      src[i].resize(n);
      dst[i].assign(n + 1, -1.0);
      for (long long k = 0; k < n; ++k) src[i][k] = (double)(rng() % 1000003) - 500000.0;
      if (n > 0 && rng() % 4 == 0) {
        src[i][rng() % n] = (rng() & 1) ? std::numeric_limits<double>::quiet_NaN()
                                        : -std::numeric_limits<double>::infinity();
        want_finite = false;
      }
      // An odd offset of one double: destinations that are not 16-byte aligned.
      pieces.push_back(CopyPiece{dst[i].data() + (n & 1), src[i].data(), n});
    }
    const bool finite = hv::copy_finite_parallel(pieces.data(), (int)pieces.size());
    if (finite != want_finite) ++*errors;
    for (int i = 0; i < count; ++i) {
      const long long n = (long long)src[i].size(), off = n & 1;
      for (long long k = 0; k < n; ++k) {
        const double a = src[i][k], b = dst[i][off + k];
        if (!(a == b || (std::isnan(a) && std::isnan(b)))) {
          ++*errors;
          break;
        }
      }
      // Nothing written outside the piece.
      if (off == 1 && dst[i][0] != -1.0) ++*errors;
      if (off == 0 && dst[i][n] != -1.0) ++*errors;
    }
  }
}

int main() {
  int e0 = 0;
  caller(1, 20, &e0);
  require(e0 == 0, "single caller copies and flags");
  const int threads = 4;
  std::vector<int> errors(threads, 0);
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) ts.emplace_back(caller, 100 + t, 40, &errors[t]);
  for (auto& t : ts) t.join();
  int total = 0;
  for (int e : errors) total += e;
  require(total == 0, "concurrent callers copy and flag independently");
  CopyPiece none{nullptr, nullptr, 0};
  require(hv::copy_finite_parallel(&none, 1), "an empty piece is finite");
  require(hv::copy_finite_parallel(nullptr, 0), "no pieces is finite");
  if (failures) return 1;
  std::printf("host copy tests passed\n");
  return 0;
}
