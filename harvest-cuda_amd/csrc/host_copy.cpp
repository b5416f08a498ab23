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
