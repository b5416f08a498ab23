/*
 * Harvest oracle: plain-C restatement of WORLD Harvest.  TEST INFRASTRUCTURE
 * ONLY (see harvest_oracle.h for who may load it and how each function is
 * pinned).  Line citations refer to the reference's vendored WORLD sources:
 * third_party/world/src/{harvest,matlabfunctions,common}.cpp.
 *
 * Build with -O2 -ffp-contract=off and without -ffast-math: every expression
 * keeps WORLD's evaluation order so that rounding matches wherever the
 * algorithm is restated literally.
 */
#include "harvest_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define HO_PI 3.1415926535897932384          /* constantnumbers.h:18 */
#define HO_LOG2 0.69314718055994529          /* constantnumbers.h:24 */
#define HO_GUARD 0.000000000001              /* constantnumbers.h:19 */
#define HO_PI_L 3.14159265358979323846264338327950288L

/* matlab_round, matlabfunctions.cpp:206-208 */
static int mround(double x) {
  return x > 0 ? (int)(x + 0.5) : (int)(x - 0.5);
}
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* e^{2 pi i q / n} for q < n (n a power of two), each entry rounded once from
 * an extended-precision evaluation; cached per n. */
typedef struct {
  double* c;
  double* s;
} twiddles;
static twiddles g_tw[40];
static pthread_mutex_t g_tw_lock = PTHREAD_MUTEX_INITIALIZER;

static const twiddles* twiddle_table(int n) {
  int k = 0;
  while ((1 << k) < n) ++k;
  pthread_mutex_lock(&g_tw_lock);
  if (!g_tw[k].c) {
    double* c = (double*)malloc(sizeof(double) * n);
    double* s = (double*)malloc(sizeof(double) * n);
    for (int q = 0; q < n; ++q) {
      long double ang = 2.0L * HO_PI_L * (long double)q / n;
      c[q] = (double)cosl(ang);
      s[q] = (double)sinl(ang);
    }
    g_tw[k].c = c;
    g_tw[k].s = s;
  }
  pthread_mutex_unlock(&g_tw_lock);
  return &g_tw[k];
}

/* GetSuitableFFTSize, common.cpp:51-54 */
static int suitable_fft_size(int sample) {
  return (int)pow(2.0, (int)(log((double)sample) / HO_LOG2) + 1.0);
}

/* HarvestGeneralBody prologue, harvest.cpp:1149-1181, Harvest harvest.cpp:1226-1228,
   GetSamplesForHarvest harvest.cpp:1219-1221. */
int ho_params_get(int64_t length, int fs, double f0_floor, double f0_ceil,
                  double frame_period, ho_params* p) {
  double adjusted_floor = f0_floor * 0.9;
  double adjusted_ceil = f0_ceil * 1.1;
  p->bands = 1 + (int)(log(adjusted_ceil / adjusted_floor) / HO_LOG2 * 40.0);
  int ratio = mround(fs / 8000.0);
  p->ratio = imax(imin(ratio, 12), 1);
  p->y_length = (int)ceil((double)length / p->ratio);
  p->fs_work = (double)fs / p->ratio;
  double b0 = adjusted_floor * pow(2.0, 1 / 40.0);
  p->fft_size = suitable_fft_size(p->y_length + 5 +
                                  2 * (int)(2.0 * p->fs_work / b0));
  p->frames = (int)(1000.0 * (double)length / fs / 1.0) + 1;
  p->out_frames = (int)(1000.0 * (double)length / fs / frame_period) + 1;
  p->capacity = mround(p->bands / 10.0);
  p->slots = p->capacity * 7;
  return 0;
}

void ho_boundaries(const ho_params* p, double f0_floor, double* boundaries) {
  double adjusted_floor = f0_floor * 0.9;
  for (int i = 0; i < p->bands; ++i)
    boundaries[i] = adjusted_floor * pow(2.0, (i + 1) / 40.0);
}

/* ---------------------------------------------------------------------------
 * Decimation: FilterForDecimate matlabfunctions.cpp:27-125 (coefficients
 * :29-113, recursion :116-124), decimate :178-204, and the replicate padding
 * of GetWaveformAndSpectrumSub harvest.cpp:43-66.
 * Rows: a0, a1, a2, b0, b1 for ratio r (rows 0 and 1 unused).
 */
static const double kDecimation[13][5] = {
  {0, 0, 0, 0, 0},
  {0, 0, 0, 0, 0},
  {0.041156734567757189, -0.42599112459189636, 0.041037215479961225,
   0.16797464681802227, 0.50392394045406674},
  {0.95039378983237421, -0.67429146741526791, 0.15412211621346475,
   0.071221945171178636, 0.21366583551353591},
  {1.4499664446880227, -0.98943497080950582, 0.24578252340690215,
   0.036710750339322612, 0.11013225101796784},
  {1.7610939654280557, -1.2554914843859768, 0.3237186507788215,
   0.021334858522387423, 0.06400457556716227},
  {1.9715352749512141, -1.4686795689225347, 0.3893908434965701,
   0.013469181309343825, 0.040407543928031475},
  {2.1225239019534703, -1.6395144861046302, 0.44469707800587366,
   0.0090366882681608418, 0.027110064804482525},
  {2.2357462340187593, -1.7780899984041358, 0.49152555365968692,
   0.0063522763407111993, 0.019056829022133598},
  {2.3236003491759578, -1.8921545617463598, 0.53148928133729068,
   0.0046331164041389372, 0.013899349212416812},
  {2.3936475118069387, -1.9873904075111861, 0.5658879979027055,
   0.0034818622251927556, 0.010445586675578267},
  {2.450743295230728, -2.06794904601978, 0.59574774438332101,
   0.0026822508007163792, 0.0080467524021491377},
  {2.4981398605924205, -2.1368928194784025, 0.62187513816221485,
   0.0021097275904709001, 0.0063291827714127002},
};

static void iir_decimation_filter(const double* x, int n, int r, double* y) {
  const double* c = kDecimation[r];
  double w0 = 0.0, w1 = 0.0, w2 = 0.0, wt;
  for (int i = 0; i < n; ++i) {
    wt = x[i] + c[0] * w0 + c[1] * w1 + c[2] * w2;
    y[i] = c[3] * wt + c[4] * w0 + c[4] * w1 + c[3] * w2;
    w2 = w1;
    w1 = w0;
    w0 = wt;
  }
}

static void matlab_decimate(const double* x, int xl, int r, double* y) {
  const int nfact = 9;
  int m = xl + 2 * nfact;
  double* a = (double*)malloc(sizeof(double) * m);
  double* b = (double*)malloc(sizeof(double) * m);
  for (int i = 0; i < nfact; ++i) a[i] = 2 * x[0] - x[nfact - i];
  for (int i = nfact; i < nfact + xl; ++i) a[i] = x[i - nfact];
  for (int i = nfact + xl; i < m; ++i)
    a[i] = 2 * x[xl - 1] - x[xl - 2 - (i - (nfact + xl))];
  iir_decimation_filter(a, m, r, b);
  for (int i = 0; i < m; ++i) a[i] = b[m - i - 1];
  iir_decimation_filter(a, m, r, b);
  for (int i = 0; i < m; ++i) a[i] = b[m - i - 1];
  int nout = (xl - 1) / r + 1;
  int nbeg = r - r * nout + xl;
  int count = 0;
  for (int i = nbeg; i < xl + nfact; i += r) y[count++] = a[i + nfact - 1];
  free(a);
  free(b);
}

void ho_decimate(const double* x, int length, int ratio, int y_length,
                 double* y) {
  if (ratio == 1) {
    for (int i = 0; i < length; ++i) y[i] = x[i];
    return;
  }
  int lag = (int)(ceil(140.0 / ratio) * ratio);
  int padded = length + lag * 2;
  double* px = (double*)malloc(sizeof(double) * padded);
  double* py = (double*)calloc((size_t)padded, sizeof(double));
  for (int i = 0; i < lag; ++i) px[i] = x[0];
  for (int i = lag; i < lag + length; ++i) px[i] = x[i - lag];
  for (int i = lag + length; i < padded; ++i) px[i] = x[length - 1];
  matlab_decimate(px, padded, ratio, py);
  for (int i = 0; i < y_length; ++i) y[i] = py[lag / ratio + i];
  free(px);
  free(py);
}

/* GetWaveformAndSpectrum harvest.cpp:81-85 */
void ho_remove_mean(double* y, int y_length) {
  double mean = 0.0;
  for (int i = 0; i < y_length; ++i) mean += y[i];
  mean /= y_length;
  for (int i = 0; i < y_length; ++i) y[i] -= mean;
}

/* ---------------------------------------------------------------------------
 * Band-pass filter: NuttallWindow common.cpp:113-121 times the cosine
 * modulation of GetFilteredSignal harvest.cpp:101-106.
 */
static int filter_half_length(double fs, double boundary) {
  return mround(fs / boundary * 2.0);
}

static void band_pass_filter(int half, double fs, double boundary, double* c) {
  int len = half * 2 + 1;
  for (int i = 0; i < len; ++i) {
    double tmp = i / (len - 1.0);
    c[i] = 0.355768 - 0.487396 * cos(2.0 * HO_PI * tmp) +
           0.144232 * cos(4.0 * HO_PI * tmp) -
           0.012604 * cos(6.0 * HO_PI * tmp);
  }
  for (int i = -half; i <= half; ++i)
    c[i + half] *= cos(2 * HO_PI * boundary * i / fs);
}

/* GetFilteredSignal harvest.cpp:99-148, restated as a linear convolution.
 * With n > N + 2h the circular FFT convolution has no wrap, so WORLD's
 * filtered_signal[i] = n * sum_m c[m] y[i + h + 1 - m] + Q(i + h + 1), where Q
 * is the contribution of the two overwritten bins (harvest.cpp:122-133):
 * bins n/2-1 and n/2 both become S[n/2] * P[n/2-1], P = S * H; c2r weights
 * bin n/2-1 by 2 Re(. e^{2 pi i k (n/2-1)/n}) and bin n/2 by its real part
 * times (-1)^k (fft.cpp:26-35).  The oracle returns WORLD's samples divided
 * by n (a power of two, so the division is exact). */
void ho_filtered_band(const double* y, int y_length, int fft_size,
                      double fs_work, double boundary, double* out) {
  int half = filter_half_length(fs_work, boundary);
  int len = 2 * half + 1;
  double* c = (double*)malloc(sizeof(double) * len);
  band_pass_filter(half, fs_work, boundary, c);

  /* Two bins of the signal and filter spectra, FFTW sign convention
     (fft.cpp:49-60): X[k] = sum_j x_j e^{-2 pi i j k / n}, with
     e^{-2 pi i j (n/2-1)/n} = (-1)^j e^{2 pi i j / n}. */
  const twiddles* tw = twiddle_table(fft_size);
  long double s_half = 0, s_re = 0, s_im = 0;
  for (int j = 0; j < y_length; ++j) {
    long double v = (j & 1) ? -(long double)y[j] : (long double)y[j];
    s_half += v;
    s_re += v * tw->c[j];
    s_im += v * tw->s[j];
  }
  long double h_half = 0, h_re = 0, h_im = 0;
  for (int m = 0; m < len; ++m) {
    long double v = (m & 1) ? -(long double)c[m] : (long double)c[m];
    h_half += v;
    h_re += v * tw->c[m];
    h_im += v * tw->s[m];
  }
  double S_half = (double)s_half;
  double P1re = (double)(s_re * h_re - s_im * h_im);
  double P1im = (double)(s_re * h_im + s_im * h_re);
  double P2 = (double)(s_half * h_half);
  double d1re = S_half * P1re - P1re, d1im = S_half * P1im - P1im;
  double d2 = S_half * P1re - P2;

  for (int i = 0; i < y_length; ++i) {
    int k = i + half + 1;
    double acc = 0.0;
    for (int m = 0; m < len; ++m) {
      int s = k - m;
      if (s >= 0 && s < y_length) acc += c[m] * y[s];
    }
    int kk = k % fft_size;
    double q = 2.0 * (d1re * tw->c[kk] + d1im * tw->s[kk]) + d2;
    if (k & 1) q = -q;
    out[i] = acc + q / fft_size;
  }
  free(c);
}

/* ---------------------------------------------------------------------------
 * Zero crossings: ZeroCrossingEngine harvest.cpp:162-197 and
 * GetFourZeroCrossingIntervals harvest.cpp:206-238.
 */
static int zero_crossing_engine(const double* v, int len, double fs,
                                double* locations, double* intervals) {
  int* edges = (int*)malloc(sizeof(int) * (len > 0 ? len : 1));
  int count = 0;
  for (int i = 0; i < len - 1; ++i)
    if (0.0 < v[i] && v[i + 1] <= 0.0) edges[count++] = i + 1;
  if (count < 2) {
    free(edges);
    return 0;
  }
  double* fine = (double*)malloc(sizeof(double) * count);
  for (int i = 0; i < count; ++i)
    fine[i] = edges[i] - v[edges[i] - 1] / (v[edges[i]] - v[edges[i] - 1]);
  for (int i = 0; i < count - 1; ++i) {
    intervals[i] = fs / (fine[i + 1] - fine[i]);
    locations[i] = (fine[i] + fine[i + 1]) / 2.0 / fs;
  }
  free(fine);
  free(edges);
  return count - 1;
}

/* histc matlabfunctions.cpp:136-155 */
static void histc(const double* x, int x_length, const double* edges,
                  int edges_length, int* index) {
  int count = 1;
  int i = 0;
  for (; i < edges_length; ++i) {
    index[i] = 1;
    if (edges[i] >= x[0]) break;
  }
  for (; i < edges_length; ++i) {
    if (edges[i] < x[count]) {
      index[i] = count;
    } else {
      index[i--] = count++;
    }
    if (count == x_length) break;
  }
  count--;
  for (i++; i < edges_length; ++i) index[i] = count;
}

/* interp1 matlabfunctions.cpp:157-176 (linear, extrapolating at both ends) */
static void interp1(const double* x, const double* y, int x_length,
                    const double* xi, int xi_length, double* yi) {
  double* h = (double*)malloc(sizeof(double) * (x_length - 1));
  int* k = (int*)calloc((size_t)xi_length, sizeof(int));
  for (int i = 0; i < x_length - 1; ++i) h[i] = x[i + 1] - x[i];
  histc(x, x_length, xi, xi_length, k);
  for (int i = 0; i < xi_length; ++i) {
    double s = (xi[i] - x[k[i] - 1]) / h[k[i] - 1];
    yi[i] = y[k[i] - 1] + s * (y[k[i]] - y[k[i] - 1]);
  }
  free(k);
  free(h);
}

/* GetF0CandidateFromRawEvent harvest.cpp:312-329 with GetFourZeroCrossing-
 * Intervals :206-238, GetF0CandidateContour :260-293 and
 * GetF0CandidateContourSub :240-254. */
void ho_band_candidates(const double* y, int y_length, int fft_size,
                        double fs_work, double boundary, double f0_floor,
                        double f0_ceil, int frames, double* out) {
  int n = y_length;
  double* f = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  ho_filtered_band(y, y_length, fft_size, fs_work, boundary, f);
  double* loc[4];
  double* val[4];
  int num[4];
  for (int e = 0; e < 4; ++e) {
    loc[e] = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    val[e] = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  }
  num[0] = zero_crossing_engine(f, n, fs_work, loc[0], val[0]);
  for (int i = 0; i < n; ++i) f[i] = -f[i];
  num[1] = zero_crossing_engine(f, n, fs_work, loc[1], val[1]);
  for (int i = 0; i < n - 1; ++i) f[i] = f[i] - f[i + 1];
  num[2] = zero_crossing_engine(f, n - 1, fs_work, loc[2], val[2]);
  for (int i = 0; i < n - 1; ++i) f[i] = -f[i];
  num[3] = zero_crossing_engine(f, n - 1, fs_work, loc[3], val[3]);

  if (!(num[0] > 2 && num[1] > 2 && num[2] > 2 && num[3] > 2)) {
    for (int i = 0; i < frames; ++i) out[i] = 0.0;
  } else {
    double* t = (double*)malloc(sizeof(double) * frames);
    double* v[4];
    for (int i = 0; i < frames; ++i) t[i] = i * 1 / 1000.0;
    for (int e = 0; e < 4; ++e) {
      v[e] = (double*)malloc(sizeof(double) * frames);
      interp1(loc[e], val[e], num[e], t, frames, v[e]);
    }
    double upper = boundary * 1.1, lower = boundary * 0.9;
    for (int i = 0; i < frames; ++i) {
      double c = (v[0][i] + v[1][i] + v[2][i] + v[3][i]) / 4.0;
      if (c > upper || c < lower || c > f0_ceil || c < f0_floor) c = 0.0;
      out[i] = c;
    }
    for (int e = 0; e < 4; ++e) free(v[e]);
    free(t);
  }
  for (int e = 0; e < 4; ++e) {
    free(loc[e]);
    free(val[e]);
  }
  free(f);
}

/* GetRawF0Candidates harvest.cpp:334-343 */
void ho_raw_candidates(const double* y, const ho_params* p,
                       const double* boundaries, double f0_floor,
                       double f0_ceil, double* raw) {
  for (int b = 0; b < p->bands; ++b)
    ho_band_candidates(y, p->y_length, p->fft_size, p->fs_work, boundaries[b],
                       f0_floor, f0_ceil, p->frames,
                       raw + (size_t)b * p->frames);
}

/* DetectOfficialF0Candidates harvest.cpp:348-412.  cand rows have p->slots
 * entries; columns [0, capacity) receive candidates, the rest of every row is
 * zero-filled as Sub2 does up to max_candidates (harvest.cpp:380). */
int ho_official(const double* raw, const ho_params* p, double* cand) {
  int B = p->bands, F = p->frames, S = p->slots;
  int* vuv = (int*)malloc(sizeof(int) * B);
  int* st = (int*)malloc(sizeof(int) * B);
  int* ed = (int*)malloc(sizeof(int) * B);
  int count = 0;
  for (int i = 0; i < F; ++i) {
    for (int j = 0; j < B; ++j) vuv[j] = raw[(size_t)j * F + i] > 0 ? 1 : 0;
    vuv[0] = vuv[B - 1] = 0;
    int sections = 0;
    for (int j = 1; j < B; ++j) {
      int d = vuv[j] - vuv[j - 1];
      if (d == 1) st[sections] = j;
      if (d == -1) ed[sections++] = j;
    }
    double* row = cand + (size_t)i * S;
    int n = 0;
    for (int s = 0; s < sections; ++s) {
      if (ed[s] - st[s] < 10) continue;
      double sum = 0.0;
      for (int j = st[s]; j < ed[s]; ++j) sum += raw[(size_t)j * F + i];
      sum /= (ed[s] - st[s]);
      row[n++] = sum;
    }
    for (int j = n; j < S; ++j) row[j] = 0.0;
    count = imax(count, n);
  }
  free(vuv);
  free(st);
  free(ed);
  return count;
}

/* OverlapF0Candidates harvest.cpp:417-429 */
void ho_overlap(const ho_params* p, int count, double* cand) {
  int F = p->frames, S = p->slots;
  for (int g = 1; g <= 3; ++g)
    for (int j = 0; j < count; ++j) {
      for (int k = g; k < F; ++k)
        cand[(size_t)k * S + j + count * g] = cand[(size_t)(k - g) * S + j];
      for (int k = 0; k < F - g; ++k)
        cand[(size_t)k * S + j + count * (g + 3)] = cand[(size_t)(k + g) * S + j];
    }
}

/* ---------------------------------------------------------------------------
 * Refinement: GetRefinedF0 harvest.cpp:589-617 -> GetMeanF0 :541-584 ->
 * GetBaseIndex :434-441, GetMainWindow :446-456, GetDiffWindow :462-468,
 * GetSpectra :474-505 (evaluated at the bins FixF0 reads) and FixF0 :507-536.
 */
static void dft_bin(const double* wave, int len, int fft_size, int k,
                    double* re, double* im) {
  const twiddles* tw = twiddle_table(fft_size);
  long double sr = 0, si = 0;
  for (int m = 0; m < len; ++m) {
    long long q = ((long long)k * m) % fft_size;
    sr += (long double)wave[m] * tw->c[q];
    si -= (long double)wave[m] * tw->s[q];
  }
  *re = (double)sr;
  *im = (double)si;
}

void ho_refine_one(const double* y, int y_length, double fs, double t,
                   double f0, double f0_floor, double f0_ceil,
                   double* refined, double* score) {
  if (f0 <= 0.0) {
    *refined = 0.0;
    *score = 0.0;
    return;
  }
  int half = (int)(1.5 * fs / f0 + 1.0);
  double window_time = (2.0 * half + 1.0) / fs;
  int len = half * 2 + 1;
  int fft_size = (int)pow(2.0, 2.0 + (int)(log(half * 2.0 + 1.0) / HO_LOG2));
  double base_time0 = (-half + 0) / fs;
  int basic = mround((t + base_time0) * fs + 0.001);

  double* w = (double*)malloc(sizeof(double) * len);
  double* d = (double*)malloc(sizeof(double) * len);
  double* xm = (double*)malloc(sizeof(double) * len);
  double* xd = (double*)malloc(sizeof(double) * len);
  for (int i = 0; i < len; ++i) {
    double tmp = (basic + i - 1.0) / fs - t;
    w[i] = 0.42 + 0.5 * cos(2.0 * HO_PI * tmp / window_time) +
           0.08 * cos(4.0 * HO_PI * tmp / window_time);
  }
  d[0] = -w[1] / 2.0;
  for (int i = 1; i < len - 1; ++i) d[i] = -(w[i + 1] - w[i - 1]) / 2.0;
  d[len - 1] = w[len - 2] / 2.0;
  for (int i = 0; i < len; ++i) {
    int safe = imax(0, imin(y_length - 1, basic + i - 1));
    xm[i] = y[safe] * w[i];
    xd[i] = y[safe] * d[i];
  }

  int harmonics = imin((int)(fs / 2.0 / f0), 6);
  double amp[6], ifreq[6];
  for (int i = 0; i < harmonics; ++i) {
    int index = mround(f0 * fft_size / fs * (i + 1));
    double mre, mim, dre, dim;
    dft_bin(xm, len, fft_size, index, &mre, &mim);
    dft_bin(xd, len, fft_size, index, &dre, &dim);
    double num = mre * dim - mim * dre;
    double pw = mre * mre + mim * mim;
    ifreq[i] = pw == 0.0 ? 0.0
                         : (double)index * fs / fft_size +
                               num / pw * fs / 2.0 / HO_PI;
    amp[i] = sqrt(pw);
  }
  double den = 0.0, numer = 0.0, s = 0.0;
  for (int i = 0; i < harmonics; ++i) {
    numer += amp[i] * ifreq[i];
    den += amp[i] * (i + 1.0);
    s += fabs((ifreq[i] / (i + 1.0) - f0) / f0);
  }
  double r = numer / (den + HO_GUARD);
  double sc = 1.0 / (s / harmonics + HO_GUARD);
  if (r < f0_floor || r > f0_ceil || sc < 2.5) {
    r = 0.0;
    sc = 0.0;
  }
  *refined = r;
  *score = sc;
  free(w);
  free(d);
  free(xm);
  free(xd);
}

/* RefineF0Candidates harvest.cpp:622-631 */
void ho_refine(const double* y, int y_length, double fs_work, int frames,
               int stride, int ncand, double f0_floor, double f0_ceil,
               double* f0, double* score) {
  for (int i = 0; i < frames; ++i)
    for (int j = 0; j < ncand; ++j) {
      size_t at = (size_t)i * stride + j;
      ho_refine_one(y, y_length, fs_work, i * 1 / 1000.0, f0[at], f0_floor,
                    f0_ceil, &f0[at], &score[at]);
    }
}

/* SelectBestF0 harvest.cpp:636-650 */
static double select_best(double reference, const double* cands, int n,
                          double allowed, double* best_error) {
  double best = 0.0;
  *best_error = allowed;
  for (int i = 0; i < n; ++i) {
    double tmp = fabs(reference - cands[i]) / reference;
    if (tmp > *best_error) continue;
    best = cands[i];
    *best_error = tmp;
  }
  return best;
}

/* RemoveUnreliableCandidates harvest.cpp:652-688 */
void ho_prune(int frames, int stride, int ncand, double* f0, double* score) {
  double* tmp = (double*)malloc(sizeof(double) * (size_t)frames * stride + 1);
  memcpy(tmp, f0, sizeof(double) * (size_t)frames * stride);
  for (int i = 1; i < frames - 1; ++i)
    for (int j = 0; j < ncand; ++j) {
      double ref = f0[(size_t)i * stride + j];
      if (ref == 0) continue;
      double e1, e2;
      select_best(ref, tmp + (size_t)(i + 1) * stride, ncand, 1.0, &e1);
      select_best(ref, tmp + (size_t)(i - 1) * stride, ncand, 1.0, &e2);
      double e = e1 < e2 ? e1 : e2;
      if (e <= 0.05) continue;
      f0[(size_t)i * stride + j] = 0;
      score[(size_t)i * stride + j] = 0;
    }
  free(tmp);
}

/* SearchF0Base harvest.cpp:693-705 */
void ho_base(const double* f0, const double* score, int frames, int stride,
             int ncand, double* base) {
  for (int i = 0; i < frames; ++i) {
    double best = 0.0;
    base[i] = 0.0;
    for (int j = 0; j < ncand; ++j)
      if (score[(size_t)i * stride + j] > best) {
        base[i] = f0[(size_t)i * stride + j];
        best = score[(size_t)i * stride + j];
      }
  }
}

/* ---------------------------------------------------------------------------
 * Contour fixing: FixF0Contour harvest.cpp:1027-1044 with FixStep1 :710-722,
 * GetBoundaryList :727-743, FixStep2 :748-762, GetMultiChannelF0 :767-778,
 * ExtendF0 :791-820, Swap :826-838, ExtendSub :840-856, Extend :861-878,
 * MakeSortedOrder :883-896, SearchScore :901-907, MergeF0Sub :912-932,
 * MergeF0 :937-963, FixStep3 :968-995, FixStep4 :1000-1022.  Candidate rows
 * are addressed as (row pointer, ncand) exactly as WORLD's double** are.
 */
typedef struct {
  const double* f0;
  const double* score;
  int stride;
  int ncand;
} cand_view;

static void fix_step1(const double* base, int F, double allowed, double* out) {
  for (int i = 0; i < F; ++i) out[i] = 0.0;
  for (int i = 2; i < F; ++i) {
    if (base[i] == 0.0) continue;
    double ref = base[i - 1] * 2 - base[i - 2];
    out[i] = fabs((base[i] - ref) / ref) > allowed &&
                     fabs((base[i] - base[i - 1])) / base[i - 1] > allowed
                 ? 0.0
                 : base[i];
  }
}

static int boundary_list(const double* f0, int F, int* list) {
  int n = 0;
  int* vuv = (int*)malloc(sizeof(int) * (F > 0 ? F : 1));
  for (int i = 0; i < F; ++i) vuv[i] = f0[i] > 0 ? 1 : 0;
  vuv[0] = vuv[F - 1] = 0;
  for (int i = 1; i < F; ++i)
    if (vuv[i] - vuv[i - 1] != 0) {
      list[n] = i - n % 2;
      n++;
    }
  free(vuv);
  return n;
}

static void fix_step2(const double* in, int F, int minimum, double* out) {
  for (int i = 0; i < F; ++i) out[i] = in[i];
  int* list = (int*)malloc(sizeof(int) * F);
  int n = boundary_list(in, F, list);
  for (int i = 0; i < n / 2; ++i) {
    if (list[i * 2 + 1] - list[i * 2] >= minimum) continue;
    for (int j = list[i * 2]; j <= list[i * 2 + 1]; ++j) out[j] = 0.0;
  }
  free(list);
}

static void multi_channel(const double* f0, int F, const int* list, int n,
                          double** mc) {
  for (int i = 0; i < n / 2; ++i) {
    for (int j = 0; j < list[i * 2]; ++j) mc[i][j] = 0.0;
    for (int j = list[i * 2]; j <= list[i * 2 + 1]; ++j) mc[i][j] = f0[j];
    for (int j = list[i * 2 + 1] + 1; j < F; ++j) mc[i][j] = 0.0;
  }
}

static int extend_f0(int origin, int last_point, int shift,
                     const cand_view* cv, double allowed, double* ext) {
  const int threshold = 4;
  double tmp_f0 = ext[origin];
  int shifted = origin;
  int distance = abs(last_point - origin);
  int count = 0;
  double dummy;
  for (int i = 0; i <= distance; ++i) {
    int idx = origin + shift * i + shift;
    ext[idx] = select_best(tmp_f0, cv->f0 + (size_t)idx * cv->stride,
                           cv->ncand, allowed, &dummy);
    if (ext[idx] == 0.0) {
      count++;
    } else {
      tmp_f0 = ext[idx];
      count = 0;
      shifted = idx;
    }
    if (count == threshold) break;
  }
  return shifted;
}

static void swap_channels(int a, int b, double** f0, int* list) {
  double* p = f0[a];
  f0[a] = f0[b];
  f0[b] = p;
  int t = list[a * 2];
  list[a * 2] = list[b * 2];
  list[b * 2] = t;
  t = list[a * 2 + 1];
  list[a * 2 + 1] = list[b * 2 + 1];
  list[b * 2 + 1] = t;
}

static int extend_sub(double** ext, int* list, int sections) {
  const double threshold = 2200.0;
  int count = 0;
  double mean = 0.0;
  for (int i = 0; i < sections; ++i) {
    int st = list[i * 2], ed = list[i * 2 + 1];
    for (int j = st; j < ed; ++j) mean += ext[i][j];
    mean /= ed - st;
    if (threshold / mean < ed - st) swap_channels(count++, i, ext, list);
  }
  return count;
}

static int extend_all(double** mc, int sections, int F, int* list,
                      const cand_view* cv, double allowed) {
  const int threshold = 100;
  for (int i = 0; i < sections; ++i) {
    list[i * 2 + 1] = extend_f0(list[i * 2 + 1],
                                imin(F - 2, list[i * 2 + 1] + threshold), 1,
                                cv, allowed, mc[i]);
    list[i * 2] = extend_f0(list[i * 2], imax(1, list[i * 2] - threshold), -1,
                            cv, allowed, mc[i]);
  }
  return extend_sub(mc, list, sections);
}

static void sorted_order(const int* list, int n, int* order) {
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int i = 1; i < n; ++i)
    for (int j = i - 1; j >= 0; --j)
      if (list[order[j] * 2] > list[order[i] * 2]) {
        int t = order[i];
        order[i] = order[j];
        order[j] = t;
      } else {
        break;
      }
}

static double search_score(double f0, const double* c, const double* s, int n) {
  double score = 0.0;
  for (int i = 0; i < n; ++i)
    if (f0 == c[i] && score < s[i]) score = s[i];
  return score;
}

static int merge_sub(const double* f1, int st1, int ed1, const double* f2,
                     int st2, int ed2, const cand_view* cv, double* merged) {
  if (st1 <= st2 && ed1 >= ed2) return ed1;
  double s1 = 0.0, s2 = 0.0;
  for (int i = st2; i <= ed1; ++i) {
    const double* c = cv->f0 + (size_t)i * cv->stride;
    const double* s = cv->score + (size_t)i * cv->stride;
    s1 += search_score(f1[i], c, s, cv->ncand);
    s2 += search_score(f2[i], c, s, cv->ncand);
  }
  if (s1 > s2)
    for (int i = ed1; i <= ed2; ++i) merged[i] = f2[i];
  else
    for (int i = st2; i <= ed2; ++i) merged[i] = f2[i];
  return ed2;
}

static void merge_f0(double** mc, int* list, int channels, int F,
                     const cand_view* cv, double* merged) {
  int* order = (int*)malloc(sizeof(int) * channels);
  sorted_order(list, channels, order);
  for (int i = 0; i < F; ++i) merged[i] = mc[0][i];
  for (int i = 1; i < channels; ++i)
    if (list[order[i] * 2] - list[1] > 0) {
      for (int j = list[order[i] * 2]; j <= list[order[i] * 2 + 1]; ++j)
        merged[j] = mc[order[i]][j];
      list[0] = list[order[i] * 2];
      list[1] = list[order[i] * 2 + 1];
    } else {
      list[1] = merge_sub(merged, list[0], list[1], mc[order[i]],
                          list[order[i] * 2], list[order[i] * 2 + 1], cv,
                          merged);
    }
  free(order);
}

static void fix_step3(const double* in, int F, const cand_view* cv,
                      double allowed, double* out) {
  for (int i = 0; i < F; ++i) out[i] = in[i];
  int* list = (int*)malloc(sizeof(int) * F);
  int n = boundary_list(in, F, list);
  int sections = n / 2;
  double** mc = (double**)malloc(sizeof(double*) * (sections > 0 ? sections : 1));
  for (int i = 0; i < sections; ++i) mc[i] = (double*)malloc(sizeof(double) * F);
  multi_channel(in, F, list, n, mc);
  int channels = extend_all(mc, sections, F, list, cv, allowed);
  if (channels != 0) merge_f0(mc, list, channels, F, cv, out);
  for (int i = 0; i < sections; ++i) free(mc[i]);
  free(mc);
  free(list);
}

static void fix_step4(const double* in, int F, int threshold, double* out) {
  for (int i = 0; i < F; ++i) out[i] = in[i];
  int* list = (int*)malloc(sizeof(int) * F);
  int n = boundary_list(in, F, list);
  for (int i = 0; i < n / 2 - 1; ++i) {
    int distance = list[(i + 1) * 2] - list[i * 2 + 1] - 1;
    if (distance >= threshold) continue;
    double t0 = in[list[i * 2 + 1]] + 1;
    double t1 = in[list[(i + 1) * 2]] - 1;
    double coef = (t1 - t0) / (distance + 1.0);
    int count = 1;
    for (int j = list[i * 2 + 1] + 1; j <= list[(i + 1) * 2] - 1; ++j)
      out[j] = t0 + coef * count++;
  }
  free(list);
}

void ho_fix_contour(const double* f0, const double* score, int frames,
                    int stride, int ncand, double* out) {
  cand_view cv = {f0, score, stride, ncand};
  double* a = (double*)malloc(sizeof(double) * frames);
  double* b = (double*)malloc(sizeof(double) * frames);
  ho_base(f0, score, frames, stride, ncand, a);
  fix_step1(a, frames, 0.008, b);
  fix_step2(b, frames, 6, a);
  fix_step3(a, frames, &cv, 0.18, b);
  fix_step4(b, frames, 9, out);
  free(a);
  free(b);
}

/* FilteringF0 harvest.cpp:1049-1074 and SmoothF0Contour :1079-1113 */
static void filtering_f0(const double* a, const double* b, double* x, int n,
                         int st, int ed, double* y) {
  double w0 = 0.0, w1 = 0.0, wt;
  double* tmp = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < st; ++i) x[i] = x[st];
  for (int i = ed + 1; i < n; ++i) x[i] = x[ed];
  for (int i = 0; i < n; ++i) {
    wt = x[i] + a[0] * w0 + a[1] * w1;
    tmp[n - i - 1] = b[0] * wt + b[1] * w0 + b[0] * w1;
    w1 = w0;
    w0 = wt;
  }
  w0 = w1 = 0.0;
  for (int i = 0; i < n; ++i) {
    wt = tmp[i] + a[0] * w0 + a[1] * w1;
    y[n - i - 1] = b[0] * wt + b[1] * w0 + b[0] * w1;
    w1 = w0;
    w0 = wt;
  }
  free(tmp);
}

void ho_smooth(const double* f0, int frames, double* out) {
  const double b[2] = {0.0078202080334971724, 0.015640416066994345};
  const double a[2] = {1.7347257688092754, -0.76600660094326412};
  const int lag = 300;
  int n = frames + lag * 2;
  double* c = (double*)calloc((size_t)n, sizeof(double));
  for (int i = lag; i < lag + frames; ++i) c[i] = f0[i - lag];
  int* list = (int*)malloc(sizeof(int) * n);
  int nb = boundary_list(c, n, list);
  int sections = nb / 2;
  double** mc = (double**)malloc(sizeof(double*) * (sections > 0 ? sections : 1));
  for (int i = 0; i < sections; ++i) mc[i] = (double*)malloc(sizeof(double) * n);
  multi_channel(c, n, list, nb, mc);
  for (int i = 0; i < frames; ++i) out[i] = 0.0;
  for (int i = 0; i < sections; ++i) {
    filtering_f0(a, b, mc[i], n, list[i * 2], list[i * 2 + 1], c);
    for (int j = list[i * 2]; j <= list[i * 2 + 1]; ++j) out[j - lag] = c[j];
  }
  for (int i = 0; i < sections; ++i) free(mc[i]);
  free(mc);
  free(c);
  free(list);
}

/* Harvest harvest.cpp:1223-1255 with HarvestGeneralBody :1145-1215 */
int ho_harvest(const double* x, int length, int fs, double f0_floor,
               double f0_ceil, double frame_period, double* f0, double* t) {
  ho_params p;
  ho_params_get(length, fs, f0_floor, f0_ceil, frame_period, &p);
  double* bounds = (double*)malloc(sizeof(double) * p.bands);
  ho_boundaries(&p, f0_floor, bounds);
  double* y = (double*)malloc(sizeof(double) * p.y_length);
  ho_decimate(x, length, p.ratio, p.y_length, y);
  ho_remove_mean(y, p.y_length);
  double* raw = (double*)malloc(sizeof(double) * (size_t)p.bands * p.frames);
  ho_raw_candidates(y, &p, bounds, f0_floor, f0_ceil, raw);
  double* cand = (double*)malloc(sizeof(double) * (size_t)p.frames * p.slots);
  double* score = (double*)calloc((size_t)p.frames * p.slots, sizeof(double));
  int count = ho_official(raw, &p, cand);
  ho_overlap(&p, count, cand);
  int ncand = count * 7;
  ho_refine(y, p.y_length, p.fs_work, p.frames, p.slots, ncand, f0_floor,
            f0_ceil, cand, score);
  ho_prune(p.frames, p.slots, ncand, cand, score);
  double* best = (double*)malloc(sizeof(double) * p.frames);
  double* basic = (double*)malloc(sizeof(double) * p.frames);
  ho_fix_contour(cand, score, p.frames, p.slots, ncand, best);
  ho_smooth(best, p.frames, basic);
  for (int i = 0; i < p.out_frames; ++i) {
    t[i] = i * frame_period / 1000.0;
    f0[i] = basic[imin(p.frames - 1, mround(t[i] * 1000.0))];
  }
  free(bounds);
  free(y);
  free(raw);
  free(cand);
  free(score);
  free(best);
  free(basic);
  return p.out_frames;
}
