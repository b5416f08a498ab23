/*
 * Harvest oracle: a plain-C restatement of WORLD's Harvest F0 estimator.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path never calls it.
 *
 * Every function follows WORLD as vendored by the reference
 * (third_party/world/src/harvest.cpp, matlabfunctions.cpp, common.cpp,
 * pyworld 0.3.5) and cites the lines it restates.  Two stages are restated in
 * a mathematically identical form instead of WORLD's literal FFT code:
 *   - the band-pass filterbank (harvest.cpp:99-148) is a direct linear
 *     convolution plus the exact two-bin correction that reproduces WORLD's
 *     overwrite of bins n/2-1 and n/2 (harvest.cpp:122-133); WORLD's
 *     unnormalised c2r (fft.cpp:26-35) scales every sample by n, a power of
 *     two, which leaves every zero crossing and fine edge bit-identical, so the
 *     oracle works unscaled;
 *   - the refinement spectra (harvest.cpp:474-505) are DFTs evaluated at the
 *     <= 6 harmonic bins FixF0 reads (harvest.cpp:514-519) instead of full
 *     FFTs.
 *
 * Pinning (each function is checked by tests/test_oracle.py):
 *   ho_params, ho_resample_times    closed forms; WORLD tap (oracle/_ref)
 *   ho_decimate                      scipy lfilter on the documented padding;
 *                                    WORLD tap; impulse response
 *   ho_filtered_band                 numpy FFT convolution with the two-bin
 *                                    overwrite applied by hand; WORLD tap
 *   ho_band_candidates               brute force events/interp1 in Python;
 *                                    WORLD tap (raw candidate matrix)
 *   ho_official, ho_overlap          brute force; WORLD tap
 *   ho_refine                        pure sinusoid closed form; numpy full-FFT
 *                                    re-expression; WORLD tap
 *   ho_prune, ho_base                brute force; WORLD tap
 *   ho_fix_contour, ho_smooth        WORLD tap, bit-exact, random matrices
 *   ho_harvest                       WORLD Harvest end to end (oracle/_ref)
 */
#ifndef HARVEST_ORACLE_H_
#define HARVEST_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int bands;          /* B: 1 + int(log2(1.1 ceil / 0.9 floor) * 40)          */
  int ratio;          /* r: clamp(round(fs / 8000), 1, 12)                     */
  double fs_work;     /* fs / r                                                */
  int y_length;       /* N = ceil(L / r)                                       */
  int fft_size;       /* n = GetSuitableFFTSize(N + 5 + 2 int(2 fs_w / b0))    */
  int frames;         /* F = int(1000 L / fs) + 1 (1 ms frames)                */
  int out_frames;     /* int(1000 L / fs / period) + 1                         */
  int capacity;       /* round(B / 10): official candidates per frame          */
  int slots;          /* 7 * capacity                                          */
} ho_params;

int ho_params_get(int64_t length, int fs, double f0_floor, double f0_ceil,
                  double frame_period, ho_params* out);
void ho_boundaries(const ho_params* p, double f0_floor, double* boundaries);

/* y[N] = decimated waveform before mean removal (harvest.cpp:43-66). */
void ho_decimate(const double* x, int length, int ratio, int y_length,
                 double* y);
/* y -= mean(y) with WORLD's sequential sum (harvest.cpp:81-85). */
void ho_remove_mean(double* y, int y_length);

/* Filtered (unscaled) signal of one band over N samples (harvest.cpp:99-148). */
void ho_filtered_band(const double* y, int y_length, int fft_size,
                      double fs_work, double boundary, double* out);

/* Raw candidate contour of one band over F frames (harvest.cpp:312-329). */
void ho_band_candidates(const double* y, int y_length, int fft_size,
                        double fs_work, double boundary, double f0_floor,
                        double f0_ceil, int frames, double* out);

/* raw[B][F] for all bands (harvest.cpp:334-343). */
void ho_raw_candidates(const double* y, const ho_params* p,
                       const double* boundaries, double f0_floor,
                       double f0_ceil, double* raw);

/* cand[F][slots] (first `capacity` columns used) and count
   (harvest.cpp:388-412).  cand must have F * p->slots doubles. */
int ho_official(const double* raw, const ho_params* p, double* cand);
/* Spread candidates to +-3 frames inside cand[F][slots] (harvest.cpp:417-429). */
void ho_overlap(const ho_params* p, int count, double* cand);

/* Refine every slot j < ncand of every frame (harvest.cpp:589-631). */
void ho_refine(const double* y, int y_length, double fs_work, int frames,
               int stride, int ncand, double f0_floor, double f0_ceil,
               double* f0, double* score);
/* One slot: refined f0 and score for candidate `f0` at time t. */
void ho_refine_one(const double* y, int y_length, double fs_work, double t,
                   double f0, double f0_floor, double f0_ceil,
                   double* refined, double* score);

void ho_prune(int frames, int stride, int ncand, double* f0, double* score);
void ho_base(const double* f0, const double* score, int frames, int stride,
             int ncand, double* base);
void ho_fix_contour(const double* f0, const double* score, int frames,
                    int stride, int ncand, double* out);
void ho_smooth(const double* f0, int frames, double* out);

/* End-to-end Harvest: f0[out_frames], t[out_frames] (harvest.cpp:1223-1255). */
int ho_harvest(const double* x, int length, int fs, double f0_floor,
               double f0_ceil, double frame_period, double* f0, double* t);

#ifdef __cplusplus
}
#endif
#endif
