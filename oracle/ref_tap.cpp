// Stage taps into the reference's vendored WORLD Harvest (pyworld 0.3.5).
//
// TEST INFRASTRUCTURE ONLY.  This translation unit textually includes
// third_party/world/src/harvest.cpp from the reference checkout (the include
// path is given by oracle/Makefile) so that WORLD's internal static routines
// can be called stage by stage, as the reference's own src/engine.cpp:13 and
// tests/cpp/test_contour.cpp:3 do.  Nothing here is copied: the build reads
// the reference sources where they lie and writes only oracle/_ref/.
#include "harvest.cpp"

#include <vector>

namespace {
struct Rows {
  std::vector<double*> rows;
  Rows(double* base, int n, int stride) : rows(n) {
    for (int i = 0; i < n; ++i) rows[i] = base + static_cast<size_t>(i) * stride;
  }
  double** data() { return rows.data(); }
};

struct Setup {
  int bands, ratio, y_length, fft_size, frames, capacity;
  double fs_work;
  std::vector<double> boundaries;
  Setup(int length, int fs, double lo, double hi) {
    double af = lo * 0.9, ac = hi * 1.1;
    bands = 1 + static_cast<int>(log(ac / af) / world::kLog2 * 40);
    boundaries.resize(bands);
    for (int i = 0; i < bands; ++i) boundaries[i] = af * pow(2.0, (i + 1) / 40.0);
    ratio = MyMaxInt(MyMinInt(matlab_round(fs / 8000.0), 12), 1);
    y_length = static_cast<int>(ceil(static_cast<double>(length) / ratio));
    fs_work = static_cast<double>(fs) / ratio;
    fft_size = GetSuitableFFTSize(y_length + 5 +
                                  2 * static_cast<int>(2.0 * fs_work / boundaries[0]));
    frames = GetSamplesForHarvest(fs, length, 1);
    capacity = matlab_round(bands / 10.0) * 7;
  }
};
}  // namespace

extern "C" {

// Harvest harvest.cpp:1223-1255.
void ref_harvest(const double* x, int length, int fs, double floor, double ceil,
                 double period, double* f0, double* t) {
  HarvestOption option{floor, ceil, period};
  Harvest(x, length, fs, &option, t, f0);
}

// y[N] after GetWaveformAndSpectrum (decimation and mean removal,
// harvest.cpp:71-93).
void ref_waveform(const double* x, int length, int fs, double floor,
                  double ceil, double* y_out) {
  Setup s(length, fs, floor, ceil);
  std::vector<double> y(s.fft_size);
  std::vector<fft_complex> spectrum(s.fft_size);
  GetWaveformAndSpectrum(x, length, s.y_length, s.fs_work, s.fft_size, s.ratio,
                         y.data(), spectrum.data());
  for (int i = 0; i < s.y_length; ++i) y_out[i] = y[i];
}

// WORLD's filtered signal of band `band` (GetFilteredSignal harvest.cpp:99-148),
// first N samples, in WORLD's scale (n times the convolution).
void ref_filtered_band(const double* x, int length, int fs, double floor,
                       double ceil, int band, double* out) {
  Setup s(length, fs, floor, ceil);
  std::vector<double> y(s.fft_size), filtered(s.fft_size);
  std::vector<fft_complex> spectrum(s.fft_size);
  GetWaveformAndSpectrum(x, length, s.y_length, s.fs_work, s.fft_size, s.ratio,
                         y.data(), spectrum.data());
  GetFilteredSignal(s.boundaries[band], s.fft_size, s.fs_work, spectrum.data(),
                    s.y_length, filtered.data());
  for (int i = 0; i < s.y_length; ++i) out[i] = filtered[i];
}

// raw[B][F] (GetRawF0Candidates harvest.cpp:334-343).
void ref_raw_candidates(const double* x, int length, int fs, double floor,
                        double ceil, double* raw) {
  Setup s(length, fs, floor, ceil);
  std::vector<double> y(s.fft_size), times(s.frames);
  std::vector<fft_complex> spectrum(s.fft_size);
  GetWaveformAndSpectrum(x, length, s.y_length, s.fs_work, s.fft_size, s.ratio,
                         y.data(), spectrum.data());
  for (int i = 0; i < s.frames; ++i) times[i] = i * 1 / 1000.0;
  Rows rows(raw, s.bands, s.frames);
  GetRawF0Candidates(s.boundaries.data(), s.bands, s.fs_work, s.y_length,
                     times.data(), s.frames, spectrum.data(), s.fft_size,
                     floor, ceil, rows.data());
}

// Official + overlapped candidates cand[F][capacity] from raw[B][F];
// returns count (DetectOfficialF0Candidates :388-412, Overlap :417-429).
int ref_official_overlap(const double* raw, int bands, int frames,
                         int capacity, double* cand) {
  std::vector<const double*> rrows(bands);
  for (int i = 0; i < bands; ++i) rrows[i] = raw + static_cast<size_t>(i) * frames;
  Rows rows(cand, frames, capacity);
  int count = DetectOfficialF0Candidates(rrows.data(), bands, frames, capacity,
                                         rows.data());
  OverlapF0Candidates(frames, count, rows.data());
  return count;
}

// RefineF0Candidates harvest.cpp:622-631 over cand[F][stride], first ncand.
void ref_refine(const double* y, int y_length, double fs_work, int frames,
                int stride, int ncand, double floor, double ceil, double* f0,
                double* score) {
  std::vector<double> times(frames);
  for (int i = 0; i < frames; ++i) times[i] = i * 1 / 1000.0;
  Rows a(f0, frames, stride), b(score, frames, stride);
  RefineF0Candidates(y, y_length, fs_work, times.data(), frames, ncand, floor,
                     ceil, a.data(), b.data());
}

// RemoveUnreliableCandidates harvest.cpp:672-688.
void ref_prune(int frames, int stride, int ncand, double* f0, double* score) {
  Rows a(f0, frames, stride), b(score, frames, stride);
  RemoveUnreliableCandidates(frames, ncand, a.data(), b.data());
}

// FixF0Contour harvest.cpp:1027-1044.
void ref_fix_contour(double* f0, double* score, int frames, int stride,
                     int ncand, double* out) {
  Rows a(f0, frames, stride), b(score, frames, stride);
  FixF0Contour(a.data(), b.data(), frames, ncand, out);
}

// SmoothF0Contour harvest.cpp:1079-1113 (output must be zero-initialised, as
// HarvestGeneralBody does at :1174-1177).
void ref_smooth(const double* f0, int frames, double* out) {
  for (int i = 0; i < frames; ++i) out[i] = 0.0;
  SmoothF0Contour(f0, frames, out);
}

}  // extern "C"
