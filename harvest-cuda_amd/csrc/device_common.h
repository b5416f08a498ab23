// Shared device-side definitions of the gfx950 Harvest backend.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hv {

// WORLD constants (world/constantnumbers.h:18-24).
constexpr double kPi = 3.1415926535897932384;
constexpr double kLog2 = 0.69314718055994529;
constexpr double kGuard = 0.000000000001;

constexpr int kThreads = 256;        // workgroup size of every kernel (4 waves)
constexpr int kTile = 2048;          // filterbank output samples per workgroup
// Tiles start kTileStride samples apart and own the crossings of their first
// kTileStride samples; the two extra outputs are the next tile's first two,
// which the derivative crossings of the last owned samples read.
constexpr int kTileStride = kTile - 2;
constexpr int kSlabCap = kTile / 2;  // max crossings of one type per tile
constexpr int kChunk = 64;          // IIR samples per scan chunk
constexpr int kMeanBlock = 4096;     // samples per deterministic partial sum
constexpr int kWindowMargin = 4;     // crossings kept either side of a tile
// Frames per thread in k_interp: a tile spans kTileStride r 1000 / fs frames, at
// most 385 for the smallest working rate (fs / r >= 5333 Hz); launches check it.
constexpr int kInterpFramesPerThread = 2;
// Work counters are spread over this many addresses (by workgroup) so that
// the per-wave atomics of a launch do not queue on one address; the host sums.
// Counter k (the host's HV_STAT_SLOTS counters) occupies words
// [k kWorkSpread, (k + 1) kWorkSpread); after them, the refinement's flag for
// a window or bin beyond its precomputed bounds.
constexpr int kWorkSpread = 64;
constexpr int kStatWindowSamples = 0, kStatBinSamples = 1, kStatBins = 2, kStatTablePairs = 3;
constexpr int kStatFault = 4;

// Per-utterance layout of a batch.  All offsets index the concatenated
// device arrays of the batch.
struct Utt {
  long long x_off;   // input samples
  long long length;  // L
  long long m_off;   // decimation pass buffer (M samples)
  long long m;       // L + 2 lag + 18 when ratio > 1
  long long y_off;   // decimated samples
  long long f_off;   // 1 ms frames
  long long t_off;   // filterbank tiles
  long long c_off;   // IIR scan chunks
  long long p_off;   // partial-sum blocks
  int n;             // y_length N
  int fft;           // WORLD's global FFT size n
  int frames;        // F
  int tiles;
  int chunks;
  int pblocks;
  int nbeg;          // decimate(): r - r*nout + padded length
  int pad;
};

// Launch-wide constants.
struct Config {
  int count;         // utterances
  int ratio;
  int lag;           // ceil(140 / r) * r
  int fs;
  double fs_work;
  double f0_floor, f0_ceil;
  int bands;
  int capacity;      // official candidates per frame
  int slots;         // 7 * capacity
  int twiddle_n;     // size of the refinement twiddle table (0: none)
  int lds_limit;     // dynamic LDS bytes a workgroup may use on this device
  int hmax;          // largest refinement half window (candidates >= f0_floor)
  int win_pairs;     // 1: the refinement windows depend on h alone (see k_window_pairs)
  // Bin-table refinement (refine_table.hip), on when bt_hr > 0: half windows
  // [bt_hlo, bt_hlo + bt_hr), at most bt_nbmax bins per half window; the
  // DFT kernel's persistent grid has at most bt_groups workgroups.
  int bt_hlo, bt_hr, bt_nbmax, bt_groups;
  long long phase_n; // size P of the filterbank's two-level phase table (>= every fft)
  int phase_bits;    // its fine part holds 2^phase_bits entries
  double a0, a1, a2, b0, b1;  // decimation IIR
  double chunk_pow[9];        // A^kChunk of the IIR state transition
};

// FFT size of GetMeanF0 for half window h: 2^(2 + floor(log2(2h + 1)))
// (harvest.cpp:598-599; 2h + 1 is odd, so log / kLog2 there is never within
// rounding of an integer).  window_fft_log2 is its exponent.
__host__ __device__ inline int window_fft_log2(int h) {
  return 2 + (31 - __builtin_clz((unsigned)(2 * h + 1)));
}
__host__ __device__ inline int window_fft(int h) { return 1 << window_fft_log2(h); }

// GetMainWindow (harvest.cpp:446-456) evaluates 0.42 + 0.5 cos phi +
// 0.08 cos 2 phi at phi = 2 pi ((basic + m - 1) / fs - t) / (len / fs) with
// basic = round(t fs - h), so phi = 2 pi (m - h - 1) / len whenever t fs is an
// integer, which holds at every 1 ms frame when fs_work is a multiple of
// 1000 Hz.  Sample m of such a window, from sincospi of the exact rational
// argument.
__device__ __forceinline__ double main_window_at(int m, int h) {
  double c1, c2, sn;
  sincospi(2.0 * (double)(m - h - 1) / (2 * h + 1), &sn, &c1);
  sincospi(4.0 * (double)(m - h - 1) / (2 * h + 1), &sn, &c2);
  return 0.42 + 0.5 * c1 + 0.08 * c2;
}

// (w[m], d[m]) of the window of half length h > 0 above: d by GetDiffWindow's
// differences (harvest.cpp:459-468): d[0] = -w[1] / 2,
// d[m] = -(w[m + 1] - w[m - 1]) / 2, d[len - 1] = w[len - 2] / 2.
__device__ __forceinline__ double2 window_pair(int m, int h) {
  const int last = 2 * h;
  double d;
  if (m == 0) d = -main_window_at(1, h) / 2.0;
  else if (m == last) d = main_window_at(last - 1, h) / 2.0;
  else d = -(main_window_at(m + 1, h) - main_window_at(m - 1, h)) / 2.0;
  return make_double2(main_window_at(m, h), d);
}

__host__ __device__ inline int mround(double x) {
  return x > 0 ? (int)(x + 0.5) : (int)(x - 0.5);
}

// Largest u in [0, count) with off[u] <= g (off non-decreasing, off[0] = 0).
__device__ __forceinline__ int find_segment(const long long* off, int count,
                                            long long g) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace hv
