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
constexpr int kWorkSpread = 64;

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
  long long phase_n; // size P of the filterbank's two-level phase table (>= every fft)
  int phase_bits;    // its fine part holds 2^phase_bits entries
  double a0, a1, a2, b0, b1;  // decimation IIR
  double chunk_pow[9];        // A^kChunk of the IIR state transition
};

__device__ __forceinline__ int mround(double x) {
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
