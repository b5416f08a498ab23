// Bin-table refinement: WORLD's RefineF0Candidates (harvest.cpp:622-631;
// GetRefinedF0 :589-617, GetMeanF0 :541-584, FixF0 :507-536) for analyses
// whose 1 ms frames start at whole samples, i.e. fs_work a multiple of
// 1000 Hz (every WORLD working rate of 8 kHz).
//
// A slot's refinement depends on its frame, on the half window h of its
// candidate (the window and the FFT size n follow from h) and on the bins k
// FixF0 reads.  With whole-sample frame offsets the window of half length h
// is the same at every frame (window_pair), and the bins that candidates of
// one h can read form a short range per harmonic: the candidates giving h lie
// in one narrow f0 interval.  The work is therefore split as
//   k_bt_codes    one lane per official candidate: its half window h,
//                 harmonic count and FixF0 bins, packed in one word, 1/f0,
//                 and h again as a byte (for k_bt_dft's slot search);
//   k_bt_plan     one lane per frame: the half windows its slots use (a
//                 bitmask per frame) and, per block of 64 frames and per h,
//                 the frames that use h;
//   scan + k_bt_scatter  the (frame, h) pairs listed per h in frame order;
//   k_bt_dft      one lane per pair and one wavefront per 64 pairs of one h:
//                 the main and derivative windowed spectra at every bin of
//                 h's table by Goertzel recurrences over the window's
//                 samples, whose window values and recurrence coefficients
//                 are uniform across the wavefront (scalar operands), with
//                 the samples staged in LDS by direct loads; then FixF0
//                 and GetRefinedF0's gate of every slot that reads the
//                 pair, from the bins in LDS (one pass over the window, at
//                 most kBtPassBins bins) or read back from res (more bins);
//   k_bt_finish   one wavefront per frame with more than kBtPairs half
//                 windows (not listed): FixF0 of its slots, bins computed
//                 one at a time.
// Each bin is one fixed sequence of operations (samples in order), so the
// results depend neither on the batch nor on which kernel computed the bin.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.h"

namespace hv {
namespace {

#define LDS __attribute__((address_space(3)))

// Read-only tables addressed uniformly: the constant address space lets the
// compiler load them with scalar loads and use them as scalar operands.
#define CONSTANT_AS __attribute__((address_space(4)))
template <class T> __device__ __forceinline__ const CONSTANT_AS T* as_constant(const T* p) {
  return (const CONSTANT_AS T*)p;
}

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("HIP kernel launch failed (") + what +
                             "): " + hipGetErrorString(e));
}

void record(hipEvent_t ev, hipStream_t s) {
  hipError_t e = hipEventRecord(ev, s);
  if (e != hipSuccess)
    throw std::runtime_error(std::string("hipEventRecord: ") + hipGetErrorString(e));
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l) {
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}

// Source frame of overlap slot group g (OverlapF0Candidates harvest.cpp:417-429):
// g = 0 the frame itself, 1..3 the frames before, 4..6 the frames after.
__device__ __forceinline__ long long overlap_source(long long i, int g) {
  return g == 0 ? i : (g <= 3 ? i - g : i + (g - 3));
}

// Half window of a positive candidate (harvest.cpp:596).
__device__ __forceinline__ int half_window(double fs, double f0) {
  return (int)(1.5 * fs / f0 + 1.0);
}

// Goertzel recurrences of NQ bins for the windowed sample x w (a) and x d
// (b): s_m = x_m + 2 cos(w) s_{m-1} - s_{m-2} from s_{-1} = s_{-2} = 0, so
// that sum_{m<N} x_m e^{-i w m} = e^{-i w (N-1)} (s_{N-1} - e^{-i w} s_{N-2}).
// Steps come in pairs, whose two updates alternate between the two state
// registers (after each pair a1 = s_m, a2 = s_{m-1}), then one last step for
// the odd window length, which leaves s_{N-1} in a2 and s_{N-2} in a1.
template <int NQ> struct Goertzel {
  double c2[NQ];  // 2 cos(w) of each bin
  double a1[NQ], a2[NQ], b1[NQ], b2[NQ];
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int q = 0; q < NQ; ++q) a1[q] = a2[q] = b1[q] = b2[q] = 0.0;
  }
  __device__ __forceinline__ void step2(double xw0, double xd0, double xw1, double xd1) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      a2[q] = fma(c2[q], a1[q], xw0) - a2[q];
      b2[q] = fma(c2[q], b1[q], xd0) - b2[q];
      a1[q] = fma(c2[q], a2[q], xw1) - a1[q];
      b1[q] = fma(c2[q], b2[q], xd1) - b1[q];
    }
  }
  __device__ __forceinline__ void step_last(double xw, double xd) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      a2[q] = fma(c2[q], a1[q], xw) - a2[q];
      b2[q] = fma(c2[q], b1[q], xd) - b2[q];
    }
  }
};

// (power, IF numerator) of GetMeanF0 (harvest.cpp:567-572) from the final
// states (s_{N-1}, s_{N-2} of each signal) of one bin and its record (2 cos w, cos w, sin w, cos w(N-1),
// sin w(N-1)): main spectrum M and derivative spectrum D by the closing
// formula above (WORLD's forward FFT takes e^{-i theta}).
__device__ __forceinline__ double2 goertzel_result(double a1, double a2, double b1, double b2,
                                                   double cw, double sw, double cl, double sl) {
  const double yr = a1 - cw * a2, yi = sw * a2;
  const double zr = b1 - cw * b2, zi = sw * b2;
  const double mre = cl * yr + sl * yi, mim = cl * yi - sl * yr;
  const double dre = cl * zr + sl * zi, dim = cl * zi - sl * zr;
  return make_double2(mre * mre + mim * mim, mre * dim - mim * dre);
}

// FixF0's amplitude sqrt(P) and instantaneous frequency k fs / n + Nu / P
// fs / (2 pi) (harvest.cpp:515-522; 0 where P is 0) of one bin from its
// (power, IF numerator) and k fs / n (the record's entry 5).  They depend
// on the (frame, h, k) alone, so every slot that reads the bin shares them.
__device__ __forceinline__ double2 bin_amp_if(double2 pn, double kfn, double fs) {
  constexpr double kInvTwoPi = 0.5 / kPi;
  const double ifr = pn.x == 0.0 ? 0.0 : kfn + pn.y / pn.x * fs * kInvTwoPi;
  return make_double2(sqrt(pn.x), ifr);
}

// FixF0 (harvest.cpp:507-536) and GetRefinedF0's gate (:610-614) of one
// slot: (refined f0, score) from its candidate code, f0 and 1 / f0, with
// bin(q) the (amplitude, IF) of bin q of its half window's table.  The
// error term's divisions by j + 1 and by f0 (and the frequency's by 2 pi, in
// bin_amp_if) are products with reciprocals: within two ulps of WORLD's
// quotients, far below the spectra's own rounding; the score they feed is
// compared with 2.5, so a slot could change only if its score fell within a
// few ulps of that threshold.
template <class Bin>
__device__ __forceinline__ double2 fix_f0_slot(const Config& c, unsigned long long code,
                                               double f0, double inv_f0, Bin bin) {
  const int H = (int)(code >> 9) & 7;
  double numer = 0.0, den = 0.0, dev = 0.0;
  for (int hh = 0; hh < H; ++hh) {
    const double2 pn = bin((int)(code >> (12 + 7 * hh)) & 127);
    const double amp = pn.x, ifr = pn.y;
    const double inv_h = hh == 0 ? 1.0 : hh == 1 ? 0.5 : hh == 2 ? 1.0 / 3.0
                       : hh == 3 ? 0.25 : hh == 4 ? 0.2 : 1.0 / 6.0;
    numer += amp * ifr;
    den += amp * (hh + 1.0);
    dev += fabs((ifr * inv_h - f0) * inv_f0);
  }
  double refined = numer / (den + kGuard);
  double score = 1.0 / (dev / H + kGuard);
  if (refined < c.f0_floor || refined > c.f0_ceil || score < 2.5) {
    refined = 0.0;
    score = 0.0;
  }
  return make_double2(refined, score);
}

}  // namespace

// ---------------------------------------------------------------------------
// Table construction, one workgroup per half window h: the window rows
// (w[m], d[m]) (window_pair) and, per bin k of h, the record (2 cos w,
// cos w, sin w, cos w(2h), sin w(2h), k fs / n) of w = 2 pi k / n, the
// trigonometric entries from sincospi of the exact rational argument.
__global__ __launch_bounds__(kThreads) void k_bt_tables(Config c, BinTable t, const int* kq,
                                                        double2* win, double* bins) {
  const int hb = blockIdx.x;
  const int4 info = t.info[hb];
  const int h = c.bt_hlo + hb, len = 2 * h + 1, n = info.x, nb = info.y;
  for (int m = threadIdx.x; m < len; m += kThreads) win[info.z + m] = window_pair(m, h);
  for (int q = threadIdx.x; q < nb; q += kThreads) {
    const long long k = kq[hb * c.bt_nbmax + q];
    double sn, cs, sl, cl;
    sincospi(2.0 * (double)(k & (n - 1)) / n, &sn, &cs);
    sincospi(2.0 * (double)((k * (len - 1)) & (n - 1)) / n, &sl, &cl);
    double* r = bins + (long long)(info.w + q) * kBtRecord;
    r[0] = 2.0 * cs;
    r[1] = cs;
    r[2] = sn;
    r[3] = cl;
    r[4] = sl;
    r[5] = (double)k * c.fs_work * ldexp(1.0, -window_fft_log2(h));  // 1 / n exact
    for (int e = 6; e < kBtRecord; ++e) r[e] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Candidate codes: for official candidate jj of frame f (official[f][jj]),
// code = 1 | hb << 1 | H << 9 | q_1 << 12 | ... | q_6 << 42 with hb = h -
// bt_hlo for the half window h (harvest.cpp:596), H the harmonics FixF0 uses
// (:577-578) and q_j the index, among h's table bins, of the bin k_j of
// harmonic j (:515; 7 bits each); 0 for a candidate that is not positive or
// not one of the utterance's count.
// inv[f][jj] = 1 / f0, hbs[f][jj] = hb (0xFF without a code; the row's
// padding to bt_hb_stride bytes is 0xFF too).  A half window or bin outside
// the tables sets the fault flag.  The lane also clears frame f's slots g cnt
// + jj (g < 7) of rf0 and rscore: slots without a candidate stay 0, the
// others are written by k_bt_dft or k_bt_finish.
__global__ __launch_bounds__(kThreads) void k_bt_codes(
    Config c, const long long* f_off, long long total_frames, const double* official,
    const int* count, const int4* harm, unsigned long long* codes, double* inv,
    unsigned char* hbs, double* rf0, double* rscore, unsigned long long* work) {
  const long long e = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (e >= total_frames * c.capacity) return;
  const long long gf = e / c.capacity;
  const int jj = (int)(e - gf * c.capacity);
  const int ui = find_segment(f_off, c.count, gf);
  const double f0 = official[e];
  unsigned long long code = 0;
  double r = 0.0;
  bool fault = false;
  if (jj < count[ui] && f0 > 0.0) {
    const double fs = c.fs_work;
    const int h = half_window(fs, f0), hb = h - c.bt_hlo;
    if (hb < 0 || hb >= c.bt_hr) {
      fault = true;
    } else {
      const int n = window_fft(h), H = min((int)(fs / 2.0 / f0), 6);
      const double qv = f0 * n / fs;
      code = 1ull | (unsigned long long)hb << 1 | (unsigned long long)H << 9;
      for (int j = 0; j < H; ++j) {
        const int k = mround(qv * (j + 1));
        const int4 hr = harm[hb * 6 + j];  // (first bin, bins, first result index)
        if (k - hr.x < 0 || k - hr.x >= hr.y) fault = true;
        code |= (unsigned long long)((hr.z + k - hr.x) & 127) << (12 + 7 * j);
      }
      r = 1.0 / f0;
      if (fault) code = 0;
    }
  }
  codes[e] = code;
  inv[e] = r;
  const int cnt_u = count[ui];
  if (jj < cnt_u)
#pragma unroll
    for (int g = 0; g < 7; ++g) {
      rf0[gf * c.slots + g * cnt_u + jj] = 0.0;
      rscore[gf * c.slots + g * cnt_u + jj] = 0.0;
    }
  const long long stride = bt_hb_stride(c.capacity);
  hbs[gf * stride + jj] = (code & 1ull) ? (unsigned char)((code >> 1) & 0xFF) : 0xFF;
  if (jj == c.capacity - 1)
    for (long long b = c.capacity; b < stride; ++b) hbs[gf * stride + b] = 0xFF;
  if (fault && work) atomicOr(&work[kStatFault * kWorkSpread + (blockIdx.x % kWorkSpread)], 1ull);
}

// ---------------------------------------------------------------------------
// Plan: lane = frame of a block of 64 frames (one wavefront per block).
// needs[f] holds bit (h - bt_hlo) for every half window h of frame f's slots;
// mask[hb][blk] the block's frames whose pairs are listed for h (frames with
// more than kBtPairs windows are left to k_bt_finish), cnt its popcount.
__global__ __launch_bounds__(kThreads) void k_bt_plan(
    Config c, const Utt* utt, const long long* f_off, long long total_frames, long long nblk,
    const unsigned long long* codes, const int* count, unsigned long long* needs,
    unsigned long long* mask, int* cnt) {
  const int lane = threadIdx.x & 63;
  const long long blk = (long long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (blk >= nblk) return;
  const long long gf = blk * 64 + lane;
  unsigned long long bits[kBtWords] = {};
  if (gf < total_frames) {
    const int ui = find_segment(f_off, c.count, gf);
    const int cnt_u = count[ui];
    const long long uf = utt[ui].f_off, frames = utt[ui].frames;
    const long long i = gf - uf;
    for (int g = 0; g < 7; ++g) {
      const long long src = overlap_source(i, g);
      if (src < 0 || src >= frames) continue;
      const unsigned long long* row = codes + (uf + src) * c.capacity;
      for (int jj = 0; jj < cnt_u; ++jj) {
        const unsigned long long code = row[jj];
        if (!(code & 1ull)) continue;
        const int hb = (int)(code >> 1) & 0xFF;
#pragma unroll
        for (int w = 0; w < kBtWords; ++w)
          if ((hb >> 6) == w) bits[w] |= 1ull << (hb & 63);
      }
    }
#pragma unroll
    for (int w = 0; w < kBtWords; ++w) needs[gf * kBtWords + w] = bits[w];
  }
  int pairs = 0;
#pragma unroll
  for (int w = 0; w < kBtWords; ++w) pairs += __popcll(bits[w]);
  if (pairs > kBtPairs)
#pragma unroll
    for (int w = 0; w < kBtWords; ++w) bits[w] = 0;
#pragma unroll
  for (int w = 0; w < kBtWords; ++w) {
    if (w * 64 >= c.bt_hr) break;
    unsigned long long mine = 0;
    for (int b = 0; b < 64; ++b) {
      const unsigned long long bal = __ballot((bits[w] >> b) & 1ull);
      if (lane == b) mine = bal;
    }
    const int hb = w * 64 + lane;
    if (hb < c.bt_hr) {
      mask[hb * nblk + blk] = mine;
      cnt[hb * nblk + blk] = __popcll(mine);
    }
  }
}

// Pair lists: the pairs of h (hb = h - bt_hlo) occupy list positions
// [off[hb nblk], off[(hb + 1) nblk]) in frame order.  Entry (frame,
// utterance << 5 | rank of h among the frame's half windows).
__global__ __launch_bounds__(kThreads) void k_bt_scatter(
    Config c, const long long* f_off, long long total_frames, long long nblk,
    const unsigned long long* mask, const int* off, int2* list) {
  const int lane = threadIdx.x & 63;
  const long long blk = (long long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (blk >= nblk) return;
  const long long gf = blk * 64 + lane;
  const int ui = gf < total_frames ? find_segment(f_off, c.count, gf) : 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  int r = 0;
  for (int base = 0; base < c.bt_hr; base += 64) {
    const int hl = base + lane;
    unsigned long long mk = 0;
    int of = 0;
    if (hl < c.bt_hr) {
      mk = mask[hl * nblk + blk];
      of = off[hl * nblk + blk];
    }
    const int nh = min(64, c.bt_hr - base);
    for (int k = 0; k < nh; ++k) {
      const unsigned long long m = readlane_u64(mk, k);
      if (!m) continue;
      const int o = __builtin_amdgcn_readlane(of, k);
      if ((m >> lane) & 1ull) {
        list[o + __popcll(m & below)] = make_int2((int)gf, (ui << 5) | r);
        ++r;
      }
    }
  }
}

// Chunks of 64 pairs, longest windows first (their chunks cost the most):
// order index k takes h = bt_hlo + bt_hr - 1 - k.  cinfo holds cpre[0..HR]
// (first chunk of each order index, cpre[HR] the total), then the list start
// and end of each order index.  Also clears the chunk counter of k_bt_dft.
__global__ __launch_bounds__(kBtMaxHr) void k_bt_chunks(Config c, long long nblk, const int* off,
                                                        int* cinfo, int* sched) {
  __shared__ int wave_sums[kBtMaxHr / 64];
  const int HR = c.bt_hr, k = threadIdx.x, lane = k & 63, wave = k >> 6;
  int nch = 0, st = 0, en = 0;
  if (k < HR) {
    const int hb = HR - 1 - k;
    st = off[hb * nblk];
    en = off[(hb + 1) * nblk];
    nch = (en - st + 63) >> 6;
  }
  int incl = nch;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  int before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kBtMaxHr / 64; ++w) {
    if (w < wave) before += wave_sums[w];
    total += wave_sums[w];
  }
  if (k < HR) {
    cinfo[k] = before + incl - nch;
    cinfo[HR + 1 + k] = st;
    cinfo[2 * HR + 1 + k] = en;
  }
  if (k == 0) {
    cinfo[HR] = total;
    sched[0] = 0;
  }
}

namespace {

// Steps m, m + 1 (pair) or the last step m of a pass: window rows (uniform,
// (w, d)) applied to the samples.
template <int NQ>
__device__ __forceinline__ void bt_step2(const CONSTANT_AS double* win, int m, double y0,
                                         double y1, Goertzel<NQ>& g) {
  g.step2(y0 * win[2 * m], y0 * win[2 * m + 1], y1 * win[2 * m + 2], y1 * win[2 * m + 3]);
}
template <int NQ>
__device__ __forceinline__ void bt_last(const CONSTANT_AS double* win, int m, double yv,
                                        Goertzel<NQ>& g) {
  g.step_last(yv * win[2 * m], yv * win[2 * m + 1]);
}

// Direct load of 16 bytes per lane from p into LDS at byte address
// lds + 16 lane (global_load_lds_dwordx4; M0 holds the LDS address).  It is
// written as inline assembly so that the compiler's wait-count model, which
// would wait for every such load before any LDS read, leaves the waits to
// bt_wait; it has no register outputs, and the memory clobber keeps the
// compiler's own memory accesses on their side of it.
__device__ __forceinline__ void lds_dma16(const double* p, unsigned lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(p), "s"(lds)
               : "memory", "m0");
}
// Waits until at most one block's direct loads (kBtBlock / 2) or no vector
// memory operations of this wave are outstanding; loads complete in issue
// order.
template <bool kOneBlock> __device__ __forceinline__ void bt_wait() {
  if (kOneBlock) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(kBtBlock / 2) : "memory");
  else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// A pass over the window's len (odd) samples yu[clamp(first + m, 0, last)]:
// blocks of kBtBlock samples are staged in LDS (lane l's samples 2j, 2j + 1
// of a block at buf[(64 j + l) 2]), the next block's loads in flight while
// the current one is summed.  When every lane's blocks lie inside its
// utterance (!clamp) they go to LDS by direct loads, 16 bytes per lane;
// otherwise through registers with the clamped indices.  The last block's
// samples past the window are staged, not used.
template <int NQ>
__device__ __forceinline__ void bt_pass(const double* yu, long long first, long long last,
                                        bool clamp, int len, const CONSTANT_AS double* win,
                                        LDS double* buf0, LDS double* buf1, int lane,
                                        Goertzel<NQ>& g) {
  constexpr int kLoads = kBtBlock / 2;
  const unsigned lds0 = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)buf0);
  const unsigned lds1 = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)buf1);
  // Loads block [m0, m0 + kBtBlock) into buf (LDS address lds).
  auto issue = [&](int m0, unsigned lds, LDS double* buf) {
    if (!clamp) {
#pragma unroll
      for (int j = 0; j < kLoads; ++j) lds_dma16(yu + first + m0 + 2 * j, lds + 1024 * j);
    } else {
#pragma unroll
      for (int j = 0; j < kLoads; ++j) {
        long long i0 = first + m0 + 2 * j, i1 = i0 + 1;
        i0 = i0 < 0 ? 0 : (i0 > last ? last : i0);
        i1 = i1 < 0 ? 0 : (i1 > last ? last : i1);
        const double v0 = yu[i0], v1 = yu[i1];
        buf[128 * j + 2 * lane] = v0;
        buf[128 * j + 2 * lane + 1] = v1;
      }
    }
  };
  // One loop over the sample pairs (the Goertzel state stays in the same
  // registers throughout); at each block's first pair the next block's loads
  // are issued into the other buffer.  Block b is in buf0 when b is even.
  constexpr int kPairs = kBtBlock / 2;
  // The wave's earlier reads of this space (the last pass's samples or bin
  // results) are done before the first block's loads overwrite it.
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  issue(0, lds0, buf0);
  const int npairs = (len - 1) / 2;
  for (int p = 0; p < npairs; ++p) {
    const int blk = p / kPairs, jj = p % kPairs;
    if (jj == 0) {
      // The previous block's reads are done before its buffer is refilled.
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      if ((blk + 1) * kBtBlock < len) {
        if (blk & 1) issue((blk + 1) * kBtBlock, lds0, buf0);
        else issue((blk + 1) * kBtBlock, lds1, buf1);
        bt_wait<true>();
      } else {
        bt_wait<false>();
      }
    }
    const LDS double* v = ((blk & 1) ? buf1 : buf0) + 128 * jj + 2 * lane;
    const double v0 = v[0], v1 = v[1];
    bt_step2<NQ>(win, 2 * p, v0, v1, g);
  }
  // The last sample, m = len - 1 (its block's loads may still be in flight).
  const int m = len - 1, blk = m / kBtBlock, jj = (m % kBtBlock) / 2;
  bt_wait<false>();
  bt_last<NQ>(win, m, ((blk & 1) ? buf1 : buf0)[128 * jj + 2 * lane], g);
}

template <int NQ>
__device__ __forceinline__ void bt_chunk(bool clamp, const double* yu, long long first,
                                         long long last, int len,
                                         const CONSTANT_AS double* win,
                                         const CONSTANT_AS double* rec, LDS double* buf0,
                                         LDS double* buf1, int lane, double2* out, bool act,
                                         double fs) {
  Goertzel<NQ> g;
#pragma unroll
  for (int q = 0; q < NQ; ++q) g.c2[q] = rec[q * kBtRecord];
  g.clear();
  bt_pass<NQ>(yu, first, last, clamp, len, win, buf0, buf1, lane, g);
  if (act)
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const CONSTANT_AS double* r = rec + q * kBtRecord;
      out[q] = bin_amp_if(goertzel_result(g.a2[q], g.a1[q], g.b2[q], g.b1[q], r[1], r[2], r[3],
                                          r[4]),
                          r[5], fs);
    }
}

// LDS per wavefront of k_bt_dft (doubles): two staging buffers of kBtBlock
// samples per lane, later the (amplitude, IF) of up to kBtPassBins bins per lane.
constexpr int kBtStage = (kBtBlock > kBtPassBins ? kBtBlock : kBtPassBins) * 128;

// Matched slots whose candidates one lane loads together (bt_pair_slots).
constexpr int kBtSlotBatch = 8;

// Where the slots of a pair find their candidates and put their results.
struct SlotArgs {
  const int* count;
  const unsigned long long* codes;
  const double* official;
  const double* inv;
  const unsigned char* hbs;
  double* rf0;
  double* rscore;
};

// The slots of frame i (utterance frames [uf, uf + frames)) whose candidate
// gives half window hb are those of the official candidates of frames i - 3
// .. i + 3 (OverlapF0Candidates harvest.cpp:417-429, slot g cnt + jj for
// candidate jj of source group g) with that half window.  hbs rows hold
// each candidate's hb as a byte: the seven rows' bytes of candidates j0 ..
// j0 + 15 are loaded together and turned into a bitmask of matches by
// zero-byte tests of the rows XOR hb; bit 16 (g & 3) + j of m[g >> 2] is
// candidate j0 + j of group g.
__device__ __forceinline__ void bt_match_mask(const Config& c, const SlotArgs& a, long long i,
                                              long long uf, long long frames, int cnt_u, int hb,
                                              int j0, unsigned long long m[2]) {
  const long long stride = bt_hb_stride(c.capacity);
  const unsigned pat = (unsigned)hb * 0x01010101u;
  const unsigned valid = j0 >= cnt_u ? 0u
                       : cnt_u - j0 >= 16 ? 0xFFFFu : (1u << (cnt_u - j0)) - 1u;
  m[0] = m[1] = 0ull;
#pragma unroll
  for (int g = 0; g < 7; ++g) {
    const long long src = overlap_source(i, g);
    uint4 w = make_uint4(~0u, ~0u, ~0u, ~0u);
    if (valid && src >= 0 && src < frames)
      w = reinterpret_cast<const uint4*>(a.hbs + (uf + src) * stride)[j0 >> 4];
    const unsigned words[4] = {w.x, w.y, w.z, w.w};
    unsigned bits = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const unsigned x = words[d] ^ pat;
      // 0x80 in each byte of x that is zero (no carry crosses a byte).
      const unsigned z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
      bits |= (((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u))
              << (4 * d);
    }
    m[g >> 2] |= (unsigned long long)(bits & valid) << (16 * (g & 3));
  }
}

// Position of the n-th (from 0) set bit of w, n < popcount(w).
__device__ __forceinline__ int select_bit(unsigned long long w, int n) {
  int pos = 0;
#pragma unroll
  for (int width = 32; width > 0; width >>= 1) {
    const int cnt = __popcll((w >> pos) & ((1ull << width) - 1ull));
    if (n >= cnt) {
      n -= cnt;
      pos += width;
    }
  }
  return pos;
}

__device__ __forceinline__ unsigned long long shfl_u64(unsigned long long v, int l) {
  const unsigned lo = __shfl((unsigned)v, l, 64), hi = __shfl((unsigned)(v >> 32), l, 64);
  return ((unsigned long long)hi << 32) | lo;
}

// FixF0 of every slot of the wave's pairs (lane l: frame gf, half window
// hb, active when act) whose bins are in LDS, rr[2 (64 q + l) + {0, 1}]:
// the lanes' matches (bt_match_mask) are numbered across the wave by a
// prefix sum of their counts, and each round every lane takes one of them
// (its owner lane by a binary search of the prefix sums, its slot by the
// owner's mask), so that the wave's rounds are its total match count / 64
// rather than its largest per-lane count.
__device__ __forceinline__ void bt_wave_slots(const Config& c, const SlotArgs& a, long long gf,
                                              long long uf, long long frames, int cnt_u, int hb,
                                              bool act, int lane, const LDS double* rr) {
  const long long i = gf - uf;
  int jmax = act ? cnt_u : 0;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) jmax = max(jmax, __shfl_xor(jmax, d, 64));
  for (int j0 = 0; j0 < jmax; j0 += 16) {
    unsigned long long m[2];
    bt_match_mask(c, a, i, uf, frames, act ? cnt_u : 0, hb, j0, m);
    const int cnt = __popcll(m[0]) + __popcll(m[1]);
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(incl, d, 64);
      if (lane >= d) incl += o;
    }
    const int excl = incl - cnt;
    const int total = __shfl(incl, 63, 64);
    for (int k0 = 0; k0 < total; k0 += 64) {
      const int k = k0 + lane;
      int o = 0;  // owner: the last lane whose first match index is <= k
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(excl, o + step, 64) <= k) o += step;
      const int n = k - __shfl(excl, o, 64);
      const unsigned long long m0 = shfl_u64(m[0], o), m1 = shfl_u64(m[1], o);
      const long long gf_o = (long long)shfl_u64((unsigned long long)gf, o);
      const long long uf_o = (long long)shfl_u64((unsigned long long)uf, o);
      const int cnt_o = __shfl(cnt_u, o, 64);
      if (k >= total) continue;
      const int c0 = __popcll(m0);
      const bool upper = n >= c0;
      const int b = select_bit(upper ? m1 : m0, upper ? n - c0 : n);
      const int g = 4 * upper + (b >> 4), jj = j0 + (b & 15);
      const long long e = (uf_o + overlap_source(gf_o - uf_o, g)) * c.capacity + jj;
      const double2 v = fix_f0_slot(c, a.codes[e], a.official[e], a.inv[e], [&](int q) {
        const int at = 2 * (64 * q + o);
        return make_double2(rr[at], rr[at + 1]);
      });
      const long long sl = gf_o * c.slots + g * cnt_o + jj;
      a.rf0[sl] = v.x;
      a.rscore[sl] = v.y;
    }
  }
}

// FixF0 of every slot of the lane's pair (frame gf, half window hb), one
// lane alone (bt_match_mask, then the matches in order); bin(q) is the
// pair's bin q.
template <class Bin>
__device__ __forceinline__ void bt_pair_slots(const Config& c, const SlotArgs& a, long long gf,
                                              long long uf, long long frames, int cnt_u, int hb,
                                              Bin bin) {
  const long long i = gf - uf;
  double* out_f0 = a.rf0 + gf * c.slots;
  double* out_sc = a.rscore + gf * c.slots;
  for (int j0 = 0; j0 < cnt_u; j0 += 16) {
    unsigned long long m[2];
    bt_match_mask(c, a, i, uf, frames, cnt_u, hb, j0, m);
    // The matches in order (slot g cnt + jj), kBtSlotBatch at a time: the
    // batch's candidates are all loaded before any FixF0, so that one load
    // latency serves the batch.
    auto take = [&](int& sl, long long& e) -> bool {
      if (!(m[0] | m[1])) return false;
      const int half = m[0] ? 0 : 1;
      const unsigned long long mh = half ? m[1] : m[0];
      const int b = __builtin_ctzll(mh);
      if (half) m[1] = mh & (mh - 1);
      else m[0] = mh & (mh - 1);
      const int g = 4 * half + (b >> 4), jj = j0 + (b & 15);
      sl = g * cnt_u + jj;
      e = (uf + overlap_source(i, g)) * c.capacity + jj;
      return true;
    };
    while (m[0] | m[1]) {
      int sl[kBtSlotBatch];
      bool ok[kBtSlotBatch];
      unsigned long long code[kBtSlotBatch];
      double f0[kBtSlotBatch], inv_f0[kBtSlotBatch];
#pragma unroll
      for (int k = 0; k < kBtSlotBatch; ++k) {
        long long e = 0;
        sl[k] = 0;
        ok[k] = take(sl[k], e);
        code[k] = 0;
        f0[k] = inv_f0[k] = 0.0;
        if (ok[k]) {
          code[k] = a.codes[e];
          f0[k] = a.official[e];
          inv_f0[k] = a.inv[e];
        }
      }
#pragma unroll
      for (int k = 0; k < kBtSlotBatch; ++k)
        if (ok[k]) {
          const double2 v = fix_f0_slot(c, code[k], f0[k], inv_f0[k], bin);
          out_f0[sl[k]] = v.x;
          out_sc[sl[k]] = v.y;
        }
    }
  }
}

// A pair whose half window has at most kBtPassBins bins: one pass over the
// window, the bins' (amplitude, IF) to LDS (over the staging space, whose
// reads are done), then FixF0 of the pair's slots from them.
template <int NQ>
__device__ __forceinline__ void bt_chunk_slots(bool clamp, const double* yu, long long first,
                                               long long last, int len,
                                               const CONSTANT_AS double* win,
                                               const CONSTANT_AS double* rec, LDS double* buf0,
                                               LDS double* buf1, int lane, bool act, double fs,
                                               const Config& c, const SlotArgs& a, long long gf,
                                               long long uf, long long frames, int cnt_u, int hb) {
  static_assert(128 * NQ <= kBtStage, "bin results exceed the staging space");
  Goertzel<NQ> g;
#pragma unroll
  for (int q = 0; q < NQ; ++q) g.c2[q] = rec[q * kBtRecord];
  g.clear();
  bt_pass<NQ>(yu, first, last, clamp, len, win, buf0, buf1, lane, g);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const CONSTANT_AS double* r = rec + q * kBtRecord;
    const double2 v = bin_amp_if(
        goertzel_result(g.a2[q], g.a1[q], g.b2[q], g.b1[q], r[1], r[2], r[3], r[4]), r[5], fs);
    buf0[2 * (64 * q + lane)] = v.x;
    buf0[2 * (64 * q + lane) + 1] = v.y;
  }
  bt_wave_slots(c, a, gf, uf, frames, cnt_u, hb, act, lane, buf0);
}

}  // namespace

// One wavefront per chunk of 64 pairs of one h.  A workgroup takes four
// consecutive chunks at a time from a counter until none is left (a
// persistent grid): its waves then mostly share one h, so that they stream
// the same table rows through the scalar cache together.  Then the pair's
// slots are finished here (bt_pair_slots writes their refined f0 and
// score) from the (amplitude, instantaneous frequency) of its bins
// (bin_amp_if): in LDS when h has at most kBtPassBins bins (one pass),
// otherwise in res[(frame kBtPairs + rank) bt_nbmax + q], written by the
// passes and read back by the same lane.
__global__ __launch_bounds__(kThreads) void k_bt_dft(
    Config c, const Utt* utt, const double* y, const int2* list, const int* cinfo, int* sched,
    BinTable t, double2* res, SlotArgs slots, unsigned long long* work) {
  __shared__ int s_info[3 * kBtMaxHr + 1];
  __shared__ int s_next;
  // Per wave: two sample staging buffers, later the bin results.
  __shared__ double stage[kThreads / 64][kBtStage];
  const int HR = c.bt_hr;
  for (int k = threadIdx.x; k < 3 * HR + 1; k += kThreads) s_info[k] = cinfo[k];
  __syncthreads();
  const int* cpre = s_info;
  const int* hstart = s_info + HR + 1;
  const int* hend = s_info + 2 * HR + 1;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  LDS double* buf0 = (LDS double*)stage[wave];
  LDS double* buf1 = buf0 + kBtBlock * 64;
  const int total = cpre[HR];
  const int spf = (int)(c.fs_work / 1000.0);  // samples per 1 ms frame (whole)
  unsigned long long wsamp = 0, bsamp = 0, nbins = 0, npairs = 0;
  const CONSTANT_AS double* wins = as_constant(reinterpret_cast<const double*>(t.win));
  const CONSTANT_AS double* bins = as_constant(t.bins);
  while (true) {
    if (threadIdx.x == 0) s_next = atomicAdd(sched, 1) * (kThreads / 64);
    __syncthreads();
    const int first_ch = s_next;
    __syncthreads();
    if (first_ch >= total) break;  // uniform across the workgroup
    const int ch = first_ch + wave;
    if (ch >= total) continue;
    int lo = 0, hi = HR - 1;  // order index: largest k with cpre[k] <= ch
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cpre[mid] <= ch) lo = mid;
      else hi = mid - 1;
    }
    const int k = __builtin_amdgcn_readfirstlane(lo);
    const int hb = HR - 1 - k, h = c.bt_hlo + hb, len = 2 * h + 1;
    const int pos0 = hstart[k] + ((ch - cpre[k]) << 6);
    const int pos = pos0 + lane;
    const bool act = pos < hend[k];
    // Inactive lanes repeat lane 0's pair (always active) and store nothing.
    const int2 e = list[act ? pos : pos0];
    const int ui = e.y >> 5, r = e.y & 31;
    const long long gf = e.x;
    const long long yoff = utt[ui].y_off, uf = utt[ui].f_off;
    const long long last = (long long)utt[ui].n - 1;
    const long long first = (long long)spf * (gf - uf) - h - 1;  // GetBaseIndex - 1
    const int rows = (len + kBtBlock - 1) / kBtBlock * kBtBlock;  // samples staged
    const bool clamp = __ballot(first < 0 || first + rows - 1 > last) != 0ull;
    const double* yu = y + yoff;
    const int4 info = t.info[hb];
    const int nb = __builtin_amdgcn_readfirstlane(info.y);
    const CONSTANT_AS double* win = wins + 2LL * __builtin_amdgcn_readfirstlane(info.z);
    const CONSTANT_AS double* rec0 = bins + (long long)kBtRecord * __builtin_amdgcn_readfirstlane(info.w);
    static_assert(kBtPassBins <= 12, "the pass switches below have cases for up to 12 bins");
    if (nb <= kBtPassBins) {
      const int cnt_u = slots.count[ui];
      const long long frames = utt[ui].frames;
#define BT_SLOTS(NQ)                                                                      \
  case NQ:                                                                               \
    bt_chunk_slots<NQ>(clamp, yu, first, last, len, win, rec0, buf0, buf1, lane, act,    \
                       c.fs_work, c, slots, gf, uf, frames, cnt_u, hb);                  \
    break;
      switch (nb) {
        BT_SLOTS(1) BT_SLOTS(2) BT_SLOTS(3) BT_SLOTS(4)
        BT_SLOTS(5) BT_SLOTS(6) BT_SLOTS(7) BT_SLOTS(8)
        BT_SLOTS(9) BT_SLOTS(10) BT_SLOTS(11) BT_SLOTS(12)
        default: break;
      }
#undef BT_SLOTS
    }
    double2* out = res + (gf * kBtPairs + r) * c.bt_nbmax;
    // Passes of at most kBtPassBins bins, split evenly.
    const int npass = nb <= kBtPassBins ? 0 : (nb + kBtPassBins - 1) / kBtPassBins;
    for (int p = 0; p < npass; ++p) {
      const int q0 = p * nb / npass, nq = (p + 1) * nb / npass - q0;
      const CONSTANT_AS double* rec = rec0 + (long long)q0 * kBtRecord;
#define BT_CHUNK(NQ)                                                                        \
  case NQ:                                                                                 \
    bt_chunk<NQ>(clamp, yu, first, last, len, win, rec, buf0, buf1, lane, out + q0, act, \
                 c.fs_work);                                                                \
    break;
      switch (nq) {
        BT_CHUNK(1) BT_CHUNK(2) BT_CHUNK(3) BT_CHUNK(4)
        BT_CHUNK(5) BT_CHUNK(6) BT_CHUNK(7) BT_CHUNK(8)
        BT_CHUNK(9) BT_CHUNK(10) BT_CHUNK(11) BT_CHUNK(12)
        default: break;
      }
#undef BT_CHUNK
    }
    if (npass > 0 && act)
      bt_pair_slots(c, slots, gf, uf, utt[ui].frames, slots.count[ui], hb,
                    [&](int q) { return out[q]; });
    const unsigned long long na = (unsigned long long)__popcll(__ballot(act));
    npairs += na;
    wsamp += na * len;
    bsamp += na * len * nb;
    nbins += na * nb;
  }
  if (work && lane == 0) {
    const int spread = blockIdx.x % kWorkSpread;
    atomicAdd(&work[kStatWindowSamples * kWorkSpread + spread], wsamp);
    atomicAdd(&work[kStatBinSamples * kWorkSpread + spread], bsamp);
    atomicAdd(&work[kStatBins * kWorkSpread + spread], nbins);
    atomicAdd(&work[kStatTablePairs * kWorkSpread + spread], npairs);
  }
}

namespace {

// Bin q of half window h at frame `first` computed alone, by the same
// operations as k_bt_dft (frames with more than kBtPairs windows).
__device__ double2 bt_bin_direct(const double* yu, long long first, long long last, int h,
                                 const double2* win, const double* rec, double fs) {
  const int len = 2 * h + 1;
  Goertzel<1> g;
  g.c2[0] = rec[0];
  g.clear();
  auto at = [&](int m) -> double {
    long long i = first + m;
    i = i < 0 ? 0 : (i > last ? last : i);
    return yu[i];
  };
  int m = 0;
  for (; m + 1 < len; m += 2) {
    const double y0 = at(m), y1 = at(m + 1);
    g.step2(y0 * win[m].x, y0 * win[m].y, y1 * win[m + 1].x, y1 * win[m + 1].y);
  }
  const double yv = at(m);
  g.step_last(yv * win[m].x, yv * win[m].y);
  return bin_amp_if(
      goertzel_result(g.a2[0], g.a1[0], g.b2[0], g.b1[0], rec[1], rec[2], rec[3], rec[4]), rec[5],
      fs);
}

}  // namespace

// FixF0 (harvest.cpp:507-536) and GetRefinedF0's gate (:610-614) of the
// frames k_bt_dft leaves: those with more than kBtPairs half windows, whose
// pairs are not listed.  One wavefront per frame (lanes over slots); the
// bins are computed here (bt_bin_direct); other frames return at once.
__global__ __launch_bounds__(kThreads) void k_bt_finish(
    Config c, const Utt* utt, const long long* f_off, long long total_frames, const double* y,
    const double* official, const unsigned long long* codes, const double* inv,
    const int* count, const unsigned long long* needs, BinTable t, double* rf0,
    double* rscore) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const long long gf = (long long)blockIdx.x * (kThreads / 64) + wave;
  if (gf >= total_frames) return;
  int pairs = 0;
#pragma unroll
  for (int w = 0; w < kBtWords; ++w) pairs += __popcll(needs[gf * kBtWords + w]);
  if (pairs <= kBtPairs) return;
  const int ui = find_segment(f_off, c.count, gf);
  const int cnt_u = count[ui];
  const int ncand = 7 * cnt_u;
  const long long uf = utt[ui].f_off, frames = utt[ui].frames;
  const long long i = gf - uf;
  const double fs = c.fs_work;
  double* out_f0 = rf0 + gf * c.slots;
  double* out_sc = rscore + gf * c.slots;
  const int spf = (int)(fs / 1000.0);
  for (int s = lane; s < ncand; s += 64) {
    const int g = s / cnt_u, jj = s - g * cnt_u;
    const long long src = overlap_source(i, g);
    const long long e = (uf + src) * c.capacity + jj;
    const unsigned long long code = (src >= 0 && src < frames) ? codes[e] : 0ull;
    if (!(code & 1ull)) continue;  // cleared by k_bt_codes
    const int hb = (int)(code >> 1) & 0xFF, h = c.bt_hlo + hb;
    const int4 info = t.info[hb];
    const double2 v = fix_f0_slot(c, code, official[e], inv[e], [&](int q) {
      return bt_bin_direct(y + utt[ui].y_off, (long long)spf * i - h - 1,
                           (long long)utt[ui].n - 1, h, t.win + info.z,
                           t.bins + (long long)(info.w + q) * kBtRecord, fs);
    });
    out_f0[s] = v.x;
    out_sc[s] = v.y;
  }
}

// ---------------------------------------------------------------------------
// Launchers.

void launch_bt_tables(hipStream_t s, const Config& c, const BinTable& t, const int* kq,
                      double2* win, double* bins) {
  if (c.bt_hr <= 0) return;
  k_bt_tables<<<c.bt_hr, kThreads, 0, s>>>(c, t, kq, win, bins);
  check_launch("bt_tables");
}

long long bt_blocks(long long total_frames) { return (total_frames + 63) / 64; }

void launch_refine_table(hipStream_t s, const Config& c, const Utt* utt, const long long* f_off,
                         long long total_frames, const double* y, const double* official,
                         const int* count, const BinTable& t, const BtBufs& b, double* rf0,
                         double* rscore, unsigned long long* work, hipEvent_t dft_begin,
                         hipEvent_t dft_end) {
  if (c.bt_hr <= 0 || c.bt_hr > kBtMaxHr) throw std::runtime_error("refine_table: no bin table");
  if (total_frames <= 0) return;
  const long long nblk = bt_blocks(total_frames);
  const unsigned grid_blk = (unsigned)((nblk + kThreads / 64 - 1) / (kThreads / 64));
  const long long entries = total_frames * c.capacity;
  k_bt_codes<<<(unsigned)((entries + kThreads - 1) / kThreads), kThreads, 0, s>>>(
      c, f_off, total_frames, official, count, t.harm, b.codes, b.inv, b.hbs, rf0, rscore, work);
  check_launch("bt_codes");
  k_bt_plan<<<grid_blk, kThreads, 0, s>>>(c, utt, f_off, total_frames, nblk, b.codes, count,
                                          b.needs, b.mask, b.cnt);
  check_launch("bt_plan");
  launch_exclusive_scan(s, b.cnt, b.off, (long long)c.bt_hr * nblk, b.scan_tmp, b.scan_len);
  k_bt_scatter<<<grid_blk, kThreads, 0, s>>>(c, f_off, total_frames, nblk, b.mask, b.off, b.list);
  check_launch("bt_scatter");
  k_bt_chunks<<<1, kBtMaxHr, 0, s>>>(c, nblk, b.off, b.cinfo, b.sched);
  check_launch("bt_chunks");
  // A persistent grid: enough workgroups to fill the device, fewer for short
  // analyses (at most one per four possible chunks: kBtPairs pairs per
  // frame, 64 per chunk, a partial chunk per h).
  long long groups = (total_frames * kBtPairs / 64 + c.bt_hr + 3) / 4;
  const long long cap = c.bt_groups > 0 ? c.bt_groups : kBtDftGroups;
  groups = groups < 64 ? 64 : (groups > cap ? cap : groups);
  if (dft_begin) record(dft_begin, s);
  const SlotArgs slots{count, b.codes, official, b.inv, b.hbs, rf0, rscore};
  k_bt_dft<<<(unsigned)groups, kThreads, 0, s>>>(c, utt, y, b.list, b.cinfo, b.sched, t, b.res,
                                                 slots, work);
  check_launch("bt_dft");
  if (dft_end) record(dft_end, s);
  k_bt_finish<<<(unsigned)((total_frames + 3) / 4), kThreads, 0, s>>>(
      c, utt, f_off, total_frames, y, official, b.codes, b.inv, count, b.needs, t, rf0, rscore);
  check_launch("bt_finish");
}

}  // namespace hv
