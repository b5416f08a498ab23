// Hand-written gfx950 kernels of the Harvest hot path.  All arithmetic is FP64
// and the file is compiled with -ffp-contract=off: wherever WORLD's expression
// order matters for a discrete decision it is kept, and fma() appears only in
// accumulations whose order WORLD's FFTs do not define anyway.
//
// Every kernel processes a batch of utterances laid out by hv::Utt; a plain
// call is a batch of one, and every per-utterance reduction has a fixed order
// that does not depend on the batch, so batched results are bitwise equal to
// plain ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "kernels.h"

namespace hv {

// LDS-qualified pointers: accesses through them compile to ds_* operations
// however the pointers are passed around.
#define LDS __attribute__((address_space(3)))
namespace {

void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("HIP kernel launch failed (") + what +
                             "): " + hipGetErrorString(e));
}

unsigned grid_for(long long n) {
  return (unsigned)((n + kThreads - 1) / kThreads);
}

// Fixed-order tree reduction of one double per thread (result in thread 0).
__device__ double block_sum(double v, double* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride) lds[threadIdx.x] += lds[threadIdx.x + stride];
    __syncthreads();
  }
  double r = lds[0];
  __syncthreads();
  return r;
}

// Exclusive scan over the workgroup of an unsigned 64-bit value; `total`
// receives the sum.  Wave64 shuffles, then one LDS pass over the 4 waves.
__device__ unsigned long long block_scan_u64(unsigned long long v,
                                             unsigned long long* wave_sums,
                                             unsigned long long& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    unsigned long long o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  unsigned long long before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) {
    unsigned long long s = wave_sums[w];
    if (w < wave) before += s;
    all += s;
  }
  total = all;
  __syncthreads();
  return before + incl - v;
}

// Inclusive prefix sum of v over the wave by DPP: row shifts by 1, 2, 4, 8
// within each 16-lane row, then the row totals broadcast to the rows after
// them (row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3).
__device__ __forceinline__ unsigned wave_scan_u32(unsigned v) {
  v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);
  return v;
}

// block_scan_u64 for values of 16-bit fields whose totals over the block
// stay below 2^16: the two 32-bit halves never carry into each other, so
// each is scanned on its own.
__device__ unsigned long long block_scan_fields(unsigned long long v,
                                                unsigned long long* wave_sums,
                                                unsigned long long& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned lo = wave_scan_u32((unsigned)v), hi = wave_scan_u32((unsigned)(v >> 32));
  const unsigned long long incl = ((unsigned long long)hi << 32) | lo;
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  unsigned long long before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) {
    unsigned long long s = wave_sums[w];
    if (w < wave) before += s;
    all += s;
  }
  total = all;
  __syncthreads();
  return before + incl - v;
}

constexpr int kXcds = 8;  // XCDs that workgroups are dispatched to in turn

// Workgroup b of G is dispatched to XCD b mod 8; renumbered so that each XCD
// takes one contiguous range of logical blocks (its L2 then holds one slice
// of the data those blocks share).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G) {
  const unsigned x = b % kXcds, per = G / kXcds, rem = G % kXcds;
  return x * per + min(x, rem) + b / kXcds;
}

// ---------------------------------------------------------------------------
// Decimation.  WORLD runs a 3rd-order IIR forward over the padded signal,
// then again over the reversed output, and keeps every r-th sample
// (matlabfunctions.cpp:178-204 with the padding of harvest.cpp:50-62).  The
// IIR state s = (w_{t-1}, w_{t-2}, w_{t-3}) evolves as s' = A s + x e1, so a
// chunk of kChunk samples maps s to P s + z with P = A^kChunk.  Phase 1
// computes z per chunk, phase 2 scans the chunk carries per utterance, phase
// 3 replays every chunk from its exact carried state with WORLD's expression.

// Pass 0 reads the replicate-padded input: WORLD pads x by `lag` copies of its
// end samples and decimate() adds 9 odd reflections that, on a constant run,
// equal the end samples again (2a - a = a exactly).  Pass 1 reads pass 0's
// output backwards.
//
// The transpose (for the adjoint): decimation is y = S F R F P x with P the
// padding, F the causal IIR from rest, R the reversal and S the selection of
// every r-th sample.  F^T = R F R exactly, so x_bar = P^T R F R F R S^T y_bar:
// pass 2 reads S^T y_bar reversed (the gradient scattered to pass-1 indices),
// pass 3 reads pass 2's output backwards, and k_iir_fold applies R and P^T.
__device__ __forceinline__ long long select_base(const Config& c, const Utt& u) {
  // decimate(): y[count] = tmp1[nbeg + count r + 8] of the twice-reversed
  // output, i.e. pass-1 index m - 9 - nbeg - count r; harvest.cpp:62 keeps
  // count = lag / r + j.
  return u.m - 9 - u.nbeg - (long long)(c.lag / c.ratio) * c.ratio;
}

// Orders this wave's LDS accesses: writes by any lane before it are visible
// to reads by any lane after it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sequential loops load this many inputs ahead, so that the loads' latency
// overlaps instead of stalling every step of the recursion.
constexpr int kIirAhead = 8;

template <int PASS>
__device__ __forceinline__ double iir_input(const Config& c, const Utt& u,
                                            const double* x, const double* w,
                                            long long i) {
  if (PASS == 0) {
    long long j = i - 9 - c.lag;
    j = j < 0 ? 0 : (j >= u.length ? u.length - 1 : j);
    return x[u.x_off + j];
  }
  if (PASS == 2) {  // x is the y-gradient here
    const long long d = select_base(c, u) - (u.m - 1 - i);
    if (d >= 0 && d % c.ratio == 0 && d / c.ratio < u.n) return x[u.y_off + d / c.ratio];
    return 0.0;
  }
  return w[u.m_off + (u.m - 1 - i)];
}

template <int PASS>
__global__ __launch_bounds__(kThreads) void k_iir_chunk_state(
    Config c, const Utt* utt, const long long* c_off, long long total,
    const double* x, const double* w, double* states) {
  long long cg = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (cg >= total) return;
  const Utt u = utt[find_segment(c_off, c.count, cg)];
  long long start = (cg - u.c_off) * kChunk;
  long long end = min(start + (long long)kChunk, u.m);
  double w0 = 0.0, w1 = 0.0, w2 = 0.0;
  // Inputs are loaded kAhead at a time so that their latency overlaps.
  for (long long i0 = start; i0 < end; i0 += kIirAhead) {
    double in[kIirAhead];
#pragma unroll
    for (int q = 0; q < kIirAhead; ++q)
      in[q] = i0 + q < end ? iir_input<PASS>(c, u, x, w, i0 + q) : 0.0;
#pragma unroll
    for (int q = 0; q < kIirAhead; ++q) {
      if (i0 + q >= end) break;
      double wt = in[q] + c.a0 * w0 + c.a1 * w1 + c.a2 * w2;
      w2 = w1;
      w1 = w0;
      w0 = wt;
    }
  }
  states[cg * 3 + 0] = w0;
  states[cg * 3 + 1] = w1;
  states[cg * 3 + 2] = w2;
}

__device__ __forceinline__ void mat3_mul(const double* a, const double* b, double* r) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r[i * 3 + j] = a[i * 3 + 0] * b[0 * 3 + j] + a[i * 3 + 1] * b[1 * 3 + j] +
                     a[i * 3 + 2] * b[2 * 3 + j];
}

__device__ __forceinline__ void mat3_apply(const double* m, const double* v,
                                           const double* add, double* r) {
  double t0 = m[0] * v[0] + m[1] * v[1] + m[2] * v[2] + add[0];
  double t1 = m[3] * v[0] + m[4] * v[1] + m[5] * v[2] + add[1];
  double t2 = m[6] * v[0] + m[7] * v[1] + m[8] * v[2] + add[2];
  r[0] = t0;
  r[1] = t1;
  r[2] = t2;
}

// Carry-in state of every chunk from the chunk end states (zero start), in
// three parallel steps.  An utterance's chunks form blocks of kCarryBlock =
// 256 runs of kCarryRun chunks.
//   k_carry_blocks<false>: per block, the end state from a zero start (runs
//     by Horner, then a Kogge-Stone scan over runs with (P^kCarryRun)^(2^j));
//   k_carry_top: per utterance, the block start states, sequentially with
//     Q = P^(kCarryRun * 256);
//   k_carry_blocks<true>: per block, each run's start state
//     P^(kCarryRun t) cin + (scan of the runs before it), then the run
//     replays its chunks and writes their carry-in states.
constexpr int kCarryRun = 4, kCarryBlock = kThreads * kCarryRun;
static_assert(kCarryBlock == 1024, "k_carry_top forms P^1024 by ten squarings");

__device__ void carry_powers(const double* P, double (*qpow)[9]) {
  // qpow[0] = P^kCarryRun, qpow[j] = qpow[j - 1]^2 (qpow[8] = Q).
  double t[9];
  mat3_mul(P, P, t);
  mat3_mul(t, t, qpow[0]);
  for (int j = 1; j < 9; ++j) mat3_mul(qpow[j - 1], qpow[j - 1], qpow[j]);
}

template <bool kApply>
__global__ __launch_bounds__(kThreads) void k_carry_blocks(Config c, const Utt* utt,
                                                           const double* states, int max_blocks,
                                                           double* block_agg,
                                                           const double* block_cin,
                                                           double* carries) {
  __shared__ double qpow[9][9];
  __shared__ double scan[kThreads][3];
  const int ui = blockIdx.y, b = blockIdx.x;
  const Utt u = utt[ui];
  const long long C = u.chunks;
  const long long b0 = (long long)b * kCarryBlock;
  if (b0 >= C) return;  // uniform over the workgroup
  const double* P = c.chunk_pow;
  if (threadIdx.x == 0) carry_powers(P, qpow);
  const long long first = b0 + (long long)threadIdx.x * kCarryRun;
  const long long last = min(first + kCarryRun, C);
  const double* z = states + u.c_off * 3;
  double zz[kCarryRun][3];
#pragma unroll
  for (int q = 0; q < kCarryRun; ++q)
    for (int r = 0; r < 3; ++r) zz[q][r] = first + q < last ? z[(first + q) * 3 + r] : 0.0;
  double agg[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < kCarryRun; ++q)
    if (first + q < last) mat3_apply(P, agg, zz[q], agg);
  scan[threadIdx.x][0] = agg[0];
  scan[threadIdx.x][1] = agg[1];
  scan[threadIdx.x][2] = agg[2];
  __syncthreads();
  for (int j = 0; j < 8; ++j) {
    const int d = 1 << j;
    double nv[3] = {scan[threadIdx.x][0], scan[threadIdx.x][1], scan[threadIdx.x][2]};
    if ((int)threadIdx.x >= d) {
      double prev[3] = {scan[threadIdx.x - d][0], scan[threadIdx.x - d][1],
                        scan[threadIdx.x - d][2]};
      mat3_apply(qpow[j], prev, nv, nv);
    }
    __syncthreads();
    scan[threadIdx.x][0] = nv[0];
    scan[threadIdx.x][1] = nv[1];
    scan[threadIdx.x][2] = nv[2];
    __syncthreads();
  }
  const long long slot = ((long long)ui * max_blocks + b) * 3;
  if (!kApply) {
    // The end state of a full block (a partial last block's is never used).
    if (threadIdx.x == kThreads - 1)
      for (int r = 0; r < 3; ++r) block_agg[slot + r] = scan[threadIdx.x][r];
    return;
  }
  if (first >= C) return;
  // P^(kCarryRun t) cin from the binary digits of t, then the runs before t.
  double carry[3] = {block_cin[slot], block_cin[slot + 1], block_cin[slot + 2]};
  const double zero[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < 8; ++j)
    if ((threadIdx.x >> j) & 1) mat3_apply(qpow[j], carry, zero, carry);
  if (threadIdx.x > 0)
    for (int r = 0; r < 3; ++r) carry[r] += scan[threadIdx.x - 1][r];
  double* out = carries + u.c_off * 3;
#pragma unroll
  for (int q = 0; q < kCarryRun; ++q) {
    const long long k = first + q;
    if (k >= last) break;
    out[k * 3 + 0] = carry[0];
    out[k * 3 + 1] = carry[1];
    out[k * 3 + 2] = carry[2];
    mat3_apply(P, carry, zz[q], carry);
  }
}

// One thread per utterance: block start states in block order.
__global__ __launch_bounds__(kThreads) void k_carry_top(Config c, const Utt* utt,
                                                        int max_blocks,
                                                        const double* block_agg,
                                                        double* block_cin) {
  const int ui = blockIdx.x * kThreads + threadIdx.x;
  if (ui >= c.count) return;
  const Utt u = utt[ui];
  const int nb = (int)((u.chunks + kCarryBlock - 1) / kCarryBlock);
  double q[9], t[9];  // Q = P^1024 by ten squarings
  for (int k = 0; k < 9; ++k) q[k] = c.chunk_pow[k];
  for (int j = 0; j < 10; ++j) {
    mat3_mul(q, q, t);
    for (int k = 0; k < 9; ++k) q[k] = t[k];
  }
  double cur[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < nb; ++b) {
    const long long slot = ((long long)ui * max_blocks + b) * 3;
    for (int r = 0; r < 3; ++r) block_cin[slot + r] = cur[r];
    const double a[3] = {block_agg[slot], block_agg[slot + 1], block_agg[slot + 2]};
    mat3_apply(q, cur, a, cur);
  }
}

// Full-rate outputs (passes 0, 2, 3) of one wave's chunks: each lane's
// kIirAhead outputs go to LDS, then the wave stores them kIirAhead lanes
// per instruction, each lane group covering one lane's kIirAhead
// consecutive samples (whole 64-byte segments rather than one element of
// 64 scattered lines per store).
struct IirStage {
  double v[64 * (kIirAhead + 1)];  // lane l's output q at l (kIirAhead + 1) + q
  long long base[64];              // lane l's first output index
  int len[64];                     // lane l's outputs (0 for a lane past the end)
};

template <int PASS>
__global__ __launch_bounds__(kThreads) void k_iir_replay(
    Config c, const Utt* utt, const long long* c_off, long long total,
    const double* x, double* w, const double* carries, double* y) {
  long long cg = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (PASS != 1) {
    __shared__ IirStage stage[kThreads / 64];
    IirStage& st = stage[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const bool live = cg < total;
    const Utt u = utt[live ? find_segment(c_off, c.count, cg) : 0];
    const long long start = live ? (cg - u.c_off) * kChunk : 0;
    const long long end = live ? min(start + (long long)kChunk, u.m) : 0;
    double* out = PASS == 3 ? y : w;
    st.base[lane] = u.m_off + start;
    st.len[lane] = (int)(end - start);
    double w0 = 0.0, w1 = 0.0, w2 = 0.0;
    if (live) {
      w0 = carries[cg * 3 + 0];
      w1 = carries[cg * 3 + 1];
      w2 = carries[cg * 3 + 2];
    }
    // Every lane takes kChunk / kIirAhead rounds (a short chunk's extra
    // samples are computed and not stored), so that the wave stores together.
    for (int r0 = 0; r0 < kChunk; r0 += kIirAhead) {
      double in[kIirAhead];
#pragma unroll
      for (int q = 0; q < kIirAhead; ++q)
        in[q] = start + r0 + q < end ? iir_input<PASS>(c, u, x, w, start + r0 + q) : 0.0;
#pragma unroll
      for (int q = 0; q < kIirAhead; ++q) {
        double wt = in[q] + c.a0 * w0 + c.a1 * w1 + c.a2 * w2;
        st.v[lane * (kIirAhead + 1) + q] = c.b0 * wt + c.b1 * w0 + c.b1 * w1 + c.b0 * w2;
        w2 = w1;
        w1 = w0;
        w0 = wt;
      }
      wave_sync();
#pragma unroll
      for (int g = 0; g < kIirAhead; ++g) {  // lanes g 64 / kIirAhead + lane / kIirAhead
        const int l = g * (64 / kIirAhead) + lane / kIirAhead, q = lane % kIirAhead;
        if (r0 + q < st.len[l]) out[st.base[l] + r0 + q] = st.v[l * (kIirAhead + 1) + q];
      }
      wave_sync();  // the values are stored before the next round overwrites them
    }
    return;
  }
  if (cg >= total) return;
  const Utt u = utt[find_segment(c_off, c.count, cg)];
  long long start = (cg - u.c_off) * kChunk;
  long long end = min(start + (long long)kChunk, u.m);
  double w0 = carries[cg * 3 + 0], w1 = carries[cg * 3 + 1], w2 = carries[cg * 3 + 2];
  // Pass 1 keeps output sample i when d = out_base - i is a nonnegative
  // multiple of r: d falls by one per sample, so its residue and quotient are
  // stepped instead of divided.
  long long d = select_base(c, u) - start;
  int rem = 0;
  long long j = -1;
  if (PASS == 1 && d >= 0) {
    rem = (int)(d % c.ratio);
    j = d / c.ratio;
  }
  for (long long i0 = start; i0 < end; i0 += kIirAhead) {
    double in[kIirAhead];
#pragma unroll
    for (int q = 0; q < kIirAhead; ++q)
      in[q] = i0 + q < end ? iir_input<PASS>(c, u, x, w, i0 + q) : 0.0;
#pragma unroll
    for (int q = 0; q < kIirAhead; ++q) {
      const long long i = i0 + q;
      if (i >= end) break;
      double wt = in[q] + c.a0 * w0 + c.a1 * w1 + c.a2 * w2;
      double out = c.b0 * wt + c.b1 * w0 + c.b1 * w1 + c.b0 * w2;
      w2 = w1;
      w1 = w0;
      w0 = wt;
      if (PASS == 0 || PASS == 2) {
        w[u.m_off + i] = out;
      } else if (PASS == 3) {  // y is the full-length output buffer here
        y[u.m_off + i] = out;
      } else {
        if (rem == 0 && j >= 0 && j < u.n) y[u.y_off + j] = out;
        if (rem == 0) {
          rem = c.ratio - 1;
          --j;
        } else {
          --rem;
        }
      }
    }
  }
}

// x_bar = P^T R z for z = pass 3's output: interior samples take one value,
// the two end samples collect their padding copies (lag + 10 each).
__global__ __launch_bounds__(kThreads) void k_iir_fold(Config c, const Utt* utt,
                                                      const long long* x_off, long long total,
                                                      const double* z, double* gx) {
  const long long g = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (g >= total) return;
  const Utt u = utt[find_segment(x_off, c.count, g)];
  const long long k = g - u.x_off;
  const long long shift = 9 + c.lag;  // padded index of x[0]
  auto v = [&](long long t) { return z[u.m_off + (u.m - 1 - t)]; };
  double s = 0.0;
  if (u.length == 1) {
    for (long long t = 0; t < u.m; ++t) s += v(t);
  } else if (k == 0) {
    for (long long t = 0; t <= shift; ++t) s += v(t);
  } else if (k == u.length - 1) {
    for (long long t = k + shift; t < u.m; ++t) s += v(t);
  } else {
    s = v(k + shift);
  }
  gx[g] = s;
}

__global__ __launch_bounds__(kThreads) void k_copy_input(
    Config c, const Utt* utt, const long long* y_off, long long total,
    const double* x, double* y) {
  long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= total) return;
  const Utt u = utt[find_segment(y_off, c.count, i)];
  y[i] = x[u.x_off + (i - u.y_off)];
}

// ---------------------------------------------------------------------------
// Mean removal (harvest.cpp:81-85) with a deterministic two-level sum.
__global__ __launch_bounds__(kThreads) void k_partial_sum(
    Config c, const Utt* utt, const long long* p_off, const double* y,
    double* partial) {
  __shared__ double lds[kThreads];
  long long pb = blockIdx.x;
  const Utt u = utt[find_segment(p_off, c.count, pb)];
  long long start = (pb - u.p_off) * kMeanBlock;
  long long end = min(start + (long long)kMeanBlock, (long long)u.n);
  double s = 0.0;
  for (long long i = start + threadIdx.x; i < end; i += kThreads) s += y[u.y_off + i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) partial[pb] = s;
}

__global__ __launch_bounds__(kThreads) void k_mean_final(Config c, const Utt* utt,
                                                        const double* partial,
                                                        double* mean) {
  __shared__ double lds[kThreads];
  const Utt u = utt[blockIdx.x];
  double s = 0.0;
  for (int k = threadIdx.x; k < u.pblocks; k += kThreads) s += partial[u.p_off + k];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) mean[blockIdx.x] = s / u.n;
}

__global__ __launch_bounds__(kThreads) void k_subtract(Config c, const long long* y_off,
                                                      long long total, double* y,
                                                      const double* mean) {
  long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= total) return;
  y[i] -= mean[find_segment(y_off, c.count, i)];
}

// ---------------------------------------------------------------------------
// WORLD overwrites bins n/2-1 and n/2 of every band's product spectrum with
// S[n/2] * P[n/2-1] (harvest.cpp:122-133).  The filtered signal is therefore
// the linear convolution plus (-1)^k [2 Re(D1 e^{-2 pi i k / n}) + D2] / n
// with D1 = (S[n/2] - 1) P[n/2-1] and D2 = S[n/2] Re P[n/2-1] - P[n/2].
__global__ __launch_bounds__(kThreads) void k_quirk_partial(
    Config c, const Utt* utt, const long long* p_off, const double* y,
    double* partial3) {
  __shared__ double lds[kThreads];
  long long pb = blockIdx.x;
  const Utt u = utt[find_segment(p_off, c.count, pb)];
  long long start = (pb - u.p_off) * kMeanBlock;
  long long end = min(start + (long long)kMeanBlock, (long long)u.n);
  double sh = 0.0, sr = 0.0, si = 0.0;
  for (long long i = start + threadIdx.x; i < end; i += kThreads) {
    double v = y[u.y_off + i];
    if (i & 1) v = -v;
    double sn, cs;
    sincospi(2.0 * (double)i / u.fft, &sn, &cs);
    sh += v;
    sr = fma(v, cs, sr);
    si = fma(v, sn, si);
  }
  sh = block_sum(sh, lds);
  sr = block_sum(sr, lds);
  si = block_sum(si, lds);
  if (threadIdx.x == 0) {
    partial3[pb * 3 + 0] = sh;
    partial3[pb * 3 + 1] = sr;
    partial3[pb * 3 + 2] = si;
  }
}

__global__ __launch_bounds__(kThreads) void k_quirk_final(Config c, const Utt* utt,
                                                         const double* partial3,
                                                         double* sums3) {
  __shared__ double lds[kThreads];
  const Utt u = utt[blockIdx.x];
  double a[3] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < u.pblocks; k += kThreads)
    for (int q = 0; q < 3; ++q) a[q] += partial3[(u.p_off + k) * 3 + q];
  for (int q = 0; q < 3; ++q) {
    double s = block_sum(a[q], lds);
    if (threadIdx.x == 0) sums3[blockIdx.x * 3 + q] = s;
  }
}

// One wavefront per (utterance, band): filter spectrum at the two bins, direct
// sums over the 2h+1 taps (coefficients are stored reversed: cr[mm] = c[2h-mm]),
// lane-strided and reduced by a fixed butterfly.
__global__ __launch_bounds__(kThreads) void k_quirk_bands(
    Config c, const Utt* utt, const double* sums3, const double* coef,
    const int* coef_off, const int* half, Quirk* quirk) {
  const long long id = (long long)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (id >= (long long)c.count * c.bands) return;
  int ui = (int)(id / c.bands), b = (int)(id % c.bands);
  const Utt u = utt[ui];
  const double* cr = coef + coef_off[b];
  int h = half[b];
  double hh = 0.0, hr = 0.0, hi = 0.0;
  for (int m = lane; m <= 2 * h; m += 64) {
    double v = cr[2 * h - m];
    if (m & 1) v = -v;
    double sn, cs;
    sincospi(2.0 * (double)m / u.fft, &sn, &cs);
    hh += v;
    hr = fma(v, cs, hr);
    hi = fma(v, sn, hi);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    hh += __shfl_xor(hh, d, 64);
    hr += __shfl_xor(hr, d, 64);
    hi += __shfl_xor(hi, d, 64);
  }
  if (lane != 0) return;
  double shalf = sums3[ui * 3 + 0], s1r = sums3[ui * 3 + 1], s1i = sums3[ui * 3 + 2];
  double p1r = s1r * hr - s1i * hi;
  double p1i = s1r * hi + s1i * hr;
  double p2 = shalf * hh;
  Quirk q;
  q.d1re = shalf * p1r - p1r;
  q.d1im = shalf * p1i - p1i;
  q.d2 = shalf * p1r - p2;
  q.pad = 0.0;
  quirk[id] = q;
}

// e^{-2 pi i q / n} is read as (cos, sin)[q] of 2 pi q / n; sincospi takes the
// exact rational argument 2q/n (n a power of two).
__global__ __launch_bounds__(kThreads) void k_twiddle(double* tw, int n) {
  long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= n) return;
  double sn, cs;
  sincospi(2.0 * (double)q / n, &sn, &cs);
  tw[2 * q] = cs;
  tw[2 * q + 1] = sn;
}

// Window rotations: (cos, sin) of 2 pi m / (2h + 1) at tab[2 (65 h + m)] for
// h <= hmax and m <= 64 (sincospi of the exact rational argument 2m / len).
__global__ __launch_bounds__(kThreads) void k_window_table(double* tab, int hmax) {
  const long long id = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (id >= 65LL * (hmax + 1)) return;
  const int h = (int)(id / 65), m = (int)(id % 65);
  double sn, cs;
  sincospi(2.0 * m / (2 * h + 1), &sn, &cs);
  tab[2 * id] = cs;
  tab[2 * id + 1] = sn;
}

// Refinement windows of frames whose Blackman phase offset is the integer
// -(h + 1) (window_pair): entry h^2 + m (m <= 2h) holds (w[m], d[m]).
__global__ __launch_bounds__(kThreads) void k_window_pairs(double2* tab, int hmax) {
  const long long id = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (id >= (long long)(hmax + 1) * (hmax + 1)) return;
  int h = (int)sqrt((double)id);
  while ((long long)h * h > id) --h;
  while ((long long)(h + 1) * (h + 1) <= id) ++h;
  const int m = (int)(id - (long long)h * h);
  tab[id] = h > 0 ? window_pair(m, h) : make_double2(0.0, 0.0);
}

// Two-level phase table of size P = 2^bits_total: fine entries
// (cos, sin)(2 pi a / P), a < 2^fine_bits, then coarse entries
// (cos, sin)(2 pi b 2^fine_bits / P), b < P / 2^fine_bits.
__global__ __launch_bounds__(kThreads) void k_phase_table(double* tab, long long P,
                                                          int fine_bits) {
  const long long id = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long F = 1LL << fine_bits;
  if (id >= F + P / F) return;
  const long long q = id < F ? id : (id - F) * F;
  double sn, cs;
  sincospi(2.0 * (double)q / (double)P, &sn, &cs);
  tab[2 * id] = cs;
  tab[2 * id + 1] = sn;
}

__global__ __launch_bounds__(kThreads) void k_fill_int(int* p, long long n, int v) {
  long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(64) void k_store_words(StoreWords words, int n,
                                                    unsigned long long* dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = words.w[i];
}

// ---------------------------------------------------------------------------
// Filterbank + crossings.  One workgroup = one (tile, band): kTile outputs
// from sample kTileStride * tile, of which the first kTileStride own their
// crossings (the last two serve only the derivative crossings before them).
// f[i] = sum_{m=0}^{2h} c[m] y[i + h + 1 - m] + Q(i + h + 1) / n reproduces
// WORLD's filtered_signal[i] / n (harvest.cpp:115-142; c2r scaling is a
// power of two and cancels in every crossing).  Each thread produces 8
// consecutive outputs with a register-blocked FIR from LDS (a rolling
// 16-sample window: 8 new samples per 64 FMAs), then the four crossing types
// of harvest.cpp:206-238 are detected per sample pair, their sub-sample edges
// computed with ZeroCrossingEngine's formula (harvest.cpp:184-186), compacted
// in order and written to the tile's slab.

// LDS position of sample s: one pad double after every 8, so that lanes 8
// samples apart (the FIR's and crossing stage's access pattern) are 9 doubles
// apart and hit distinct banks.
__device__ __forceinline__ int pad8(int s) { return s + (s >> 3); }

__host__ __device__ inline size_t pad8_size(int n) { return (size_t)n + (size_t)(n >> 3) + 1; }

// The two-bin correction, crossing detection and slab write of one tile of
// one band.  acc[r] is this thread's FIR output j0 + r (j0 = 8 threadIdx.x);
// f is the workgroup's LDS array of pad8_size(kTile) doubles, free when
// called.
__device__ void tile_crossings(const Config& c, const Utt& u, const Quirk& qk, long long g,
                               int bg, long long total_tiles, long long i0, int h,
                               const double (&acc)[8], double* f, const double2* phase,
                               unsigned long long* wave_sums, double* slab, int* counts) {
  const int j0 = threadIdx.x * 8;
  const int N = u.n;
  // Two-bin correction (-1)^k [2 Re(D1 e^{-2 pi i k / n}) + D2] / n at
  // k = i0 + j + h + 1: the twiddle of the thread's first output from the
  // two-level phase table (fine entries e^{2 pi i a / P}, a < 2^phase_bits,
  // then coarse entries e^{2 pi i b 2^phase_bits / P}; P = phase_n is a power
  // of two no smaller than n), then rotation by 2 pi / n per output.
  const double inv_n = 1.0 / u.fft;
  const long long nmask = u.fft - 1;
  const long long k0 = i0 + j0 + h + 1;
  const long long fine = (1LL << c.phase_bits) - 1;
  auto phase_at = [&](long long k, double& cs, double& sn) {  // (cos, sin)(2 pi k / n)
    const long long kk = (k & nmask) * (c.phase_n / u.fft);
    const double2 a = phase[kk & fine], b = phase[fine + 1 + (kk >> c.phase_bits)];
    cs = b.x * a.x - b.y * a.y;
    sn = b.x * a.y + b.y * a.x;
  };
  double tsn, tcs, rsn, rcs;
  phase_at(k0, tcs, tsn);
  phase_at(1, rcs, rsn);
  auto correction_at = [&](long long k, double cs, double sn) -> double {
    const double q = 2.0 * (qk.d1re * cs + qk.d1im * sn) + qk.d2;
    return (k & 1) ? -q : q;
  };
  // v[0..7]: this thread's outputs j0 + r (kept in registers); v[8], v[9]:
  // the next thread's first two, read back after the barrier.
  double v[10];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    v[r] = acc[r] + correction_at(k0 + r, tcs, tsn) * inv_n;
    f[pad8(j0 + r)] = v[r];
    const double nc = fma(tcs, rcs, -(tsn * rsn)), ns = fma(tsn, rcs, tcs * rsn);
    tcs = nc;
    tsn = ns;
  }
  __syncthreads();

  // Crossing flags: bit (type * 8 + r).  Types: 0 negative-going, 1
  // positive-going, 2 peak, 3 dip; WORLD's order (harvest.cpp:218-237).
  auto F = [&](int j) -> double { return f[pad8(j)]; };
  // The last thread owns j < kTileStride only, which needs no v[8], v[9].
  const bool next = threadIdx.x + 1 < kThreads;
  v[8] = next ? F(j0 + 8) : 0.0;
  v[9] = next ? F(j0 + 9) : 0.0;
  // Every sample pair and triple of the thread lies inside the utterance.
  const bool inside = i0 + j0 + 7 <= N - 3;
  unsigned flags = 0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int j = j0 + r;
    const long long i = i0 + j;
    if (j >= kTileStride) break;  // owned by the next tile
    if (inside || i <= N - 2) {
      const double a = v[r], bb = v[r + 1];
      if (0.0 < a && bb <= 0.0) flags |= 1u << (0 * 8 + r);
      if (a < 0.0 && bb >= 0.0) flags |= 1u << (1 * 8 + r);
    }
    if (inside || i <= N - 3) {
      const double d0 = v[r + 1] - v[r], d1 = v[r + 2] - v[r + 1];
      if (0.0 < d0 && d1 <= 0.0) flags |= 1u << (2 * 8 + r);
      if (d0 < 0.0 && d1 >= 0.0) flags |= 1u << (3 * 8 + r);
    }
  }
  unsigned long long packed = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t)
    packed |= (unsigned long long)__popc((flags >> (t * 8)) & 0xffu) << (16 * t);
  unsigned long long total;
  // Per type at most kSlabCap - 1 crossings per tile (same-type crossings
  // are two samples apart or more): the 16-bit fields never overflow.
  unsigned long long excl = block_scan_fields(packed, wave_sums, total);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    unsigned bits = (flags >> (t * 8)) & 0xffu;
    int pos = (int)((excl >> (16 * t)) & 0xffffu);
    double* out = slab + ((long long)(bg * 4 + t) * total_tiles + g) * kSlabCap;
    while (bits) {
      const int r = __ffs(bits) - 1;
      bits &= bits - 1;
      const int j = j0 + r;
      const double e = (double)(i0 + j + 1);
      double fe;
      if (t < 2) {
        fe = e - F(j) / (F(j + 1) - F(j));
      } else {
        const double d0 = F(j + 1) - F(j), d1 = F(j + 2) - F(j + 1);
        fe = e - d0 / (d1 - d0);
      }
      out[pos++] = fe;
    }
  }
  if (threadIdx.x < 4)
    counts[(long long)(bg * 4 + threadIdx.x) * total_tiles + g] =
        (int)((total >> (16 * threadIdx.x)) & 0xffffu);
}

template <bool kLds>
__global__ __launch_bounds__(kThreads) void k_filter_events(
    Config c, const Utt* utt, const long long* t_off, long long total_tiles,
    int band0, int nbands, int slot0, const double* y, const double* coef, const int* coef_off,
    const int* half, const Quirk* quirk, const double2* phase, double* slab, int* counts) {
  extern __shared__ double smem[];
  __shared__ unsigned long long wave_sums[kThreads / 64];
  // Logical block = tile * nbands + band: the bands of a tile run back to back
  // on one XCD, whose L2 then serves the tile's samples to all of them.
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const long long g = lb / nbands;
  const int bi = (int)(lb % nbands);
  const int bg = slot0 + bi;  // the band's slab slot within its group
  const int b = band0 + bi;
  const int ui = find_segment(t_off, c.count, g);
  const Utt u = utt[ui];
  const long long i0 = (g - u.t_off) * kTileStride;
  const int N = u.n;
  const int h = half[b];
  const int taps8 = (2 * h + 1 + 7) & ~7;
  const double* cr = coef + coef_off[b];
  const double* yu = y + u.y_off;
  const long long base = i0 - h + 1;
  const int span = kTile + taps8 + 10;
  if (kLds) {
    for (int s = threadIdx.x; s < span; s += kThreads) {
      long long idx = base + s;
      smem[pad8(s)] = (idx >= 0 && idx < N) ? yu[idx] : 0.0;
    }
    __syncthreads();
  }
  auto Y = [&](int s) -> double {
    if (kLds) return smem[pad8(s)];
    long long idx = base + s;
    return (idx >= 0 && idx < N) ? yu[idx] : 0.0;
  };

  // acc[r] = sum over taps mm in increasing order of cr[mm] y[j0 + r + mm].
  const int j0 = threadIdx.x * 8;
  double acc[8], va[8], vb[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    acc[r] = 0.0;
    va[r] = Y(j0 + r);
  }
  for (int mm = 0; mm < taps8; mm += 16) {
#pragma unroll
    for (int q = 0; q < 8; ++q) vb[q] = Y(j0 + mm + 8 + q);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double cc = cr[mm + k];
#pragma unroll
      for (int r = 0; r < 8; ++r) acc[r] = fma(cc, r + k < 8 ? va[r + k] : vb[r + k - 8], acc[r]);
    }
    if (mm + 8 >= taps8) break;
#pragma unroll
    for (int q = 0; q < 8; ++q) va[q] = Y(j0 + mm + 16 + q);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double cc = cr[mm + 8 + k];
#pragma unroll
      for (int r = 0; r < 8; ++r) acc[r] = fma(cc, r + k < 8 ? vb[r + k] : va[r + k - 8], acc[r]);
    }
  }
  const Quirk qk = quirk[(long long)ui * c.bands + b];
  if (kLds) __syncthreads();  // the staged samples are read: smem takes the outputs
  tile_crossings(c, u, qk, g, bg, total_tiles, i0, h, acc, smem, phase, wave_sums, slab, counts);
}

// ---------------------------------------------------------------------------
// Filterbank by FFT for the long bands (overlap-save).  Tile g's input block
// is x[j] = y[i0 - D + j], j < 4096 (zero outside the utterance), with D the
// longest FFT band's half length.  For band b,
//   f[i0 + j] = (x * c_b)[j + D + h_b + 1]          (linear convolution)
// with no circular wrap for j < kTile since D >= h_b - 1 and D + h_b <= 2047.
// It is the real inverse FFT of X_g C_b at size 4096, the block spectrum
// shared by every band of the tile.  Real transforms of size 4096 run as
// complex transforms of size 2048 of the even/odd samples packed as
// re + i im, with the usual split/merge twiddles e^{-+2 pi i k / 4096}.
// Results equal the direct FIR up to rounding; every tile is computed the
// same way whatever the batch.

// v[r] = sum_m v[m] e^{SIGN 2 pi i m r / 8}, in place.
template <int SIGN>
__device__ __forceinline__ void dft8(double* re, double* im) {
  constexpr double h = 0.70710678118654752440;
  double ar[8], ai[8];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    ar[m] = re[m] + re[m + 4];
    ai[m] = im[m] + im[m + 4];
    ar[m + 4] = re[m] - re[m + 4];
    ai[m + 4] = im[m] - im[m + 4];
  }
  {  // the difference half times e^{SIGN 2 pi i m / 8}, m = 1, 2, 3
    const double r1 = ar[5], i1 = ai[5];
    ar[5] = h * (r1 - SIGN * i1);
    ai[5] = h * (i1 + SIGN * r1);
    const double r2 = ar[6], i2 = ai[6];
    ar[6] = -SIGN * i2;
    ai[6] = SIGN * r2;
    const double r3 = ar[7], i3 = ai[7];
    ar[7] = h * (-r3 - SIGN * i3);
    ai[7] = h * (-i3 + SIGN * r3);
  }
#pragma unroll
  for (int base = 0; base < 8; base += 4) {  // two 4-point DFTs: even and odd r
    const double b0r = ar[base] + ar[base + 2], b0i = ai[base] + ai[base + 2];
    const double b1r = ar[base] - ar[base + 2], b1i = ai[base] - ai[base + 2];
    const double b2r = ar[base + 1] + ar[base + 3], b2i = ai[base + 1] + ai[base + 3];
    const double b3r = ar[base + 1] - ar[base + 3], b3i = ai[base + 1] - ai[base + 3];
    const double t3r = -SIGN * b3i, t3i = SIGN * b3r;
    ar[base] = b0r + b2r;
    ai[base] = b0i + b2i;
    ar[base + 1] = b1r + t3r;
    ai[base + 1] = b1i + t3i;
    ar[base + 2] = b0r - b2r;
    ai[base + 2] = b0i - b2i;
    ar[base + 3] = b1r - t3r;
    ai[base + 3] = b1i - t3i;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    re[2 * q] = ar[q];
    im[2 * q] = ai[q];
    re[2 * q + 1] = ar[4 + q];
    im[2 * q + 1] = ai[4 + q];
  }
}

// v[r] = sum_m v[m] e^{SIGN 2 pi i m r / 4}, in place.
template <int SIGN>
__device__ __forceinline__ void dft4(double* re, double* im) {
  const double b0r = re[0] + re[2], b0i = im[0] + im[2];
  const double b1r = re[0] - re[2], b1i = im[0] - im[2];
  const double b2r = re[1] + re[3], b2i = im[1] + im[3];
  const double b3r = re[1] - re[3], b3i = im[1] - im[3];
  const double t3r = -SIGN * b3i, t3i = SIGN * b3r;
  re[0] = b0r + b2r;
  im[0] = b0i + b2i;
  re[1] = b1r + t3r;
  im[1] = b1i + t3i;
  re[2] = b0r - b2r;
  im[2] = b0i - b2i;
  re[3] = b1r - t3r;
  im[3] = b1i - t3i;
}

constexpr int kFftM = 2048;        // complex transform size (real size 4096)
constexpr int kSpecStride = 2050;  // complex entries per stored spectrum (k = 0..2048)

// Element i of a transform is stored at LDS index sw(i) = i + i / 16: one
// pad double after every 16, which spreads the butterflies' strided accesses
// over the banks (measured: 4 % faster filterbank than an XOR swizzle of
// bits 5-9 into 0-4).  A transform array holds kFftPad doubles.
constexpr int kFftPad = kFftM + kFftM / 16;
__device__ __forceinline__ int sw(int i) { return i + (i >> 4); }

// sw(base + S m) - sw(base) for the butterflies of fft2048: S >= 16 adds
// m S / 16 pads; S = 4 with base % 16 < 4 adds m / 4; S = 1 with
// base % 16 <= 12 and m < 4 adds none.  Constant once the passes are
// unrolled, so every access is one address plus an immediate offset.
__device__ __forceinline__ int sw_step(int S, int m) {
  return S >= 16 ? m * (S + S / 16) : (S == 4 ? 4 * m + (m >> 2) : m);
}

// Position of frequency k after fft2048: k = r1 + 8 r2 + 64 r3 + 512 r4
// (r4 < 4) lands at element r1 256 + r2 32 + r3 4 + r4.
__device__ __forceinline__ int dig_rev(int k) {
  return ((k & 7) << 8) | (((k >> 3) & 7) << 5) | (((k >> 6) & 7) << 2) | (k >> 9);
}

// This thread's twiddles W_L^{j r}, r = 1, 2, 4, of the three radix-8 passes
// of fft2048 (as e^{+i}: the inverse sign), loaded ahead of the transform so
// that their latency overlaps the caller's input stage; r = 3, 5, 6, 7 are
// their products.
struct FftTwiddles {
  double2 w[3][3];
};
__device__ __forceinline__ FftTwiddles load_fft_twiddles(const double2* tw) {
  FftTwiddles f;
  const int t = threadIdx.x;
#pragma unroll
  for (int pass = 0; pass < 3; ++pass) {
    const int L = kFftM >> (3 * pass), S = L >> 3;
    const int step = (t % S) * (2 * kFftM / L);
#pragma unroll
    for (int q = 0; q < 3; ++q) f.w[pass][q] = tw[((1 << q) * step) & (2 * kFftM - 1)];
  }
  return f;
}

// Complex product with fused multiply-adds (the FFT filterbank's results
// equal the direct FIR's up to rounding, not bit for bit).
__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}

// Unscaled in-place DIF FFT of 2048 complex values (re, im in LDS, element
// i at sw(i), kFftPad doubles each) by radix 8, 8, 8, 4, on 256 threads;
// output in dig_rev order.
// SIGN -1 forward, +1 inverse.  Twiddles from load_fft_twiddles (tables of
// (cos, sin)(2 pi t / 4096)).  Ends with a barrier.
template <int SIGN>
__device__ void fft2048(LDS double* re, LDS double* im, const FftTwiddles& tws) {
  const int t = threadIdx.x;
#pragma unroll
  for (int pass = 0; pass < 3; ++pass) {
    const int L = kFftM >> (3 * pass), S = L >> 3;  // sub-transform length, butterfly stride
    const int j = t % S;
    const int pb = sw((t / S) * L + j);  // base % 16 = j % 16 < 4 when S = 4
    double vr[8], vi[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      vr[m] = re[pb + sw_step(S, m)];
      vi[m] = im[pb + sw_step(S, m)];
    }
    dft8<SIGN>(vr, vi);
    double2 wr[8];
    wr[1] = tws.w[pass][0];
    wr[2] = tws.w[pass][1];
    wr[4] = tws.w[pass][2];
    wr[3] = cmul2(wr[1], wr[2]);
    wr[5] = cmul2(wr[1], wr[4]);
    wr[6] = cmul2(wr[2], wr[4]);
    wr[7] = cmul2(wr[3], wr[4]);
#pragma unroll
    for (int r = 1; r < 8; ++r) {
      const double2 w = wr[r];
      const double wi = SIGN * w.y;
      const double xr = fma(vr[r], w.x, -(vi[r] * wi)), xi = fma(vr[r], wi, vi[r] * w.x);
      vr[r] = xr;
      vi[r] = xi;
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      re[pb + sw_step(S, m)] = vr[m];
      im[pb + sw_step(S, m)] = vi[m];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {  // radix 4, stride 1 (no twiddles)
    const int pb = sw(4 * (t + kThreads * q));
    double vr[4], vi[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      vr[m] = re[pb + sw_step(1, m)];
      vi[m] = im[pb + sw_step(1, m)];
    }
    dft4<SIGN>(vr, vi);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      re[pb + sw_step(1, m)] = vr[m];
      im[pb + sw_step(1, m)] = vi[m];
    }
  }
  __syncthreads();
}

// Spectrum X[k], k = 0..2048, of a real block of 4096 samples: a tile of y
// (kTileInput) or a band filter c_b[t] = cr_b[2 h_b - t] zero-padded.  One
// workgroup per tile or band (FFT bands [0, gridDim.x)).
template <bool kTileInput>
__global__ __launch_bounds__(kThreads) void k_spectrum(Config c, const Utt* utt,
                                                       const long long* t_off, int D,
                                                       const double* y, const double* coef,
                                                       const int* coef_off, const int* half,
                                                       const double2* tw, double2* spec) {
  __shared__ double buf[2 * kFftPad];
  LDS double* re = (LDS double*)buf;
  LDS double* im = re + kFftPad;
  const long long id = blockIdx.x;
  const int t = threadIdx.x;
  const FftTwiddles tws = load_fft_twiddles(tw);
  if (kTileInput) {
    const Utt u = utt[find_segment(t_off, c.count, id)];
    const long long first = (id - u.t_off) * kTileStride - D;
    const double* yu = y + u.y_off;
    for (int n = t; n < kFftM; n += kThreads) {
      const long long a = first + 2 * n, b = a + 1;
      re[sw(n)] = a >= 0 && a < u.n ? yu[a] : 0.0;
      im[sw(n)] = b >= 0 && b < u.n ? yu[b] : 0.0;
    }
  } else {
    const double* cr = coef + coef_off[id];
    const int h2 = 2 * half[id];
    for (int n = t; n < kFftM; n += kThreads) {
      re[sw(n)] = 2 * n <= h2 ? cr[h2 - 2 * n] : 0.0;
      im[sw(n)] = 2 * n + 1 <= h2 ? cr[h2 - 2 * n - 1] : 0.0;
    }
  }
  __syncthreads();
  fft2048<-1>(re, im, tws);
  // X[k] = E[k] + e^{-2 pi i k / 4096} O[k] with E = (C[k] + conj C[M - k]) / 2
  // and O = -i (C[k] - conj C[M - k]) / 2.
  double2* out = spec + id * kSpecStride;
  for (int k = t; k <= kFftM; k += kThreads) {
    const int p1 = sw(dig_rev(k & (kFftM - 1))), p2 = sw(dig_rev((kFftM - k) & (kFftM - 1)));
    const double ckr = re[p1], cki = im[p1], cmr = re[p2], cmi = -im[p2];
    const double er = 0.5 * (ckr + cmr), ei = 0.5 * (cki + cmi);
    const double dr = ckr - cmr, di = cki - cmi;
    const double orr = 0.5 * di, oi = -0.5 * dr;
    const double2 w = tw[k];  // e^{-2 pi i k / 4096} = (w.x, -w.y)
    out[k] = make_double2(er + fma(w.x, orr, w.y * oi), ei + fma(w.x, oi, -(w.y * orr)));
  }
}

// One workgroup per (tile, FFT band), the band varying fastest so that the
// workgroups of a tile share its spectrum in L2: the band's 2048 outputs of
// the tile by the inverse transform of X_g C_b, then the shared crossing stage.
__global__ __launch_bounds__(kThreads) void k_filter_fft(
    Config c, const Utt* utt, const long long* t_off, long long total_tiles, int band0,
    int nbands, int slot0, int D, const double2* tile_spec, const double2* band_spec,
    const int* half, const Quirk* quirk, const double2* tw, const double2* phase, double* slab,
    int* counts) {
  __shared__ double buf[2 * kFftPad];
  __shared__ unsigned long long wave_sums[kThreads / 64];
  LDS double* re = (LDS double*)buf;
  LDS double* im = re + kFftPad;
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);  // tile * nbands + band
  const long long g = lb / nbands;
  const int bi = (int)(lb % nbands);
  const int bg = slot0 + bi;
  const int b = band0 + bi;
  const int ui = find_segment(t_off, c.count, g);
  const Utt u = utt[ui];
  const long long i0 = (g - u.t_off) * kTileStride;
  const int h = half[b];
  const int t = threadIdx.x;
  const double2* X = tile_spec + g * kSpecStride;
  const double2* C = band_spec + (long long)b * kSpecStride;
  const FftTwiddles tws = load_fft_twiddles(tw);
  // W[k] = A + i B, A = (Z[k] + conj Z[M - k]) / 2,
  // B = (Z[k] - conj Z[M - k]) / 2 e^{2 pi i k / 4096}, Z = X C; k and M - k
  // share their loads (k = 0 pairs with the Nyquist entry M, k = M / 2 with
  // itself).
  auto emit = [&](int k, double2 zk, double2 zm) {  // W[k] from Z[k] and Z[M - k]
    const double ar = 0.5 * (zk.x + zm.x), ai = 0.5 * (zk.y - zm.y);
    const double dr = 0.5 * (zk.x - zm.x), di = 0.5 * (zk.y + zm.y);
    const double2 w = tw[k];
    const double br = fma(dr, w.x, -(di * w.y)), bi = fma(dr, w.y, di * w.x);
    re[sw(k)] = ar - bi;
    im[sw(k)] = ai + br;
  };
#pragma unroll
  for (int q = 0; q < kFftM / 2 / kThreads; ++q) {
    const int k = t + kThreads * q, m = kFftM - k;  // k < M / 2
    const double2 zk = cmul2(X[k], C[k]), zm = cmul2(X[m], C[m]);
    emit(k, zk, zm);
    if (k > 0) emit(m, zm, zk);
  }
  if (t == 0) {
    const double2 zh = cmul2(X[kFftM / 2], C[kFftM / 2]);
    emit(kFftM / 2, zh, zh);
  }
  __syncthreads();
  fft2048<1>(re, im, tws);
  // z[s] = (s even ? Re : Im) w[s / 2] / 2048; output j is z[j + D + h + 1].
  // The thread's 8 outputs read w[k0 .. k0 + 4].  Within a group of 8
  // consecutive k, dig_rev adds 256 per step (the low digit moves to the top),
  // which sw turns into 272: so w[k] is at sw(dig_rev(k0)) + 272 (k - k0)
  // before the next group's first index kw, and at sw(dig_rev(kw)) +
  // 272 (k - kw) from it on (k < kFftM since D + h + 1 + 8 t + 7 < 2 kFftM).
  const int s0 = 8 * t + D + h + 1, k0 = s0 >> 1, kw = (k0 | 7) + 1;
  const int a0 = sw(dig_rev(k0)), a1 = sw(dig_rev(kw & (kFftM - 1)));
  double acc[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int s = s0 + r, k = s >> 1;
    const int p = k < kw ? a0 + 272 * (k - k0) : a1 + 272 * (k - kw);
    acc[r] = ((s & 1) ? im[p] : re[p]) * (1.0 / kFftM);
  }
  __syncthreads();  // every value is read: buf takes the outputs
  const Quirk qk = quirk[(long long)ui * c.bands + b];
  tile_crossings(c, u, qk, g, bg, total_tiles, i0, h, acc, buf, phase, wave_sums, slab, counts);
}

// Exclusive scan of per-tile crossing counts within each utterance, for every
// (band in group, type).  One workgroup per (utterance, band*4+type).
__global__ __launch_bounds__(kThreads) void k_tile_scan(Config c, const Utt* utt,
                                                       long long total_tiles,
                                                       const int* counts, int* offs,
                                                       int* totals) {
  __shared__ unsigned long long wave_sums[kThreads / 64];
  const int ui = blockIdx.x;
  const long long q = blockIdx.y;
  const Utt u = utt[ui];
  const int* cnt = counts + q * total_tiles + u.t_off;
  int* out = offs + q * total_tiles + u.t_off;
  const int T = u.tiles;
  const int K = (T + kThreads - 1) / kThreads;
  const int first = threadIdx.x * K;
  const int last = min(first + K, T);
  unsigned long long s = 0;
  for (int k = first; k < last; ++k) s += (unsigned long long)cnt[k];
  unsigned long long total;
  unsigned long long run = block_scan_u64(s, wave_sums, total);
  for (int k = first; k < last; ++k) {
    out[k] = (int)run;
    run += (unsigned long long)cnt[k];
  }
  if (threadIdx.x == 0) totals[q * c.count + ui] = (int)total;
}

// Raw candidate contour (harvest.cpp:260-293): for each 1 ms frame whose
// sample position falls in this tile, interp1 with histc's end extrapolation
// (matlabfunctions.cpp:136-176) over the interval list of each crossing type,
// average, and gate (harvest.cpp:240-254).  Frame i of a tile only needs
// crossings of ranks [R_lo - 3, R_hi + 2] (R: ranks inside the tile, same-type
// edges are more than one sample apart), so each type's window of ranks
// [R_lo - 4, R_hi + 4] is staged in LDS.
__global__ __launch_bounds__(kThreads) void k_interp(
    Config c, const Utt* utt, const long long* t_off, long long total_tiles,
    long long total_frames, int band0, const double* boundaries,
    const double* slab, const int* counts, const int* offs, const int* totals,
    double* raw) {
  // Logical block = band * total_tiles + tile: each XCD walks a contiguous
  // run of one band's tiles, whose crossing slabs (and the adjacent tiles'
  // margins) it reads.
  // One crossing type at a time: its edges of ranks [wlo, whi] are staged,
  // then the interval locations loc_k = (e_k + e_{k+1}) / 2 / fs and values
  // val_k = fs / (e_{k+1} - e_k) (harvest.cpp:222-236) are computed once, so
  // histc's search only compares; each frame keeps its running sum of the
  // four types' interp1 values in registers, in type order.
  constexpr int kWin = kSlabCap + 2 * kWindowMargin + 2;
  constexpr int kReps = kInterpFramesPerThread;
  __shared__ double edge[kWin], loc[kWin], val[kWin];
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
  const long long g = lb % total_tiles;
  const int bg = (int)(lb / total_tiles);
  const int b = band0 + bg;
  const int ui = find_segment(t_off, c.count, g);
  const Utt u = utt[ui];
  const long long tau = g - u.t_off;
  const long long scale = (long long)kTileStride * 1000LL * c.ratio;
  long long i_lo = (tau * scale + c.fs - 1) / c.fs;
  long long i_hi = (tau == u.tiles - 1) ? (long long)u.frames - 1
                                        : ((tau + 1) * scale + c.fs - 1) / c.fs - 1;
  if (i_hi > u.frames - 1) i_hi = u.frames - 1;
  if (i_lo > i_hi) return;
  double* out = raw + (long long)b * total_frames + u.f_off;

  // Every type's counts and offsets, loaded together: its total E, this
  // tile's first rank and count, and the rank ranges of the adjacent tiles
  // (where the window margins almost always fall).
  int tE[4], tlo[4], tn[4], tprev[4], tnext[4];
  const bool has_prev = g > u.t_off, has_next = g + 1 < u.t_off + u.tiles;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const long long q = bg * 4 + t;
    tE[t] = totals[q * c.count + ui];
    tlo[t] = offs[q * total_tiles + g];
    tn[t] = counts[q * total_tiles + g];
    tprev[t] = has_prev ? counts[q * total_tiles + g - 1] : 0;
    tnext[t] = has_next ? counts[q * total_tiles + g + 1] : 0;
  }
  bool valid = true;
  for (int t = 0; t < 4; ++t) valid = valid && tE[t] >= 4;
  if (!valid) {
    for (long long i = i_lo + threadIdx.x; i <= i_hi; i += kThreads) out[i] = 0.0;
    return;
  }
  const double fs = c.fs_work;
  double sum[kReps];
  for (int ty = 0; ty < 4; ++ty) {
    const long long q = bg * 4 + ty;
    const int* qoff = offs + q * total_tiles;
    const int E = tE[ty];
    const int r_lo = tlo[ty], n_in = tn[ty];
    const int r_prev = r_lo - tprev[ty], r_end = r_lo + n_in, r_next = r_end + tnext[ty];
    const int wlo = max(0, r_lo - kWindowMargin);
    const int whi = min(E - 1, r_lo + n_in - 1 + kWindowMargin);
    if (ty > 0) __syncthreads();  // the previous type's readers are done
    for (int k = wlo + threadIdx.x; k <= whi; k += kThreads) {
      long long gt;
      int local;
      if (k >= r_lo && k < r_end) {
        gt = g;
        local = k - r_lo;
      } else if (k < r_lo && k >= r_prev) {
        gt = g - 1;
        local = k - r_prev;
      } else if (k >= r_end && k < r_next) {
        gt = g + 1;
        local = k - r_end;
      } else {
        long long lo = u.t_off, hi = u.t_off + u.tiles - 1;
        while (lo < hi) {
          long long mid = (lo + hi + 1) >> 1;
          if (qoff[mid] <= k) lo = mid; else hi = mid - 1;
        }
        gt = lo;
        local = k - qoff[gt];
      }
      edge[k - wlo] = slab[(q * total_tiles + gt) * kSlabCap + local];
    }
    __syncthreads();
    for (int k = wlo + threadIdx.x; k < whi; k += kThreads) {
      const double e0 = edge[k - wlo], e1 = edge[k + 1 - wlo];
      loc[k - wlo] = (e0 + e1) / 2.0 / fs;
      val[k - wlo] = fs / (e1 - e0);
    }
    __syncthreads();
#pragma unroll
    for (int rep = 0; rep < kReps; ++rep) {
      const long long i = i_lo + threadIdx.x + (long long)rep * kThreads;
      if (i > i_hi) continue;
      const double t = (double)i / 1000.0;
      // k* = number of interval locations <= t (histc); loc_k needs edges
      // k and k + 1, so k ranges over [wlo, min(whi, E - 1)).
      int lo = wlo;
      int hi = min(whi, E - 1);
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (loc[mid - wlo] <= t) lo = mid + 1; else hi = mid;
      }
      // histc clamps to [1, E - 2]; the window bounds hold by construction
      // (ranks either side of the tile) and are enforced here as well.
      int kc = min(max(lo, 1), E - 2);
      kc = min(max(kc, wlo + 1), whi - 1);
      const double x0 = loc[kc - 1 - wlo], x1 = loc[kc - wlo];
      const double y0 = val[kc - 1 - wlo], y1 = val[kc - wlo];
      const double s = (t - x0) / (x1 - x0);
      const double v = y0 + s * (y1 - y0);
      sum[rep] = ty == 0 ? v : sum[rep] + v;
    }
  }
  const double bound = boundaries[b];
  const double upper = bound * 1.1, lower = bound * 0.9;
#pragma unroll
  for (int rep = 0; rep < kReps; ++rep) {
    const long long i = i_lo + threadIdx.x + (long long)rep * kThreads;
    if (i > i_hi) continue;
    double cand = sum[rep] / 4.0;
    if (cand > upper || cand < lower || cand > c.f0_ceil || cand < c.f0_floor) cand = 0.0;
    out[i] = cand;
  }
}

// Official candidates (harvest.cpp:348-412): per frame, runs of >= 10 voiced
// bands (bands 0 and B-1 forced unvoiced) averaged in band order.
__global__ __launch_bounds__(kThreads) void k_official(Config c, const long long* f_off,
                                                      long long total_frames,
                                                      const double* raw,
                                                      double* official, int* count) {
  const long long gf = (long long)blockIdx.x * kThreads + threadIdx.x;
  // Lanes past the end stay active for the wave-level reduction below.
  const bool valid = gf < total_frames;
  const int ui = find_segment(f_off, c.count, valid ? gf : total_frames - 1);
  double* o = official + gf * c.capacity;
  int n = 0, st = 0;
  bool prev = false;
  double sum = 0.0;
  // Eight bands are loaded ahead of their use: one thread per frame leaves few
  // waves per CU, so the loads' latency is hidden by having them in flight.
  constexpr int kAhead = 8;
  for (int b0 = 0; valid && b0 < c.bands; b0 += kAhead) {
    double vals[kAhead];
#pragma unroll
    for (int q = 0; q < kAhead; ++q)
      vals[q] = b0 + q < c.bands ? raw[(long long)(b0 + q) * total_frames + gf] : 0.0;
#pragma unroll
    for (int q = 0; q < kAhead; ++q) {
      const int b = b0 + q;
      if (b >= c.bands) break;
      const double v = vals[q];
      const bool on = b > 0 && b < c.bands - 1 && v > 0;
      if (on) {
        if (!prev) {
          st = b;
          sum = 0.0;
        }
        sum += v;
      } else if (prev) {
        const int len = b - st;
        if (len >= 10) {
          if (n < c.capacity) o[n] = sum / len;
          ++n;
        }
      }
      prev = on;
    }
  }
  n = min(n, c.capacity);
  if (valid)
    for (int k = n; k < c.capacity; ++k) o[k] = 0.0;
  // count[u] = max over the utterance's frames: one atomic per wave when the
  // wave lies in one utterance (the usual case), not one per frame on a
  // single contended address.
  const int ui0 = __shfl(ui, 0, 64);
  if (__all(ui == ui0)) {
    int m = n;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(count + ui0, m);
  } else if (valid) {
    atomicMax(count + ui, n);
  }
}

// ---------------------------------------------------------------------------
// Refinement (GetRefinedF0 harvest.cpp:589-617, GetMeanF0 :541-584, FixF0
// :507-536), one wavefront per 1 ms frame, kRefineWaves frames per
// workgroup.  Analyses whose frames start at whole samples take the
// bin-table path (refine_table.hip) instead.
//
// Slot s < 7 count of a frame holds the overlapped candidate of
// OverlapF0Candidates (harvest.cpp:417-429).  A slot's windows depend only on
// its half length h (GetMainWindow/GetDiffWindow :446-468) and a DFT bin only
// on (h, k), so the wave
//   * groups the frame's slots by h (LDS hash, then per-window slot lists)
//     and handles one window at a time: the Blackman window by rotation of
//     one exact sincospi per lane, the derivative window from the
//     neighbouring lanes' values, and the windowed samples x w, x d staged in
//     LDS, kStage samples per pass;
//   * marks the window's distinct bins k in an LDS bitmap, whose rank
//     (popcount) gives each (slot, harmonic) its result index;
//   * sums each distinct bin with 8 lanes and a fixed butterfly: lane sub runs
//     a Goertzel recurrence at frequency 8w over its samples m = sub + 8i
//     (two operations per sample and signal, no twiddle per sample) and
//     turns the final state into its partial sum with the exact table
//     values of 8w and w m_last;
//   * keeps the frame's bin results in LDS and finishes all of its slots
//     (FixF0 and the gate) in one lane-parallel pass at the end.
// The order of operations per bin is fixed, so results do not depend on the
// batch, on the staging passes or on which path served the bin.
constexpr int kWaves = 4;        // frames per workgroup (prune)
constexpr int kRefineWaves = 2;  // frames per refinement workgroup: finer blocks balance better
constexpr int kSplit = 8;        // lanes sharing one DFT bin
constexpr int kPer = 1;          // bins per lane group per pass
constexpr int kFrameBins = 128;  // bin results of one frame kept in LDS (< 256)
constexpr int kStage = 128;      // windowed samples staged per pass (multiple of 64)
constexpr int kMaxWords = 128;   // bitmap words (bins 0..8191)
// The next window's first 128 samples (one stage) are loaded ahead; more
// cost registers (6 chunks ran 8 % slower).


// One wave's part of k_refine_waves' dynamic LDS, in carve order:
//   f0     [S]          candidate per slot
//   res    [2 kFrameBins] power and IF numerator per bin result
//   xwd    [2 kStage]   x w and x d of the staged samples
//   bits   [W]          distinct bins of the current window
//   wpre   [W]          bins below each bitmap word
//   h      [S]          half window length (0: inactive slot)
//   nh     [S]          harmonics FixF0 uses (8 bits)
//   win    [S]          window of each slot (16 bits)
//   rank   [S]          position of the slot in its window's list (16 bits)
//   wslots [S]          slots grouped by window (16 bits)
//   win_h  [S]          half length of each window
//   pk     [6 S]        bin of each (grouped slot, harmonic), -1 past nh
//   hkey   [HS]         hashed half lengths (-1: empty), later slot counts
//   hval   [HS]         window of each hashed half length, later list starts
//   bink   [kFrameBins] bins whose results are in res
//   jidx   [6 S]        result index of each (grouped slot, harmonic)
//   misc   [4]          counters
#define WAVE_LDS_FIELDS(X)                                                   \
  X(double, f0, S) X(double, res, 2 * kFrameBins)                           \
  X(double, xwd, 2 * kStage) X(unsigned long long, bits, W) X(int, wpre, W)  \
  X(int, h, S) X(unsigned char, nh, S) X(short, win, S) X(short, rank, S)    \
  X(short, wslots, S) X(int, win_h, S) X(int, hkey, HS) X(int, hval, HS)     \
  X(int, bink, kFrameBins) X(short, pk, 6 * S) X(unsigned char, jidx, 6 * S) X(int, misc, 4)

struct WaveLds {
#define X(T, name, count) LDS T* name;
  WAVE_LDS_FIELDS(X)
#undef X
};

__host__ __device__ inline size_t wave_lds_bytes(int S, int HS, int W) {
  size_t at = 0;
#define X(T, name, count) at += (sizeof(T) * (size_t)(count) + 15) & ~size_t(15);
  WAVE_LDS_FIELDS(X)
#undef X
  return at;
}

__device__ inline WaveLds carve_wave_lds(LDS char* base, int S, int HS, int W) {
  WaveLds l;
  size_t at = 0;
#define X(T, name, count)                \
  l.name = (LDS T*)(base + at);          \
  at += (sizeof(T) * (size_t)(count) + 15) & ~size_t(15);
  WAVE_LDS_FIELDS(X)
#undef X
  return l;
}

// LDS atomics of one wave's region.
__device__ __forceinline__ int lds_cas(LDS int* p, int expected, int desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
  return expected;  // the previous value, whether or not it was replaced
}
__device__ __forceinline__ int lds_add(LDS int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_or(LDS unsigned long long* p, unsigned long long v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// v of the lane that DPP control kCtrl selects (a permutation within rows),
// as two 32-bit moves: no LDS round trip.
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, kCtrl, 0xF, 0xF, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), kCtrl, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
constexpr int kDppXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141; // lane 7 - i within each 8 lanes

// Sum of v over each aligned group of 8 lanes, in every lane of the group:
// pairs, then quads, then the two quads (the same sum in all 8 lanes, as
// IEEE addition commutes).
__device__ __forceinline__ double sum8(double v) {
  v += dpp_f64<kDppXor1>(v);
  v += dpp_f64<kDppXor2>(v);
  return v + dpp_f64<kDppHalfMirror>(v);
}

// Value of v in lane `src` (uniform), without LDS traffic.
__device__ __forceinline__ double lane_value(double v, int src) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, src);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// q / d and q % d for 0 <= q < 2^16 and 1 <= d < 2^16 from m = ceil(2^32 / d):
// floor(q m / 2^32) = floor(q / d + q e / (d 2^32)) with e = m d - 2^32 < d,
// and q e / (d 2^32) < 2^-16 <= 1 / d stays below the gap to the next
// integer, so the quotient is exact.
// d = 1 is encoded as m = 0 (2^32 does not fit).
__host__ __device__ inline unsigned small_div_magic(unsigned d) {
  // ceil of 2^32 / d in FP64: a quotient that is not an integer is at least
  // 1 / d from one, far beyond its rounding.
  return d <= 1 ? 0u : (unsigned)ceil(4294967296.0 / (double)d);
}
__device__ __forceinline__ int small_div(int q, unsigned magic) {
  return magic ? (int)__umulhi((unsigned)q, magic) : q;
}

// Exclusive prefix sum of v over the wave; returns the wave total.
__device__ __forceinline__ int wave_exclusive(int v, int lane, int* excl) {
  int incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  *excl = incl - v;
  return __builtin_amdgcn_readlane(incl, 63);
}

// cos and sin of 2 pi p / n: the exact table entry (table of T >= n entries,
// stride T / n) or sincospi when no table covers n.
__device__ __forceinline__ void twiddle_at(const double* tw, int stride, int n, int p,
                                           double& cs, double& sn) {
  if (stride > 0) {
    const double2 v = reinterpret_cast<const double2*>(tw)[(long long)p * stride];
    cs = v.x;
    sn = v.y;
  } else {
    sincospi(2.0 * (double)p / n, &sn, &cs);
  }
}

// Four waves per SIMD: at most 128 VGPRs and about 10 KB of LDS per wave at
// 71-800 Hz; 17 % faster than three.  kS, kHS, kW > 0 fix the slot count,
// hash size and bitmap words at compile time (the default analysis settings):
// every LDS offset then folds into its ds_* instruction, which frees the
// scalar registers that would hold the carved pointers.  kS = 0 takes them
// from the arguments; both instances compute the same values.  kPairs reads
// each window's (w, d) from the k_window_pairs table (c.win_pairs), otherwise
// they are generated per frame by rotation.
template <int kS, int kHS, int kW, bool kPairs>
__global__ __launch_bounds__(64 * kRefineWaves, 4) void k_refine_waves(
    Config c, const Utt* utt, const long long* f_off, long long total_frames, int hs_arg,
    int w_arg, const double* y, const double* official, const int* count, const double* twiddle,
    const double* wintab, const double2* winpair, double* rf0, double* rscore,
    unsigned long long* work) {
  extern __shared__ double lds_raw[];
  const int S = kS > 0 ? kS : c.slots;
  const int HS = kS > 0 ? kHS : hs_arg;
  const int W = kS > 0 ? kW : w_arg;
  // The wave index is uniform: readfirstlane lets the compiler keep the
  // frame's scalars (frame, utterance, counts) in scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int grp = lane / kSplit, sub = lane % kSplit;
  const WaveLds L =
      carve_wave_lds((LDS char*)(char*)lds_raw + wave * wave_lds_bytes(S, HS, W), S, HS, W);

  const long long gf = (long long)xcd_block(blockIdx.x, gridDim.x) * kRefineWaves + wave;
  if (gf >= total_frames) return;
  const int ui = find_segment(f_off, c.count, gf);
  const int cnt = count[ui];
  const int ncand = 7 * cnt;
  if (ncand == 0) return;
  const Utt u = utt[ui];
  const long long i = gf - u.f_off;
  const double fs = c.fs_work;
  // The frame time is uniform: held in scalar registers, not in a VGPR pair
  // that would be spilled across the window loop.
  const double t = lane_value((double)i / 1000.0, 0);
  const double* yu = y + u.y_off;
  auto y_at = [&](int idx) -> double {  // GetBaseIndex's clamping (harvest.cpp:434-441)
    return yu[idx < 0 ? 0 : (idx > u.n - 1 ? u.n - 1 : idx)];
  };
  double* out_f0 = rf0 + gf * S;
  double* out_sc = rscore + gf * S;
  unsigned long long wsamp = 0, bsamp = 0, nbins = 0;
  bool fault = false;

  // Slots, their half lengths, harmonic counts and distinct windows.
  for (int k = lane; k < HS; k += 64) L.hkey[k] = -1;
  if (lane == 0) L.misc[0] = 0;
  wave_sync();
  const unsigned cnt_magic = small_div_magic((unsigned)cnt);
  for (int s = lane; s < ncand; s += 64) {
    const int g = small_div(s, cnt_magic), jj = s - g * cnt;
    long long src = i;
    if (g >= 1 && g <= 3) src = i - g;
    else if (g >= 4) src = i + (g - 3);
    const double f0 = (src >= 0 && src < u.frames) ? official[(u.f_off + src) * c.capacity + jj]
                                                   : 0.0;
    int hw = f0 > 0.0 ? (int)(1.5 * fs / f0 + 1.0) : 0;
    if (hw > c.hmax) {  // cannot happen for candidates in [f0_floor, f0_ceil]
      fault = true;
      hw = 0;
    }
    L.f0[s] = f0;
    L.h[s] = hw;
    int pos = -1;
    if (hw > 0) {
      L.nh[s] = min((int)(fs / 2.0 / f0), 6);
      pos = (int)(((unsigned)hw * 2654435761u) & (unsigned)(HS - 1));
      while (true) {
        const int prev = lds_cas(L.hkey + pos, -1, hw);
        if (prev == -1) {
          const int w = lds_add(L.misc, 1);
          L.hval[pos] = w;
          L.win_h[w] = hw;
          break;
        }
        if (prev == hw) break;
        pos = (pos + 1) & (HS - 1);
      }
    }
    L.win[s] = pos;
  }
  wave_sync();
  const int nwin = __builtin_amdgcn_readfirstlane(L.misc[0]);
  for (int s = lane; s < ncand; s += 64) L.win[s] = L.win[s] >= 0 ? L.hval[L.win[s]] : -1;
  // Per-window constants, lane w holding those of window w < 64: the first
  // sample index (GetBaseIndex), e^{2 pi i a / len} of the Blackman phase
  // offset a, and the rotation by 64 samples; windows from 64 on compute them
  // in the loop.
  const double2* wrot = reinterpret_cast<const double2*>(wintab);
  int pw_basic = 0;
  double pw_ac = 1.0, pw_as = 0.0, pw_gc = 1.0, pw_gs = 0.0;
  if (lane < nwin) {
    const int hw = L.win_h[lane];
    pw_basic = mround((t + (-hw) / fs) * fs + 0.001);
  }
  if (!kPairs && lane < nwin) {
    const int hw = L.win_h[lane];
    const double a = (double)(pw_basic - 1) - t * fs;
    sincospi(2.0 * a / (2 * hw + 1), &pw_as, &pw_ac);
    const double2 g = wrot[65 * hw + 64];
    pw_gc = g.x;
    pw_gs = g.y;
  }
  auto window_basic = [&](int w) -> int {
    if (w < 64) return __builtin_amdgcn_readlane(pw_basic, w);
    return mround((t + (-__builtin_amdgcn_readfirstlane(L.win_h[w])) / fs) * fs + 0.001);
  };
  // The next window's samples lane and lane + 64 are loaded ahead into
  // registers (y0, y1), overlapping their latency with the current window's
  // work, and with them their (w, d) (t0, t1), or for the rotation form
  // e^{2 pi i lane / len} (t0).
  static_assert(kStage == 128, "one stage is two 64-sample chunks");
  double y0 = 0.0, y1 = 0.0;
  double2 t0 = make_double2(1.0, 0.0), t1 = make_double2(0.0, 0.0);
  auto prefetch = [&](int w) {
    if (w >= nwin) return;
    const int hw = __builtin_amdgcn_readfirstlane(L.win_h[w]);
    const int bs = window_basic(w), len = 2 * hw + 1;
    y0 = lane < len ? y_at(bs + lane - 1) : 0.0;
    y1 = lane + 64 < len ? y_at(bs + lane + 63) : 0.0;
    if (kPairs) {
      t0 = lane < len ? winpair[hw * hw + lane] : make_double2(0.0, 0.0);
      t1 = lane + 64 < len ? winpair[hw * hw + lane + 64] : make_double2(0.0, 0.0);
    } else {
      t0 = wrot[65 * hw + lane];
    }
  };
  prefetch(0);
  wave_sync();
  // Per-window slot lists: counts (hkey), starts (hval), lists (wslots).
  LDS int* wcnt = L.hkey;
  LDS int* wbeg = L.hval;
  for (int w = lane; w < nwin; w += 64) wcnt[w] = 0;
  wave_sync();
  for (int s = lane; s < ncand; s += 64)
    if (L.win[s] >= 0) L.rank[s] = lds_add(wcnt + L.win[s], 1);
  wave_sync();
  {
    int carry = 0;
    for (int base = 0; base < nwin; base += 64) {
      const int w = base + lane;
      int ex;
      const int tot = wave_exclusive(w < nwin ? wcnt[w] : 0, lane, &ex);
      if (w < nwin) wbeg[w] = carry + ex;
      carry += tot;
    }
    if (lane == 0) wbeg[nwin] = carry;
  }
  wave_sync();
  for (int s = lane; s < ncand; s += 64)
    if (L.win[s] >= 0) {
      const int pos = wbeg[L.win[s]] + L.rank[s];
      const int H = L.nh[s];
      const double qv = L.f0[s] * window_fft(L.h[s]) / fs;
      L.wslots[pos] = s;
#pragma unroll
      for (int hh = 0; hh < 6; ++hh) L.pk[6 * pos + hh] = (short)(hh < H ? mround(qv * (hh + 1)) : -1);
    }
  wave_sync();

  // FixF0 (harvest.cpp:507-536) and GetRefinedF0's gate (:610-614) of the
  // slot at grouped position p from its harmonics' results res[jidx].
  // The divisions of FixF0 that feed no integer decision become products
  // with reciprocals (1 / (2 pi), 1 / (h + 1), 1 / f0): within two ulps of
  // WORLD's quotients, which the spectra's own rounding far exceeds.
  constexpr double kInvTwoPi = 0.5 / kPi;
  auto finish = [&](int p) {
    const int s = L.wslots[p];
    const double f0 = L.f0[s];
    const int H = L.nh[s];
    const double inv_n = ldexp(1.0, -window_fft_log2(L.h[s]));  // 1 / n, exact
    const double inv_f0 = 1.0 / f0;
    double numer = 0.0, den = 0.0, dev = 0.0;
    for (int hh = 0; hh < H; ++hh) {
      const double inv_h = hh == 0 ? 1.0 : hh == 1 ? 0.5 : hh == 2 ? 1.0 / 3.0
                         : hh == 3 ? 0.25 : hh == 4 ? 0.2 : 1.0 / 6.0;
      const int k = L.pk[6 * p + hh];
      const int j = L.jidx[6 * p + hh];
      const double pw = L.res[2 * j], nu = L.res[2 * j + 1];
      const double ifr = pw == 0.0 ? 0.0 : (double)k * fs * inv_n + nu / pw * fs * kInvTwoPi;
      const double amp = sqrt(pw);
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
    out_f0[s] = refined;
    out_sc[s] = score;
  };
  // Finishes the slots of windows [w0, w1).
  auto flush = [&](int w0, int w1) {
    const int b = __builtin_amdgcn_readfirstlane(wbeg[w0]);
    const int e = __builtin_amdgcn_readfirstlane(wbeg[w1]);
    for (int p = b + lane; p < e; p += 64) finish(p);
    wave_sync();
  };

  const int stride = c.twiddle_n > 0 ? c.twiddle_n : 0;  // divided by n per window
  int rbase = 0, wflush = 0;
  for (int w = 0; w < nwin; ++w) {
    const int hw = __builtin_amdgcn_readfirstlane(L.win_h[w]), len = 2 * hw + 1, n = window_fft(hw);
    const int b = __builtin_amdgcn_readfirstlane(wbeg[w]);
    const int e = __builtin_amdgcn_readfirstlane(wbeg[w + 1]), pairs = 6 * (e - b);
    const int tstride = stride / n;  // 0: no table covers n
    // Blackman argument (harvest.cpp:446-456) in units of pi:
    // 2 tmp / wl with tmp = (basic + m - 1) / fs - t and wl = len / fs.
    int basic;
    double ac = 1.0, as = 0.0, gc = 1.0, gs = 0.0;
    if (kPairs) {
      basic = window_basic(w);
    } else if (w < 64) {
      basic = __builtin_amdgcn_readlane(pw_basic, w);
      ac = lane_value(pw_ac, w);
      as = lane_value(pw_as, w);
      gc = lane_value(pw_gc, w);
      gs = lane_value(pw_gs, w);
    } else {
      basic = window_basic(w);
      const double a = (double)(basic - 1) - t * fs;
      sincospi(2.0 * a / len, &as, &ac);
      const double2 g = wrot[65 * hw + 64];  // advance by 64 samples
      gc = g.x;
      gs = g.y;
    }
    // Sample m = lane: e^{2 pi i (a + m) / len} = e^{2 pi i a / len} e^{2 pi i m / len}.
    const double wc0 = kPairs ? 0.0 : fma(ac, t0.x, -(as * t0.y));
    const double ws0 = kPairs ? 0.0 : fma(as, t0.x, ac * t0.y);
    // (cos, sin) delta and sin 2 delta, delta = 2 pi / len, for the derivative window.
    double dc = 0.0, ds = 0.0, dA = 0.0, dB = 0.0;
    if (!kPairs) {
      const double2 e1 = wrot[65 * hw + 1], e2 = wrot[65 * hw + 2];
      dc = e1.x;
      ds = e1.y;
      dA = 0.5 * e1.y;
      dB = 0.16 * e2.y;
    }
    const double2* wpair = winpair + hw * hw;
    const int nstage = (len + kStage - 1) / kStage;
    bool staged = false;

    // (power, IF numerator) of bins bink[j0, j0 + nbt) into res[j0, ...).
    auto dft_bins = [&](int j0, int nbt) {
      for (int jb = 0; jb < nbt; jb += kSplit * kPer) {
        // Goertzel on this lane's samples m = sub + 8 i (frequency 8 w):
        // states (p1, p2) for x w and (q1, q2) for x d, carried across the
        // staging passes; (gc, gs) = (cos, sin) 8 w and (pc, ps) = (cos, sin)
        // of w m_last, both exact table entries.
        double p1[kPer], p2[kPer], q1[kPer], q2[kPer], gc8[kPer], gs8[kPer], pc[kPer], ps[kPer];
        bool act[kPer];
        const int mlast = sub < len ? sub + kSplit * ((len - 1 - sub) / kSplit) : 0;
#pragma unroll
        for (int qq = 0; qq < kPer; ++qq) {
          const int j = jb + grp + 8 * qq;
          act[qq] = j < nbt;
          const int k = act[qq] ? L.bink[j0 + j] : 0;
          twiddle_at(twiddle, tstride, n, (k * kSplit) & (n - 1), gc8[qq], gs8[qq]);
          twiddle_at(twiddle, tstride, n, (int)(((long long)k * mlast) & (n - 1)), pc[qq], ps[qq]);
          p1[qq] = p2[qq] = q1[qq] = q2[qq] = 0.0;
        }
        const bool fill = !staged || nstage > 1;
        double wc = wc0, ws = ws0;  // (cos, sin) phi_m of this lane's sample m
        for (int c0 = 0; c0 < len; c0 += kStage) {
          const int cend = min(c0 + kStage, len);
          if (kPairs && fill) {
            // The stage's samples m0 = c0 + lane and m1 = m0 + 64: prefetched
            // for the first stage, else both loaded before either is stored.
            const int m0 = c0 + lane, m1 = m0 + 64;
            double x0 = y0, x1 = y1;
            double2 v0 = t0, v1 = t1;
            if (c0 > 0) {
              x0 = m0 < len ? y_at(basic + m0 - 1) : 0.0;
              x1 = m1 < len ? y_at(basic + m1 - 1) : 0.0;
              v0 = m0 < len ? wpair[m0] : make_double2(0.0, 0.0);
              v1 = m1 < len ? wpair[m1] : make_double2(0.0, 0.0);
            }
            if (m0 < len) {
              L.xwd[2 * lane] = x0 * v0.x;
              L.xwd[2 * lane + 1] = x0 * v0.y;
            }
            if (m1 < len) {
              L.xwd[2 * lane + 128] = x1 * v1.x;
              L.xwd[2 * lane + 129] = x1 * v1.y;
            }
            wsamp += (unsigned long long)(cend - c0);  // uniform: kept in SGPRs
            wave_sync();
          } else if (fill) {
            // Samples s0 + lane of one 64-sample chunk; xv is y there when
            // `have` (a prefetched chunk), else it is loaded.
            auto chunk = [&](int s0, double xv, bool have) {
              const int m = s0 + lane;
              if (m < len) {
                // GetMainWindow: 0.42 + 0.5 cos phi + 0.08 cos 2 phi.
                const double w = fma(wc, fma(0.16, wc, 0.5), 0.34);
                // GetDiffWindow's central difference -(w[m+1] - w[m-1]) / 2 =
                // sin phi (sin(delta) / 2 + 0.16 sin(2 delta) cos phi), delta
                // = 2 pi / len; its ends take WORLD's one-sided values
                // d[0] = -w[1] / 2 and d[len - 1] = w[len - 2] / 2.
                double d = ws * fma(dB, wc, dA);
                if (m == 0 || m == len - 1) {
                  const double side = m == 0 ? 1.0 : -1.0;  // neighbour phi + side delta
                  const double ce = fma(wc, dc, -(side * ws * ds));
                  d = -0.5 * side * fma(ce, fma(0.16, ce, 0.5), 0.34);
                }
                if (!have) xv = y_at(basic + m - 1);
                L.xwd[2 * (m - c0)] = xv * w;
                L.xwd[2 * (m - c0) + 1] = xv * d;
              }
              // On to sample m + 64.
              const double nc = fma(wc, gc, -(ws * gs)), ns = fma(ws, gc, wc * gs);
              wc = nc;
              ws = ns;
            };
            if (c0 == 0) chunk(0, y0, true);
            for (int s0 = max(c0, 64); s0 < cend; s0 += 64) chunk(s0, 0.0, false);
            wsamp += (unsigned long long)(cend - c0);  // uniform: kept in SGPRs
            wave_sync();
          }
          // Lane sub takes staged samples sub + 8 r: `full` of them on every
          // lane, one more on lanes sub < rem (uniform trip counts).
          const int full = (cend - c0) / kSplit, rem = (cend - c0) % kSplit;
#pragma unroll
          for (int qq = 0; qq < kPer; ++qq) {
            const double c2 = 2.0 * gc8[qq];
            double a1 = p1[qq], a2 = p2[qq], b1 = q1[qq], b2 = q2[qq];
            const LDS double* px = L.xwd + 2 * sub;
            int r = 0;
#pragma unroll 1
            for (; r + 4 <= full; r += 4, px += 8 * kSplit) {  // four steps, registers alternate
              double xs[8];
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                xs[2 * v] = px[2 * kSplit * v];
                xs[2 * v + 1] = px[2 * kSplit * v + 1];
              }
              a2 = fma(c2, a1, xs[0]) - a2;
              b2 = fma(c2, b1, xs[1]) - b2;
              a1 = fma(c2, a2, xs[2]) - a1;
              b1 = fma(c2, b2, xs[3]) - b1;
              a2 = fma(c2, a1, xs[4]) - a2;
              b2 = fma(c2, b1, xs[5]) - b2;
              a1 = fma(c2, a2, xs[6]) - a1;
              b1 = fma(c2, b2, xs[7]) - b1;
            }
            auto step = [&](const LDS double* q) {
              const double a0 = fma(c2, a1, q[0]) - a2;
              const double b0 = fma(c2, b1, q[1]) - b2;
              a2 = a1;
              a1 = a0;
              b2 = b1;
              b1 = b0;
            };
            for (; r < full; ++r, px += 2 * kSplit) step(px);
            if (sub < rem) step(px);
            p1[qq] = a1;
            p2[qq] = a2;
            q1[qq] = b1;
            q2[qq] = b2;
          }
          if (nstage > 1) wave_sync();
        }
        if (fill && nstage == 1) staged = true;
#pragma unroll
        for (int qq = 0; qq < kPer; ++qq) {
          // sum_i x_i e^{-i w (m0 + 8 i)} = e^{-i w m_last} (s_{N-1} - e^{-8 i w} s_{N-2}).
          const double yr = p1[qq] - gc8[qq] * p2[qq], yi = gs8[qq] * p2[qq];
          const double zr = q1[qq] - gc8[qq] * q2[qq], zi = gs8[qq] * q2[qq];
          double mre = pc[qq] * yr + ps[qq] * yi, mim = pc[qq] * yi - ps[qq] * yr;
          double dre = pc[qq] * zr + ps[qq] * zi, dim = pc[qq] * zi - ps[qq] * zr;
          static_assert(kSplit == 8, "sum8 reduces groups of 8 lanes");
          mre = sum8(mre);
          mim = sum8(mim);
          dre = sum8(dre);
          dim = sum8(dim);
          const int j = jb + grp + 8 * qq;
          if (sub == 0 && act[qq]) {
            L.res[2 * (j0 + j)] = mre * mre + mim * mim;
            L.res[2 * (j0 + j) + 1] = mre * dim - mim * dre;
          }
        }
      }
      bsamp += (unsigned long long)nbt * len;  // uniform
      nbins += (unsigned long long)nbt;
      wave_sync();
    };

    // Distinct bins of the window (bins 0..n/2) and the result index of each
    // (slot, harmonic) pair.  Up to 64 pairs sit one per lane and are matched
    // by ballots, one round per distinct bin; more pairs mark an LDS bitmap
    // whose popcounts rank the bins.  A bin's value does not depend on its
    // index.
    if (pairs <= 64) {
      const int k = lane < pairs ? (int)L.pk[6 * b + lane] : -1;
      unsigned long long left = __ballot(k >= 0);
      int nb = 0, myj = 0, mybin = 0;
      while (left) {
        const int kl = __builtin_amdgcn_readlane(k, __builtin_ctzll(left));
        const unsigned long long same = __ballot(k == kl);
        if (k == kl) myj = nb;
        if (lane == nb) mybin = kl;
        left &= ~same;
        ++nb;
      }
      if (rbase + nb > kFrameBins) {  // results buffer full: finish what is pending
        flush(wflush, w);
        wflush = w;
        rbase = 0;
      }
      if (lane < nb) L.bink[rbase + lane] = mybin;
      if (k >= 0) L.jidx[6 * b + lane] = (unsigned char)(rbase + myj);
      wave_sync();
      dft_bins(rbase, nb);
      prefetch(w + 1);
      rbase += nb;
      continue;
    }
    const int words = (n / 2 + 64) >> 6;
    int nb = kFrameBins + 1;
    if (W > 0) {
      for (int k = lane; k < words; k += 64) L.bits[k] = 0ull;
      wave_sync();
      for (int p = lane; p < pairs; p += 64) {
        const int k = L.pk[6 * b + p];
        if (k >= 0) lds_or(L.bits + (k >> 6), 1ull << (k & 63));
      }
      wave_sync();
      if (words <= 64) {
        const int pc = lane < words ? __popcll(L.bits[lane]) : 0;
        int run = 0, ex = 0;
        for (int v = 0; v < words; ++v) {
          if (lane == v) ex = run;
          run += __builtin_amdgcn_readlane(pc, v);
        }
        if (lane < words) L.wpre[lane] = ex;
        nb = run;
      } else {
        int carry = 0;
        for (int base = 0; base < words; base += 64) {
          const int wd = base + lane;
          int ex;
          const int tot = wave_exclusive(wd < words ? __popcll(L.bits[wd]) : 0, lane, &ex);
          if (wd < words) L.wpre[wd] = carry + ex;
          carry += tot;
        }
        nb = carry;
      }
    }
    wave_sync();
    if (nb <= kFrameBins) {
      if (rbase + nb > kFrameBins) {
        flush(wflush, w);
        wflush = w;
        rbase = 0;
      }
      for (int wd = lane; wd < words; wd += 64) {
        unsigned long long bb = L.bits[wd];
        int r = rbase + L.wpre[wd];
        while (bb) {
          L.bink[r++] = wd * 64 + __ffsll((long long)bb) - 1;
          bb &= bb - 1;
        }
      }
      for (int p = lane; p < pairs; p += 64) {
        const int k = L.pk[6 * b + p];
        if (k >= 0) {
          const int wd = k >> 6;
          L.jidx[6 * b + p] = (unsigned char)(rbase + L.wpre[wd] +
                                              __popcll(L.bits[wd] & ((1ull << (k & 63)) - 1ull)));
        }
      }
      wave_sync();
      dft_bins(rbase, nb);
      prefetch(w + 1);
      rbase += nb;
    } else {
      // More distinct bins than the buffer holds (or no bitmap): finish what
      // is pending, then slot by slot.
      flush(wflush, w);
      for (int p = b; p < e; ++p) {
        const int s = __builtin_amdgcn_readfirstlane(L.wslots[p]);
        const int H = __builtin_amdgcn_readfirstlane(L.nh[s]);
        if (lane < H) {
          L.bink[lane] = L.pk[6 * p + lane];
          L.jidx[6 * p + lane] = (unsigned char)lane;
        }
        wave_sync();
        dft_bins(0, H);
        if (lane == 0) finish(p);
        wave_sync();
      }
      prefetch(w + 1);
      wflush = w + 1;
      rbase = 0;
    }
  }
  flush(wflush, nwin);
  for (int s = lane; s < ncand; s += 64)
    if (L.h[s] == 0) {
      out_f0[s] = 0.0;
      out_sc[s] = 0.0;
    }
  if (work) {
    const int spread = blockIdx.x % kWorkSpread;
    if (fault) atomicOr(&work[kStatFault * kWorkSpread + spread], 1ull);
    if (lane == 0) {
      atomicAdd(&work[kStatWindowSamples * kWorkSpread + spread], wsamp);
      atomicAdd(&work[kStatBinSamples * kWorkSpread + spread], bsamp);
      atomicAdd(&work[kStatBins * kWorkSpread + spread], nbins);
    }
  }
}

// ---------------------------------------------------------------------------
// Frozen-decision refinement of given entries and its adjoint.  Entry e is a
// 1 ms frame and the official candidate its slot was refined from; that fixes
// the half window h, the FFT size n, the harmonic count H and the bins k_h
// (GetMeanF0 harvest.cpp:541-584, FixF0 :507-536).  With those held, the
// refined F0
//   r = sum_h a_h IF_h / (sum_h a_h (h+1) + 1e-12),   a_h = sqrt(P_h),
//   IF_h = k_h fs / n + Nu_h / P_h * fs / (2 pi),
//   P_h = |M_h|^2,  Nu_h = Re M_h Im D_h - Im M_h Re D_h,
//   M_h = sum_m y_m w_m e^{-2 pi i k_h m / n},  D_h likewise with d_m,
// is a smooth function of the windowed samples.  One wavefront per entry:
// lanes own samples, the 4H sums are reduced by a fixed butterfly, r is
// written to `refined` (if given), and weight_e * dr/dy is added into gy (if
// given) at the clamped sample indices.  The gradient is accumulated with
// atomics, so its last bits depend on the order of accumulation.
__global__ __launch_bounds__(64) void k_refine_entries(
    Config c, const Utt* utt, const long long* f_off, long long n_entries,
    const long long* frame, const double* f0_init, const double* weight, const double* y,
    double* refined, double* gy) {
  const long long e = blockIdx.x;
  if (e >= n_entries) return;
  const int lane = threadIdx.x;
  const long long gf = frame[e];
  const Utt u = utt[find_segment(f_off, c.count, gf)];
  const double fs = c.fs_work;
  const double t = (double)(gf - u.f_off) / 1000.0;
  const double f0 = f0_init[e];
  const int hw = (int)(1.5 * fs / f0 + 1.0), len = 2 * hw + 1;
  int lg = 0;
  while ((2 << lg) <= len) ++lg;
  const int n = 1 << (lg + 2);
  const int basic = mround((t + (-hw) / fs) * fs + 0.001);
  const double a = (double)(basic - 1) - t * fs;
  const int H = min((int)(fs / 2.0 / f0), 6);
  int kb[6];
  for (int hh = 0; hh < 6; ++hh) kb[hh] = hh < H ? mround(f0 * n / fs * (hh + 1)) : 0;
  const double* yu = y + u.y_off;
  auto idx_of = [&](int m) {
    const int idx = basic + m - 1;
    return idx < 0 ? 0 : (idx > u.n - 1 ? u.n - 1 : idx);
  };
  // Blackman window (GetMainWindow) and its central difference (GetDiffWindow).
  auto win = [&](int m) {
    const double cs = cospi(2.0 * (a + m) / len);
    return 0.42 + 0.5 * cs + 0.08 * (2.0 * cs * cs - 1.0);
  };
  auto dwin = [&](int m) {
    if (m == 0) return -win(1) / 2.0;
    if (m == len - 1) return win(len - 2) / 2.0;
    return -(win(m + 1) - win(m - 1)) / 2.0;
  };
  double mr[6], mi[6], dr[6], di[6];
  for (int hh = 0; hh < 6; ++hh) mr[hh] = mi[hh] = dr[hh] = di[hh] = 0.0;
  for (int m = lane; m < len; m += 64) {
    const double xv = yu[idx_of(m)];
    const double xw = xv * win(m), xd = xv * dwin(m);
    for (int hh = 0; hh < H; ++hh) {
      double sn, cs;
      sincospi(2.0 * (double)(((long long)kb[hh] * m) & (n - 1)) / n, &sn, &cs);
      mr[hh] = fma(xw, cs, mr[hh]);
      mi[hh] = fma(-xw, sn, mi[hh]);
      dr[hh] = fma(xd, cs, dr[hh]);
      di[hh] = fma(-xd, sn, di[hh]);
    }
  }
  for (int hh = 0; hh < 6; ++hh)
    for (int d = 32; d > 0; d >>= 1) {
      mr[hh] += __shfl_xor(mr[hh], d, 64);
      mi[hh] += __shfl_xor(mi[hh], d, 64);
      dr[hh] += __shfl_xor(dr[hh], d, 64);
      di[hh] += __shfl_xor(di[hh], d, 64);
    }
  const double cif = fs / 2.0 / kPi;
  double pw[6], nu[6], amp[6], ifr[6], numer = 0.0, den = 0.0;
  for (int hh = 0; hh < H; ++hh) {
    pw[hh] = mr[hh] * mr[hh] + mi[hh] * mi[hh];
    nu[hh] = mr[hh] * di[hh] - mi[hh] * dr[hh];
    ifr[hh] = pw[hh] == 0.0 ? 0.0 : (double)kb[hh] * fs / n + nu[hh] / pw[hh] * fs / 2.0 / kPi;
    amp[hh] = sqrt(pw[hh]);
    numer += amp[hh] * ifr[hh];
    den += amp[hh] * (hh + 1.0);
  }
  const double dg = den + kGuard;
  const double r = numer / dg;
  if (refined && lane == 0) refined[e] = r;
  if (!gy || !weight) return;
  const double g = weight[e];
  // Reverse mode: r -> (a_h, IF_h) -> (P_h, Nu_h) -> (M_h, D_h).
  double gmr[6], gmi[6], gdr[6], gdi[6];
  for (int hh = 0; hh < H; ++hh) {
    gmr[hh] = gmi[hh] = gdr[hh] = gdi[hh] = 0.0;
    if (pw[hh] == 0.0) continue;  // IF_h = 0 and a_h = 0 there: no smooth dependence
    const double g_amp = g * (ifr[hh] - r * (hh + 1.0)) / dg;
    const double g_if = g * amp[hh] / dg;
    const double g_pw = g_amp / (2.0 * amp[hh]) - g_if * cif * nu[hh] / (pw[hh] * pw[hh]);
    const double g_nu = g_if * cif / pw[hh];
    gmr[hh] = 2.0 * mr[hh] * g_pw + di[hh] * g_nu;
    gmi[hh] = 2.0 * mi[hh] * g_pw - dr[hh] * g_nu;
    gdr[hh] = -mi[hh] * g_nu;
    gdi[hh] = mr[hh] * g_nu;
  }
  for (int m = lane; m < len; m += 64) {
    double gxw = 0.0, gxd = 0.0;
    for (int hh = 0; hh < H; ++hh) {
      double sn, cs;
      sincospi(2.0 * (double)(((long long)kb[hh] * m) & (n - 1)) / n, &sn, &cs);
      gxw += gmr[hh] * cs - gmi[hh] * sn;
      gxd += gdr[hh] * cs - gdi[hh] * sn;
    }
    const double gx = win(m) * gxw + dwin(m) * gxd;
    if (gx != 0.0) atomicAdd(gy + u.y_off + idx_of(m), gx);
  }
}

// RemoveUnreliableCandidates (harvest.cpp:652-688) and SearchF0Base
// (:693-705), one wavefront per frame.  The neighbour frames' nonzero
// pre-prune values are compacted into LDS first; a zero neighbour value only
// adds the error min(|r - 0| / r, 1) = 1, which the clamp already implies, and
// is taken into account for fidelity.  min_k fl(|r - c_k| / r) =
// fl(min_k |r - c_k| / r) because rounding is monotone, so one division per
// neighbour frame suffices.  The best F0 is that of the first slot with the
// largest score (strictly greater than 0), found by a wave reduction that
// breaks score ties by the lower slot index.
__global__ __launch_bounds__(kThreads) void k_prune_best(
    Config c, const Utt* utt, const long long* f_off, long long total_frames,
    const int* count, const double* rf0, const double* rscore, double* pf0,
    double* pscore, double* best, int* row_nnz) {
  extern __shared__ double prune_raw[];
  const int S = c.slots;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  LDS double* nxt = (LDS double*)(LDS char*)(char*)prune_raw + (size_t)wave * 2 * S;
  LDS double* prv = nxt + S;
  const long long gf = (long long)blockIdx.x * kWaves + wave;
  if (gf >= total_frames) return;
  const int ui = find_segment(f_off, c.count, gf);
  const int ncand = 7 * count[ui];
  const Utt u = utt[ui];
  const long long i = gf - u.f_off;
  const bool interior = i >= 1 && i <= u.frames - 2;
  int n1 = 0, n2 = 0;
  if (interior) {
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int base = 0; base < ncand; base += 64) {
      const int k = base + lane;
      const double a = k < ncand ? rf0[(gf + 1) * S + k] : 0.0;
      const double p = k < ncand ? rf0[(gf - 1) * S + k] : 0.0;
      const unsigned long long ma = __ballot(a != 0.0), mp = __ballot(p != 0.0);
      if (a != 0.0) nxt[n1 + __popcll(ma & below)] = a;
      if (p != 0.0) prv[n2 + __popcll(mp & below)] = p;
      n1 += __popcll(ma);
      n2 += __popcll(mp);
    }
  }
  wave_sync();
  double bs = 0.0, bf = 0.0;
  int bk = 0x7fffffff, nz = 0;
  for (int s = lane; s < ncand; s += 64) {
    double r = rf0[gf * S + s], sc = rscore[gf * S + s];
    if (r != 0.0 && interior) {
      double m1 = INFINITY, m2 = INFINITY;
      for (int k = 0; k < n1; ++k) m1 = fmin(m1, fabs(r - nxt[k]));
      for (int k = 0; k < n2; ++k) m2 = fmin(m2, fabs(r - prv[k]));
      if (n1 < ncand) m1 = fmin(m1, fabs(r));
      if (n2 < ncand) m2 = fmin(m2, fabs(r));
      double e1 = m1 / r, e2 = m2 / r;
      if (e1 > 1.0) e1 = 1.0;
      if (e2 > 1.0) e2 = 1.0;
      const double e = e1 < e2 ? e1 : e2;
      if (!(e <= 0.05)) {
        r = 0.0;
        sc = 0.0;
      }
    }
    pf0[gf * S + s] = r;
    pscore[gf * S + s] = sc;
    if (sc > bs) {
      bs = sc;
      bf = r;
      bk = s;
    }
    nz += r != 0.0;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const double obs = __shfl_xor(bs, d, 64), obf = __shfl_xor(bf, d, 64);
    const int obk = __shfl_xor(bk, d, 64);
    if (obs > bs || (obs == bs && obk < bk)) {
      bs = obs;
      bf = obf;
      bk = obk;
    }
    nz += __shfl_xor(nz, d, 64);
  }
  if (lane == 0) {
    best[gf] = bs > 0.0 ? bf : 0.0;
    row_nnz[gf] = nz;
  }
}

// Order-preserving compaction of the nonzero pruned candidates, each with the
// official candidate its slot was refined from (OverlapF0Candidates layout),
// which fixes the slot's window and bins for the frozen-decision adjoint.
__global__ __launch_bounds__(kThreads) void k_compact(Config c, const Utt* utt,
                                                     const long long* f_off,
                                                     long long total_frames,
                                                     const int* count, const double* official,
                                                     const double* pf0, const double* pscore,
                                                     const int* offsets, double* of0,
                                                     double* oscore, double* oinit) {
  // One wavefront per frame: lanes take 64 slots at a time and the nonzero
  // ones are packed in slot order by ballot ranks.
  const long long gf = (long long)blockIdx.x * (kThreads / 64) +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (gf >= total_frames) return;
  const int ui = find_segment(f_off, c.count, gf);
  const int cnt = count[ui];
  const int ncand = 7 * cnt;
  const Utt u = utt[ui];
  const long long i = gf - u.f_off;
  int pos = offsets[gf];
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned cnt_magic = small_div_magic((unsigned)max(cnt, 1));
  for (int k0 = 0; k0 < ncand; k0 += 64) {
    const int k = k0 + lane;
    const double f = k < ncand ? pf0[gf * c.slots + k] : 0.0;
    const unsigned long long nz = __ballot(f != 0.0);
    if (f != 0.0) {
      const int at = pos + __popcll(nz & below);
      of0[at] = f;
      oscore[at] = pscore[gf * c.slots + k];
      const int g = small_div(k, cnt_magic), jj = k - g * cnt;
      long long src = i;
      if (g >= 1 && g <= 3) src = i - g;
      else if (g >= 4) src = i + (g - 3);
      oinit[at] = official[(u.f_off + src) * c.capacity + jj];
    }
    pos += __popcll(nz);
  }
}

// Generic exclusive scan of ints: per-block sums, one-workgroup scan of the
// sums, then per-block local scans.  kScanBlock elements per workgroup.
constexpr int kScanBlock = kThreads * 16;

__global__ __launch_bounds__(kThreads) void k_scan_reduce(const int* in, long long n,
                                                         int* block_sums) {
  __shared__ unsigned long long wave_sums[kThreads / 64];
  long long start = (long long)blockIdx.x * kScanBlock + threadIdx.x * 16;
  unsigned long long s = 0;
  for (int k = 0; k < 16; ++k)
    if (start + k < n) s += (unsigned)in[start + k];
  unsigned long long total;
  block_scan_u64(s, wave_sums, total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = (int)total;
}

__global__ __launch_bounds__(kThreads) void k_scan_sums(int* block_sums, long long nb) {
  __shared__ unsigned long long wave_sums[kThreads / 64];
  const long long K = (nb + kThreads - 1) / kThreads;
  const long long first = threadIdx.x * K, last = min(first + K, nb);
  unsigned long long s = 0;
  for (long long k = first; k < last; ++k) s += (unsigned)block_sums[k];
  unsigned long long total;
  unsigned long long run = block_scan_u64(s, wave_sums, total);
  for (long long k = first; k < last; ++k) {
    unsigned long long v = (unsigned)block_sums[k];
    block_sums[k] = (int)run;
    run += v;
  }
}

__global__ __launch_bounds__(kThreads) void k_scan_local(const int* in, int* out,
                                                        long long n,
                                                        const int* block_offs) {
  __shared__ unsigned long long wave_sums[kThreads / 64];
  long long start = (long long)blockIdx.x * kScanBlock + threadIdx.x * 16;
  unsigned long long s = 0;
  for (int k = 0; k < 16; ++k)
    if (start + k < n) s += (unsigned)in[start + k];
  unsigned long long total;
  unsigned long long run = block_scan_u64(s, wave_sums, total) + (unsigned)block_offs[blockIdx.x];
  for (int k = 0; k < 16; ++k) {
    if (start + k < n) {
      out[start + k] = (int)run;
      run += (unsigned)in[start + k];
    }
  }
  if (start < n && start + 16 >= n) out[n] = (int)run;
}

}  // namespace

// ------------------------------------------------------------------ launchers
size_t carry_scratch_doubles(int count, long long max_chunks) {
  const long long blocks = (max_chunks + kCarryBlock - 1) / kCarryBlock;
  return (size_t)2 * 3 * (size_t)count * (size_t)std::max(blocks, 1LL);
}

// The same block kernels serve every utterance length, so an utterance's
// carries do not depend on the batch it is analysed in.
void launch_carries(hipStream_t s, const Config& c, const Utt* utt, long long max_chunks,
                    const double* states, double* carries, double* scratch) {
  const int max_blocks = (int)std::max(1LL, (max_chunks + kCarryBlock - 1) / kCarryBlock);
  double* agg = scratch;
  double* cin = scratch + (size_t)3 * c.count * max_blocks;
  const dim3 grid((unsigned)max_blocks, (unsigned)c.count);
  k_carry_blocks<false><<<grid, kThreads, 0, s>>>(c, utt, states, max_blocks, agg, cin, carries);
  k_carry_top<<<grid_for(c.count), kThreads, 0, s>>>(c, utt, max_blocks, agg, cin);
  k_carry_blocks<true><<<grid, kThreads, 0, s>>>(c, utt, states, max_blocks, agg, cin, carries);
}

void launch_decimate(hipStream_t s, const Config& c, const Utt* utt,
                     const long long* c_off, long long total_chunks, long long max_chunks,
                     const double* x, double* pass_buffer, double* states,
                     double* carries, double* carry_scratch, double* y) {
  const unsigned grid = grid_for(total_chunks);
  k_iir_chunk_state<0><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, x,
                                                  pass_buffer, states);
  launch_carries(s, c, utt, max_chunks, states, carries, carry_scratch);
  k_iir_replay<0><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, x,
                                             pass_buffer, carries, y);
  k_iir_chunk_state<1><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, x,
                                                  pass_buffer, states);
  launch_carries(s, c, utt, max_chunks, states, carries, carry_scratch);
  k_iir_replay<1><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, x,
                                             pass_buffer, carries, y);
  check_launch("decimate");
}

void launch_copy_input(hipStream_t s, const Config& c, const Utt* utt,
                       const long long* y_off, long long total, const double* x,
                       double* y) {
  k_copy_input<<<grid_for(total), kThreads, 0, s>>>(c, utt, y_off, total, x, y);
  check_launch("copy_input");
}

void launch_remove_mean(hipStream_t s, const Config& c, const Utt* utt,
                        const long long* p_off, long long total_blocks,
                        const long long* y_off, long long total_y, double* y,
                        double* partial, double* mean) {
  k_partial_sum<<<(unsigned)total_blocks, kThreads, 0, s>>>(c, utt, p_off, y, partial);
  k_mean_final<<<c.count, kThreads, 0, s>>>(c, utt, partial, mean);
  k_subtract<<<grid_for(total_y), kThreads, 0, s>>>(c, y_off, total_y, y, mean);
  check_launch("remove_mean");
}

void launch_quirk(hipStream_t s, const Config& c, const Utt* utt,
                  const long long* p_off, long long total_blocks, const double* y,
                  const double* coef, const int* coef_off, const int* half,
                  double* partial3, double* sums3, Quirk* quirk) {
  k_quirk_partial<<<(unsigned)total_blocks, kThreads, 0, s>>>(c, utt, p_off, y, partial3);
  k_quirk_final<<<c.count, kThreads, 0, s>>>(c, utt, partial3, sums3);
  k_quirk_bands<<<grid_for((long long)c.count * c.bands * 64), kThreads, 0, s>>>(
      c, utt, sums3, coef, coef_off, half, quirk);
  check_launch("quirk");
}

void launch_twiddle(hipStream_t s, double* tw, int n) {
  k_twiddle<<<grid_for(n), kThreads, 0, s>>>(tw, n);
  check_launch("twiddle");
}

void launch_phase_table(hipStream_t s, double* tab, long long P, int fine_bits) {
  const long long n = (1LL << fine_bits) + (P >> fine_bits);
  k_phase_table<<<grid_for(n), kThreads, 0, s>>>(tab, P, fine_bits);
  check_launch("phase_table");
}

void launch_window_pairs(hipStream_t s, double2* tab, int hmax) {
  k_window_pairs<<<grid_for((long long)(hmax + 1) * (hmax + 1)), kThreads, 0, s>>>(tab, hmax);
  check_launch("window_pairs");
}

void launch_window_table(hipStream_t s, double* tab, int hmax) {
  k_window_table<<<grid_for(65LL * (hmax + 1)), kThreads, 0, s>>>(tab, hmax);
  check_launch("window_table");
}

void launch_store_words(hipStream_t s, const StoreWords& words, int n, unsigned long long* dst) {
  if (n <= 0) return;
  if (n > kStoreWords) throw std::runtime_error("store_words: too many words");
  k_store_words<<<1, 64, 0, s>>>(words, n, dst);
  check_launch("store_words");
}

void launch_fill_int(hipStream_t s, int* p, long long n, int value) {
  if (n <= 0) return;
  k_fill_int<<<grid_for(n), kThreads, 0, s>>>(p, n, value);
  check_launch("fill_int");
}

void launch_filter_events(hipStream_t s, const Config& c, const Utt* utt,
                          const long long* t_off, long long total_tiles,
                          int band0, int count, int slot0, int max_half, const double* y,
                          const double* coef, const int* coef_off,
                          const int* half, const Quirk* quirk, const double* phase,
                          double* slab, int* counts) {
  if (count <= 0) return;
  const double2* ph = reinterpret_cast<const double2*>(phase);
  const int taps8 = (2 * max_half + 1 + 7) & ~7;
  const size_t lds_bytes = sizeof(double) * pad8_size(kTile + taps8 + 10);
  const unsigned grid = (unsigned)(total_tiles * count);
  if (lds_bytes <= 64 * 1024) {
    k_filter_events<true><<<grid, kThreads, lds_bytes, s>>>(
        c, utt, t_off, total_tiles, band0, count, slot0, y, coef, coef_off, half, quirk, ph, slab,
        counts);
  } else {
    k_filter_events<false><<<grid, kThreads, sizeof(double) * pad8_size(kTile), s>>>(
        c, utt, t_off, total_tiles, band0, count, slot0, y, coef, coef_off, half, quirk, ph, slab,
        counts);
  }
  check_launch("filter_events");
}

void launch_tile_spectra(hipStream_t s, const Config& c, const Utt* utt, const long long* t_off,
                         long long total_tiles, int D, const double* y, const double* tw4096,
                         double* spec) {
  k_spectrum<true><<<(unsigned)total_tiles, kThreads, 0, s>>>(
      c, utt, t_off, D, y, nullptr, nullptr, nullptr, reinterpret_cast<const double2*>(tw4096),
      reinterpret_cast<double2*>(spec));
  check_launch("tile_spectra");
}

void launch_band_spectra(hipStream_t s, const Config& c, int bands, const double* coef,
                         const int* coef_off, const int* half, const double* tw4096,
                         double* spec) {
  if (bands <= 0) return;
  k_spectrum<false><<<(unsigned)bands, kThreads, 0, s>>>(
      c, nullptr, nullptr, 0, nullptr, coef, coef_off, half,
      reinterpret_cast<const double2*>(tw4096), reinterpret_cast<double2*>(spec));
  check_launch("band_spectra");
}

void launch_filter_fft(hipStream_t s, const Config& c, const Utt* utt, const long long* t_off,
                       long long total_tiles, int band0, int count, int slot0, int D,
                       const double* tile_spec, const double* band_spec, const int* half,
                       const Quirk* quirk, const double* tw4096, const double* phase,
                       double* slab, int* counts) {
  if (count <= 0) return;
  k_filter_fft<<<(unsigned)(total_tiles * count), kThreads, 0, s>>>(
      c, utt, t_off, total_tiles, band0, count, slot0, D,
      reinterpret_cast<const double2*>(tile_spec), reinterpret_cast<const double2*>(band_spec),
      half, quirk, reinterpret_cast<const double2*>(tw4096),
      reinterpret_cast<const double2*>(phase), slab, counts);
  check_launch("filter_fft");
}

void launch_tile_scan(hipStream_t s, const Config& c, const Utt* utt,
                      long long total_tiles, int group, const int* counts,
                      int* offs, int* totals) {
  dim3 grid((unsigned)c.count, (unsigned)(group * 4));
  k_tile_scan<<<grid, kThreads, 0, s>>>(c, utt, total_tiles, counts, offs, totals);
  check_launch("tile_scan");
}

void launch_interp(hipStream_t s, const Config& c, const Utt* utt,
                   const long long* t_off, long long total_tiles,
                   long long total_frames, int band0, int group,
                   const double* boundaries, const double* slab,
                   const int* counts, const int* offs, const int* totals,
                   double* raw) {
  // Frames per tile: at most ceil(kTileStride r 1000 / fs) + 1.
  const long long per_tile = ((long long)kTileStride * 1000LL * c.ratio + c.fs - 1) / c.fs + 1;
  if (per_tile > (long long)kInterpFramesPerThread * kThreads)
    throw std::runtime_error("interp: " + std::to_string(per_tile) +
                             " frames per tile exceed the kernel's threads");
  const unsigned grid = (unsigned)(total_tiles * group);
  k_interp<<<grid, kThreads, 0, s>>>(c, utt, t_off, total_tiles, total_frames, band0,
                                     boundaries, slab, counts, offs, totals, raw);
  check_launch("interp");
}

void launch_official(hipStream_t s, const Config& c, const long long* f_off,
                     long long total_frames, const double* raw, double* official,
                     int* count) {
  k_official<<<grid_for(total_frames), kThreads, 0, s>>>(c, f_off, total_frames, raw,
                                                        official, count);
  check_launch("official");
}

void launch_refine(hipStream_t s, const Config& c, const Utt* utt,
                   const long long* f_off, long long total_frames,
                   const double* y, const double* official, const int* count,
                   const double* twiddle, const double* wintab, const double2* winpair,
                   double* rf0, double* rscore, unsigned long long* work) {
  int HS = 64;  // window hash: a power of two above the slot count (never full;
                // reused for the nwin + 1 list starts)
  while (HS < c.slots + 1) HS <<= 1;
  int W = (window_fft(c.hmax) / 2 + 64) >> 6;
  if (W > kMaxWords) W = 0;  // very low floors: every window takes the slot-by-slot path
  const size_t bytes = kRefineWaves * wave_lds_bytes(c.slots, HS, W);
  if (bytes > (size_t)c.lds_limit)
    throw std::runtime_error("refinement needs " + std::to_string(bytes) +
                             " bytes of LDS per workgroup; the device allows " +
                             std::to_string(c.lds_limit));
  const unsigned grid = (unsigned)((total_frames + kRefineWaves - 1) / kRefineWaves);
  if (c.win_pairs && !winpair) throw std::runtime_error("refine: no window table");
  const bool fixed = c.slots == 105 && HS == 128 && W == 9;  // 71-800 Hz at 7350-8000 Hz
#define REFINE(S_, HS_, W_, P_)                                                          \
  k_refine_waves<S_, HS_, W_, P_><<<grid, 64 * kRefineWaves, bytes, s>>>(               \
      c, utt, f_off, total_frames, HS, W, y, official, count, twiddle, wintab, winpair, rf0, \
      rscore, work)
  if (fixed && c.win_pairs) REFINE(105, 128, 9, true);
  else if (fixed) REFINE(105, 128, 9, false);
  else if (c.win_pairs) REFINE(0, 0, 0, true);
  else REFINE(0, 0, 0, false);
#undef REFINE
  check_launch("refine");
}

void launch_prune_best(hipStream_t s, const Config& c, const Utt* utt,
                       const long long* f_off, long long total_frames,
                       const int* count, const double* rf0,
                       const double* rscore, double* pf0, double* pscore,
                       double* best, int* row_nnz) {
  const size_t bytes = (size_t)kWaves * 2 * c.slots * sizeof(double);
  if (bytes > (size_t)c.lds_limit) throw std::runtime_error("prune: too many slots for LDS");
  const unsigned grid = (unsigned)((total_frames + kWaves - 1) / kWaves);
  k_prune_best<<<grid, kThreads, bytes, s>>>(c, utt, f_off, total_frames, count, rf0, rscore,
                                             pf0, pscore, best, row_nnz);
  check_launch("prune_best");
}

long long scan_block_count(long long n) { return (n + kScanBlock - 1) / kScanBlock; }

void launch_exclusive_scan(hipStream_t s, const int* in, int* out, long long n,
                           int* block_tmp, long long block_tmp_len) {
  const long long nb = scan_block_count(n);
  if (nb > block_tmp_len) throw std::runtime_error("scan scratch too small");
  if (n == 0) {
    k_fill_int<<<1, kThreads, 0, s>>>(out, 1, 0);
    check_launch("scan");
    return;
  }
  k_scan_reduce<<<(unsigned)nb, kThreads, 0, s>>>(in, n, block_tmp);
  k_scan_sums<<<1, kThreads, 0, s>>>(block_tmp, nb);
  k_scan_local<<<(unsigned)nb, kThreads, 0, s>>>(in, out, n, block_tmp);
  check_launch("scan");
}

void launch_compact(hipStream_t s, const Config& c, const Utt* utt,
                    const long long* f_off, long long total_frames,
                    const int* count, const double* official, const double* pf0,
                    const double* pscore, const int* offsets, double* out_f0,
                    double* out_score, double* out_init) {
  k_compact<<<grid_for(total_frames * 64), kThreads, 0, s>>>(c, utt, f_off, total_frames, count,
                                                            official, pf0, pscore, offsets, out_f0,
                                                            out_score, out_init);
  check_launch("compact");
}

void launch_decimate_transpose(hipStream_t s, const Config& c, const Utt* utt,
                               const long long* c_off, long long total_chunks,
                               long long max_chunks, const long long* x_off, long long total_x,
                               const double* gy, double* pass_buffer, double* states,
                               double* carries, double* carry_scratch, double* full,
                               double* gx) {
  const unsigned grid = grid_for(total_chunks);
  k_iir_chunk_state<2><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, gy, pass_buffer,
                                                  states);
  launch_carries(s, c, utt, max_chunks, states, carries, carry_scratch);
  k_iir_replay<2><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, gy, pass_buffer,
                                             carries, full);
  k_iir_chunk_state<3><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, gy, pass_buffer,
                                                  states);
  launch_carries(s, c, utt, max_chunks, states, carries, carry_scratch);
  k_iir_replay<3><<<grid, kThreads, 0, s>>>(c, utt, c_off, total_chunks, gy, pass_buffer,
                                             carries, full);
  k_iir_fold<<<grid_for(total_x), kThreads, 0, s>>>(c, utt, x_off, total_x, full, gx);
  check_launch("decimate_transpose");
}

void launch_refine_entries(hipStream_t s, const Config& c, const Utt* utt,
                           const long long* f_off, long long n_entries, const long long* frame,
                           const double* f0_init, const double* weight, const double* y,
                           double* refined, double* gy) {
  if (n_entries <= 0) return;
  k_refine_entries<<<(unsigned)n_entries, 64, 0, s>>>(c, utt, f_off, n_entries, frame, f0_init,
                                                      weight, y, refined, gy);
  check_launch("refine_entries");
}

}  // namespace hv
