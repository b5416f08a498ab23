// Host-side launchers of the gfx950 Harvest kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.h"

namespace hv {

struct Quirk {  // WORLD's two overwritten bins, per (utterance, band)
  double d1re, d1im, d2, pad;
};

// Decimation (exact chunked scan of WORLD's zero-phase IIR).  max_chunks is
// the largest utterance's chunk count; carry_scratch holds
// carry_scratch_doubles(count, max_chunks) doubles.
size_t carry_scratch_doubles(int count, long long max_chunks);
void launch_decimate(hipStream_t s, const Config& c, const Utt* utt,
                     const long long* c_off, long long total_chunks, long long max_chunks,
                     const double* x, double* pass_buffer, double* states,
                     double* carries, double* carry_scratch, double* y);
void launch_copy_input(hipStream_t s, const Config& c, const Utt* utt,
                       const long long* y_off, long long total, const double* x,
                       double* y);
// y -= mean(y) per utterance, deterministic fixed-order reduction.
void launch_remove_mean(hipStream_t s, const Config& c, const Utt* utt,
                        const long long* p_off, long long total_blocks,
                        const long long* y_off, long long total_y, double* y,
                        double* partial, double* mean);
// Two-bin correction terms of every (utterance, band).
void launch_quirk(hipStream_t s, const Config& c, const Utt* utt,
                  const long long* p_off, long long total_blocks, const double* y,
                  const double* coef, const int* coef_off, const int* half,
                  double* partial3, double* sums3, Quirk* quirk);
void launch_fill_int(hipStream_t s, int* p, long long n, int value);
// Refinement twiddles (cos, sin) of 2 pi q / n for q < n.
void launch_twiddle(hipStream_t s, double* tw, int n);
// (cos, sin) of 2 pi m / (2h + 1) for h <= hmax, m <= 64, 65 (h + 1) entries.
void launch_window_table(hipStream_t s, double* tab, int hmax);
// (w, d) of the refinement windows with phase offset -(h + 1) for h <= hmax,
// (hmax + 1)^2 entries, window h at [h^2, (h + 1)^2).
void launch_window_pairs(hipStream_t s, double2* tab, int hmax);

// Filterbank + crossing detection for bands [band0, band0 + group).
void launch_filter_events(hipStream_t s, const Config& c, const Utt* utt,
                          const long long* t_off, long long total_tiles,
                          int band0, int count, int slot0, int max_half, const double* y,
                          const double* coef, const int* coef_off,
                          const int* half, const Quirk* quirk, const double* phase,
                          double* slab, int* counts);
// Filterbank by FFT for bands [band0, band0 + count) (slab slots from slot0):
// tile spectra of 4096-sample blocks starting D samples before each tile,
// band filter spectra, and the per-(tile, band) inverse transforms.  Spectra
// hold kSpecDoubles doubles each; tw4096 holds (cos, sin)(2 pi t / 4096).
constexpr int kSpecDoubles = 2 * 2050;
constexpr int kFftMaxHalf = 1000;  // FFT bands need D + h <= 2047
constexpr int kFftMinHalf = 64;    // shorter bands run the direct FIR (cheaper there)
void launch_tile_spectra(hipStream_t s, const Config& c, const Utt* utt, const long long* t_off,
                         long long total_tiles, int D, const double* y, const double* tw4096,
                         double* spec);
void launch_band_spectra(hipStream_t s, const Config& c, int bands, const double* coef,
                         const int* coef_off, const int* half, const double* tw4096,
                         double* spec);
void launch_filter_fft(hipStream_t s, const Config& c, const Utt* utt, const long long* t_off,
                       long long total_tiles, int band0, int count, int slot0, int D,
                       const double* tile_spec, const double* band_spec, const int* half,
                       const Quirk* quirk, const double* tw4096, const double* phase,
                       double* slab, int* counts);
// Two-level table of e^{2 pi i k / P} (P a power of two): 2^fine_bits fine
// entries then P / 2^fine_bits coarse ones, (cos, sin) each.
void launch_phase_table(hipStream_t s, double* tab, long long P, int fine_bits);
void launch_tile_scan(hipStream_t s, const Config& c, const Utt* utt,
                      long long total_tiles, int group, const int* counts,
                      int* offs, int* totals);
void launch_interp(hipStream_t s, const Config& c, const Utt* utt,
                   const long long* t_off, long long total_tiles,
                   long long total_frames, int band0, int group,
                   const double* boundaries, const double* slab,
                   const int* counts, const int* offs, const int* totals,
                   double* raw);
void launch_official(hipStream_t s, const Config& c, const long long* f_off,
                     long long total_frames, const double* raw, double* official,
                     int* count);
void launch_refine(hipStream_t s, const Config& c, const Utt* utt,
                   const long long* f_off, long long total_frames,
                   const double* y, const double* official, const int* count,
                   const double* twiddle, const double* wintab, const double2* winpair,
                   double* rf0, double* rscore, unsigned long long* work);
// Bin-table refinement (refine_table.hip): analyses whose frames start at
// whole samples (Config::win_pairs) and whose half windows span at most
// kBtMaxHr values.
constexpr int kBtPassBins = 12;  // bins per pass over a window (at most)
constexpr int kBtWords = 4;     // 64-bit words of a frame's half-window bitmask
constexpr int kBtMaxHr = kBtWords * 64;
constexpr int kBtPairs = 32;    // (frame, h) pairs per frame with stored bin results
constexpr int kBtBlock = 8;     // window samples staged per lane at a time
constexpr int kBtRecord = 8;    // doubles per bin record
constexpr long long kBtDftGroups = 1536;  // persistent grid of k_bt_dft (at most)
// Tables of half windows h = bt_hlo + hb, hb < bt_hr:
//   info[hb]     (FFT size n, bins NB, first window row, first bin record)
//   harm[6 hb + j] (first bin, bin count (0: harmonic unused), first result
//                index) of harmonic j + 1
//   win          per h, 2h + 1 rows (w[m], d[m])
//   bins         per bin k of h, a record (2 cos w, cos w, sin w, cos w(2h),
//                sin w(2h), 0, 0, 0) of w = 2 pi k / n
struct BinTable {
  const int4* info;
  const int4* harm;
  const double2* win;
  const double* bins;
};
// Per-analysis buffers: codes and inv [frames][capacity],
// hbs[frames][bt_hb_stride(capacity)] (each candidate's hb as a byte, 0xFF
// for none), needs[frames][kBtWords], mask and cnt [bt_hr][nblk] (nblk =
// bt_blocks(frames)), off[bt_hr nblk + 1] with its scan scratch,
// list[frames kBtPairs], cinfo[3 bt_hr + 1], sched[1],
// res[frames kBtPairs bt_nbmax] (bins of half windows with more than
// kBtPassBins bins; the others' slots are finished inside k_bt_dft).
__host__ __device__ inline long long bt_hb_stride(int capacity) { return (capacity + 15) / 16 * 16; }
struct BtBufs {
  unsigned long long* codes;
  double* inv;
  unsigned char* hbs;
  unsigned long long *needs, *mask;
  int *cnt, *off, *scan_tmp;
  long long scan_len;
  int2* list;
  int *cinfo, *sched;
  double2* res;
};
long long bt_blocks(long long total_frames);
// Fills the window rows and bin records from the bins kq[hb bt_nbmax + q]
// (t.info already on the device).
void launch_bt_tables(hipStream_t s, const Config& c, const BinTable& t, const int* kq,
                      double2* win, double* bins);
void launch_refine_table(hipStream_t s, const Config& c, const Utt* utt, const long long* f_off,
                         long long total_frames, const double* y, const double* official,
                         const int* count, const BinTable& t, const BtBufs& b, double* rf0,
                         double* rscore, unsigned long long* work, hipEvent_t dft_begin = nullptr,
                         hipEvent_t dft_end = nullptr);

void launch_prune_best(hipStream_t s, const Config& c, const Utt* utt,
                       const long long* f_off, long long total_frames,
                       const int* count, const double* rf0,
                       const double* rscore, double* pf0, double* pscore,
                       double* best, int* row_nnz);
// Exclusive scan of n ints into out[0..n] (out[n] = total).
void launch_exclusive_scan(hipStream_t s, const int* in, int* out, long long n,
                           int* block_tmp, long long block_tmp_len);
long long scan_block_count(long long n);
void launch_compact(hipStream_t s, const Config& c, const Utt* utt,
                    const long long* f_off, long long total_frames,
                    const int* count, const double* official, const double* pf0,
                    const double* pscore, const int* offsets, double* out_f0,
                    double* out_score, double* out_init);
// Transpose of launch_decimate (gradient of y -> gradient of x); `full` holds
// one pass-length buffer per utterance.
void launch_decimate_transpose(hipStream_t s, const Config& c, const Utt* utt,
                               const long long* c_off, long long total_chunks,
                               long long max_chunks, const long long* x_off, long long total_x,
                               const double* gy, double* pass_buffer, double* states,
                               double* carries, double* carry_scratch, double* full,
                               double* gx);
// Frozen-decision refined F0 of (frame, initial f0) entries and/or the
// weighted adjoint added into gy.
void launch_refine_entries(hipStream_t s, const Config& c, const Utt* utt,
                           const long long* f_off, long long n_entries, const long long* frame,
                           const double* f0_init, const double* weight, const double* y,
                           double* refined, double* gy);

// Up to kStoreWords 64-bit words passed by value, stored to dst[0, n) by one
// small kernel: puts a launch plan into device memory with no host staging
// buffer and no copy from host memory.
constexpr int kStoreWords = 32;
struct StoreWords {
  unsigned long long w[kStoreWords];
};
void launch_store_words(hipStream_t s, const StoreWords& words, int n, unsigned long long* dst);

}  // namespace hv
