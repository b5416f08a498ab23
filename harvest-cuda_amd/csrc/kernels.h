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
