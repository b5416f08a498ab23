// Host-side staging copy of the backend (x86-64 host code, built without the
// GPU compiler's device pass).
#pragma once

namespace hv {

// dst[i] = src[i] for i < n, written with streaming stores (the staging
// buffer is read next by the GPU's upload, not by this CPU); returns false if
// any sample is infinite or NaN, found in the same pass.  Ends with a store
// fence, so the data is globally visible before the upload is queued.
bool copy_finite(double* dst, const double* src, long long n);

// One piece of a staging copy: n samples from src to dst.
struct CopyPiece {
  double* dst;
  const double* src;
  long long n;
};

// copy_finite of every piece, split into parts of at least kCopyPart samples
// that run on a pool of host threads (created on first use, at most
// kCopyThreads, shared by all callers) and on the calling thread, which also
// takes parts itself, so the call completes even if no pool thread is free.
// Returns true if every sample is finite.  Safe to call from several threads
// at once.
constexpr long long kCopyPart = 1 << 16;
constexpr int kCopyThreads = 8;
bool copy_finite_parallel(const CopyPiece* pieces, int count);

}  // namespace hv
