// Host-side staging copy of the backend (x86-64 host code, built without the
// GPU compiler's device pass).
#pragma once

namespace hv {

// dst[i] = src[i] for i < n, written with streaming stores (the staging
// buffer is read next by the GPU's upload, not by this CPU); returns false if
// any sample is infinite or NaN, found in the same pass.  Ends with a store
// fence, so the data is globally visible before the upload is queued.
bool copy_finite(double* dst, const double* src, long long n);

}  // namespace hv
