// capi_internal.hpp -- error reporting shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace petalo {

// Copies `bytes` from pageable host memory at src to the device at dst on
// stream st and returns once the copy is complete.  Large copies go through
// a process-wide ring of pinned buffers filled by several host threads while
// the previous buffer's DMA runs (about 52 GB/s against 30 GB/s for a plain
// copy from pageable memory); small ones, and those that find the ring in
// use by another thread, are plain copies.  Throws HipError on failure.
void upload_h2d(void* dst, const void* src, size_t bytes, hipStream_t st);
// The reverse: device src to pageable host memory at dst, the DMA of one
// pinned buffer overlapping the host threads' copy out of the other (which
// also spreads the first-touch page faults of a fresh destination over the
// threads).  Work on st before the call is ordered before the copy.
void download_d2h(void* dst, const void* src, size_t bytes, hipStream_t st);

// Records the message returned by petalo_last_error() on this thread.
void set_error(const std::string& msg);
// set_error(msg) and return code.
int fail(int code, const std::string& msg);

}  // namespace petalo
