// Host-side copies of the host entry (mdmc_io.cpp).
#pragma once

#include <stddef.h>

namespace mdmc {

// Synchronous copy between pageable host memory and device memory on the
// current device through pinned chunks (dir 0: to the device, 1: to the
// host); staged_release frees the pinned chunks.
int staged_copy(void *host, void *dev, size_t bytes, int dir);
void staged_release();

// Records msg as this thread's mdmc_last_error() and returns code.
int set_error(int code, const char *msg);

}  // namespace mdmc
