// gz_runtime.h -- the opaque stream handle of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

struct gz_stream_s {
  hipStream_t s;
};
