// Spatial Meyer shearing filters built on the device (FP64).
#pragma once

#include <hip/hip_runtime.h>

namespace shx {

// The 2^level spatial filters of one scale with support n x n, as
// shearing_filters_Myer(n, level) builds them: w_dev receives 2^level x n x n
// doubles (row-major; element [k][r][c] is MATLAB w_s(r+1, c+1, k+1)).
// Synchronises s.  Returns SHX_OK, or an error code with the message set.
int meyer_filters_device(int n, int level, double* w_dev, hipStream_t s);

}  // namespace shx
