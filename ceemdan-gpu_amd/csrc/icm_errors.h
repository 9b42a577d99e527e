// Device error word bits (set with atomicOr by the kernels, read by the host
// after each forward).
#pragma once

#define ICM_DEVERR_KNOTCAP 2  // a tile's knot range exceeded its LDS capacity
#define ICM_DEVERR_INDEX 4    // the sample-index array is not 0..N-1
