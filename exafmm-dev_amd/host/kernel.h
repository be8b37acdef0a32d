// The reference's per-cell kernel API (include/kernel.h:6-14,
// kernel::P2P/P2M/M2M/M2L/L2L/L2P on one cell or cell pair) is driven by its
// CPU recursion.  On the GPU path these operators run as batched device
// kernels inside the passes (BuildTree, UpDownPass, Traversal); no per-cell
// host entry point is provided.
#ifndef kernel_h
#define kernel_h
#include "types.h"
#endif
