// include/build_tree_omp.h: the uniform Morton-key tree builder, selected
// explicitly (the GPU builder with build_tree_omp.h semantics).
#ifndef EXAFMM_HOST_BUILDER
#define EXAFMM_HOST_BUILDER 0
#endif
#include "build_tree.h"
