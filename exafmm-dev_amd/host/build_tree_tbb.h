// include/build_tree_tbb.h: the adaptive octree builder, selected explicitly
// (the GPU builder with build_tree_tbb.h semantics).
#ifndef EXAFMM_HOST_BUILDER
#define EXAFMM_HOST_BUILDER 1
#endif
#include "build_tree.h"
