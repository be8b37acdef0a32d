// Bounds and the uniform Morton-key tree (reference: include/bound_box.h and
// include/build_tree_omp.h).  Compiled with -ffp-contract=off: keys, cell
// centres and radii must round exactly as the reference does on x86-64.
//
// Layout produced (identical to build_tree_omp.h after reverseOrder,
// :200-211): cells are level-contiguous, root first; inside a level they run
// in descending Morton order; bodies run in ascending Morton order (stable).
#include <cmath>

#include "internal.h"

namespace exafmm {

// ---------------------------------------------------------------------------
// Bounds: grid-stride min/max, one partial per block, then one block.
constexpr int kBoundsBlocks = 512;

__global__ void __launch_bounds__(256) k_bounds_partial(const double4 *__restrict__ b, int n, double *__restrict__ part) {
  __shared__ double s[6][256];
  double mn[3] = {b[0].x, b[0].y, b[0].z};
  double mx[3] = {mn[0], mn[1], mn[2]};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double4 v = b[i];
    mn[0] = v.x < mn[0] ? v.x : mn[0];
    mn[1] = v.y < mn[1] ? v.y : mn[1];
    mn[2] = v.z < mn[2] ? v.z : mn[2];
    mx[0] = v.x > mx[0] ? v.x : mx[0];
    mx[1] = v.y > mx[1] ? v.y : mx[1];
    mx[2] = v.z > mx[2] ? v.z : mx[2];
  }
  for (int d = 0; d < 3; d++) {
    s[d][threadIdx.x] = mn[d];
    s[3 + d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int d = 0; d < 3; d++) {
        const double a = s[d][threadIdx.x + w];
        s[d][threadIdx.x] = a < s[d][threadIdx.x] ? a : s[d][threadIdx.x];
        const double c = s[3 + d][threadIdx.x + w];
        s[3 + d][threadIdx.x] = c > s[3 + d][threadIdx.x] ? c : s[3 + d][threadIdx.x];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) k_bounds_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
  __shared__ double s[6][256];
  double r[6];
  for (int d = 0; d < 6; d++) r[d] = part[d];
  for (int i = threadIdx.x; i < nb; i += 256)
    for (int d = 0; d < 3; d++) {
      const double a = part[6 * i + d], c = part[6 * i + 3 + d];
      r[d] = a < r[d] ? a : r[d];
      r[3 + d] = c > r[3 + d] ? c : r[3 + d];
    }
  for (int d = 0; d < 6; d++) s[d][threadIdx.x] = r[d];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) {
        const double a = s[d][threadIdx.x + w], c = s[3 + d][threadIdx.x + w];
        s[d][threadIdx.x] = a < s[d][threadIdx.x] ? a : s[d][threadIdx.x];
        s[3 + d][threadIdx.x] = c > s[3 + d][threadIdx.x] ? c : s[3 + d][threadIdx.x];
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// The bounds stay on the device for the builder; the host copy (xmin,
// xmax) arrives behind ev_bounds and is read by sync_bounds when needed.
void run_bounds(Handle &h) {
  h.box.reserve(16);
  h.bounds_dev_stale = false;
  if (h.n == 0) {
    sync_bounds(h);
    for (int d = 0; d < 3; d++) h.xmin[d] = h.xmax[d] = 0;
    h.bounds_dev_stale = true;
    h.have_bounds = true;
    return;
  }
  const int nb = std::min(kBoundsBlocks, div_up(h.n, 256));
  h.red.reserve(6 * nb);
  hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(256), 0, h.stream, h.in4.p, h.n, h.red.p);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, h.stream, h.red.p, nb, h.box.p);
  EXA_HIP(hipMemcpyAsync(h.hpin + Handle::kPinBounds, h.box.p, 6 * sizeof(double), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipEventRecord(h.ev_bounds, h.stream));
  h.bounds_pending = true;
  h.have_bounds = true;
}

void sync_bounds(Handle &h) {
  if (!h.bounds_pending) return;
  EXA_HIP(hipEventSynchronize(h.ev_bounds));
  const double *r = reinterpret_cast<const double *>(h.hpin + Handle::kPinBounds);
  for (int d = 0; d < 3; d++) {
    h.xmin[d] = r[d];
    h.xmax[d] = r[3 + d];
  }
  h.bounds_pending = false;
}

// The device bounds set from the host values (bounds given by the caller).
static void upload_bounds(Handle &h) {
  h.box.reserve(16);
  if (!h.bounds_dev_stale) return;
  const double v[6] = {h.xmin[0], h.xmin[1], h.xmin[2], h.xmax[0], h.xmax[1], h.xmax[2]};
  EXA_HIP(hipMemcpyAsync(h.box.p, v, sizeof(v), hipMemcpyHostToDevice, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));  // v lives on this stack frame
  h.bounds_dev_stale = false;
}

// ---------------------------------------------------------------------------
// Uniform builder kernels.

// getKey (build_tree_omp.h:36-57): float cell size promoted to double for
// the division, truncation to int, bit interleave in int.
// (The key has 3 * level <= 30 bits: it is kept as a 32-bit int for the
// radix sort; k_widen_keys turns the sorted keys into the 64-bit form.)
// bounds2box (build_tree_omp.h:21-33) on the device: bb = the bounds
// (xmin, xmax); box = (x0, y0, z0, d) with d the float leaf cell size of a
// depth-L grid (:218-220), the same double operations as the reference.
__global__ void k_box(const double *__restrict__ bb, int L, double *__restrict__ box) {
  double Xc[3], R0 = 0;
  for (int d = 0; d < 3; d++) Xc[d] = (bb[3 + d] + bb[d]) / 2;
  for (int d = 0; d < 3; d++) {  // std::max(a, R0) = a < R0 ? R0 : a
    const double a = Xc[d] - bb[d], b = bb[3 + d] - Xc[d];
    R0 = a < R0 ? R0 : a;
    R0 = b < R0 ? R0 : b;
  }
  R0 *= 1.00001;
  const float d = 2 * R0 / (1 << L);
  box[0] = bb[0];
  box[1] = bb[1];
  box[2] = bb[2];
  box[3] = d;
}

__global__ void k_morton32(const double4 *__restrict__ in, int n, const double *__restrict__ box, int level,
                           int *__restrict__ key, int *__restrict__ idx) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const double4 v = in[b];
  const double x0 = box[0], y0 = box[1], z0 = box[2];
  const double dd = (double)(float)box[3];
  int ix = (int)((v.x - x0) / dd);
  int iy = (int)((v.y - y0) / dd);
  int iz = (int)((v.z - z0) / dd);
  int id = 0;
  for (int l = 0; l != level; ++l) {
    id += (ix & 1) << (3 * l);
    id += (iy & 1) << (3 * l + 1);
    id += (iz & 1) << (3 * l + 2);
    ix >>= 1;
    iy >>= 1;
    iz >>= 1;
  }
  key[b] = id;
  idx[b] = b;
}

__global__ void k_widen_keys(const int *__restrict__ k32, int n, uint64_t *__restrict__ k64) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) k64[i] = (uint64_t)(int64_t)k32[i];
}

__global__ void k_gather_bodies(const double4 *__restrict__ in, const int *__restrict__ order, int n,
                                double4 *__restrict__ out, int *__restrict__ bidx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = order[i];
  out[i] = in[s];
  bidx[i] = s;
}

// Head flags of tree level t: body i starts a level-t cell when its key
// prefix (key >> shift) differs from body i-1's.  flags has n+1 entries; the
// last is 0 so that the exclusive scan yields the total at index n.
// out[i] = in[i * stride], i < n (n small).
constexpr int kMaxUniformLevels = 10;  // the reference's int keys hold 3 x 10 bits

__global__ void k_strided_gather(const int *__restrict__ in, int64_t stride, int n, int *__restrict__ out) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = in[(int64_t)i * stride];
}

__global__ void k_level_positions(const int *__restrict__ flags, const int *__restrict__ scan, int n,
                                  int *__restrict__ pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags[i]) pos[scan[i]] = i;
}

// All levels at once (level t = blockIdx.y; arrays of stride n + 1 per level).
__global__ void k_level_heads_all(const uint64_t *__restrict__ key, int n, int L, int64_t stride,
                                  int *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
  if (i > n) return;
  const int shift = 3 * (L - t);
  int f = 0;
  if (i < n) f = (i == 0) || ((key[i] >> shift) != (key[i - 1] >> shift));
  flags[t * stride + i] = f;
}

// One exclusive scan ran over all levels: subtract each level's start.
__global__ void k_level_rebase(int *__restrict__ scan, int n, int64_t stride, const int *__restrict__ base) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
  if (i <= n) scan[t * stride + i] -= base[t];
}

struct LevelOffsets {
  int v[kMaxUniformLevels + 1];
};

__global__ void k_level_positions_all(const int *__restrict__ flags, const int *__restrict__ scan, int n,
                                      int64_t stride, LevelOffsets posoff, int *__restrict__ pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
  if (i >= n) return;
  if (flags[t * stride + i]) pos[posoff.v[t] + scan[t * stride + i]] = i;
}

struct LevelArgs {
  const int *pos;      // heads of this level, ascending
  const int *scan;     // exclusive scan of this level's flags (n+1)
  const int *scan_up;  // scan of level t-1 (or nullptr at the root)
  const int *scan_dn;  // scan of level t+1 (or nullptr at the leaves)
  int count, count_up, count_dn;
  int off, off_up, off_dn;
  int shift;           // 3 * (L - t)
  int leaf;            // 1 at the finest level
  int level;
};

// bodies2leafs / leafs2cells (build_tree_omp.h:130-198) fused with the index
// arithmetic of reverseOrder: ascending rank r at level t lands at
// off_t + count_t - 1 - r.
__device__ inline void level_cell(const LevelArgs &a, const double4 *__restrict__ body,
                                  const uint64_t *__restrict__ key, int n, int nc, const double *__restrict__ box,
                                  double4 *__restrict__ cX, double *__restrict__ cRb, int *__restrict__ cParent,
                                  int *__restrict__ cChild, int *__restrict__ cNChild, int *__restrict__ cBody,
                                  int *__restrict__ cNBody, uint64_t *__restrict__ cKey, int *__restrict__ cLevel) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.count) return;
  const int start = a.pos[r];
  const int end = (r + 1 < a.count) ? a.pos[r + 1] : n;
  const int c = a.off + a.count - 1 - r;
  const double4 v = body[start];
  const double x0 = box[0], y0 = box[1], z0 = box[2];
  float dt = (float)box[3];
  for (int l = 0; l < a.shift / 3; l++) dt *= 2;  // leafs2cells doubles d per level (:166)
  const double dd = (double)dt;
  const int ix = (int)((v.x - x0) / dd);
  const int iy = (int)((v.y - y0) / dd);
  const int iz = (int)((v.z - z0) / dd);
  double4 X;
  X.x = dd * (ix + .5) + x0;
  X.y = dd * (iy + .5) + y0;
  X.z = dd * (iz + .5) + z0;
  X.w = dd * .5;
  cX[c] = X;
  cRb[c] = X.w;
  cBody[c] = start;
  cNBody[c] = end - start;
  cKey[c] = (uint64_t)(int64_t)(int)(key[start] >> a.shift);
  cLevel[c] = a.level;
  if (a.leaf) {
    cNChild[c] = 0;
    cChild[c] = nc;  // reverseOrder maps ICHILD = 0 of a leaf to numCells
  } else {
    const int lo = a.scan_dn[start], hi = a.scan_dn[end];
    cNChild[c] = hi - lo;
    cChild[c] = a.off_dn + a.count_dn - hi;
  }
  if (a.level == 0) {
    cParent[c] = nc - 1;  // reverseOrder maps the root's IPARENT = 0 to numCells-1
  } else {
    const int rp = a.scan_up[start + 1] - 1;
    cParent[c] = a.off_up + a.count_up - 1 - rp;
  }
}

struct LevelArgsAll {
  LevelArgs a[kMaxUniformLevels + 1];
};

// k_level_cells of every level in one launch (level t = blockIdx.y).
__global__ void k_level_cells_all(LevelArgsAll all, const double4 *__restrict__ body, const uint64_t *__restrict__ key,
                                  int n, int nc, const double *__restrict__ box, double4 *__restrict__ cX,
                                  double *__restrict__ cRb, int *__restrict__ cParent, int *__restrict__ cChild,
                                  int *__restrict__ cNChild, int *__restrict__ cBody, int *__restrict__ cNBody,
                                  uint64_t *__restrict__ cKey, int *__restrict__ cLevel) {
  level_cell(all.a[blockIdx.y], body, key, n, nc, box, cX, cRb, cParent, cChild, cNChild, cBody, cNBody, cKey, cLevel);
}

__global__ void k_iota(int *__restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

void run_build_uniform(Handle &h) {
  const int n = h.n;
  const int ncrit = h.cfg.ncrit;
  // build_tree_omp.h:218 (integer n / ncrit inside the log).
  const int L = n >= ncrit ? 1 + int(std::log((double)(n / ncrit)) / M_LN2 / 3) : 0;
  if (L > 10) throw Error(EXAFMM_HIP_EUNSUP, "uniform tree deeper than 10 levels overflows the reference's int keys");
  upload_bounds(h);
  hipLaunchKernelGGL(k_box, dim3(1), dim3(1), 0, h.stream, h.box.p, L, h.box.p + 8);

  h.k64a.reserve(n + 1);
  h.k64b.reserve(n + 1);
  h.ia.reserve(n + 1);
  h.ib.reserve(n + 1);
  h.body4.resize(n);
  h.bidx.resize(n);
  h.trg4.resize(n);
  const int T = 256;
  // Keys of 3 L <= 30 bits: sorted as 32-bit ints (a radix sort over 64-bit
  // keys takes several times as many passes), then widened.  k64b holds the
  // two 32-bit key buffers.
  int *k32 = reinterpret_cast<int *>(h.k64b.p), *k32alt = k32 + (n + 1);
  hipLaunchKernelGGL(k_morton32, dim3(div_up(n, T)), dim3(T), 0, h.stream, h.in4.p, n, h.box.p + 8, L, k32, h.ia.p);
  bool alt = false;
  sort_pairs_int_int(h, k32, k32alt, h.ia.p, h.ib.p, n, 3 * L, alt);
  hipLaunchKernelGGL(k_widen_keys, dim3(div_up(n, T)), dim3(T), 0, h.stream, alt ? k32alt : k32, n, h.k64a.p);
  uint64_t *key = h.k64a.p;
  int *order = alt ? h.ib.p : h.ia.p;
  hipLaunchKernelGGL(k_gather_bodies, dim3(div_up(n, T)), dim3(T), 0, h.stream, h.in4.p, order, n, h.body4.p, h.bidx.p);

  // Per-level heads and scans: (L+1) x (n+1) ints each.
  const size_t stride = (size_t)n + 1;
  h.ic.reserve(stride * (L + 1));  // flags
  h.id.reserve(stride * (L + 1));  // scans
  std::vector<int> counts(L + 1);
  std::vector<int> tot(L + 1);
  // Heads of every level, one exclusive scan over all levels, each level's
  // start subtracted: the per-level exclusive scans.
  h.cnt.reserve(2 * (L + 1));
  hipLaunchKernelGGL(k_level_heads_all, dim3(div_up(n + 1, T), L + 1), dim3(T), 0, h.stream, key, n, L,
                     (int64_t)stride, h.ic.p);
  exclusive_scan_int(h, h.ic.p, h.id.p, (int)(stride * (L + 1)));
  hipLaunchKernelGGL(k_strided_gather, dim3(1), dim3(64), 0, h.stream, h.id.p, (int64_t)stride, L + 1,
                     h.cnt.p + L + 1);
  hipLaunchKernelGGL(k_level_rebase, dim3(div_up(n + 1, T), L + 1), dim3(T), 0, h.stream, h.id.p, n, (int64_t)stride,
                     h.cnt.p + L + 1);
  // Totals: scan[n] of each level, gathered on the device, one read-back.
  hipLaunchKernelGGL(k_strided_gather, dim3(1), dim3(64), 0, h.stream, h.id.p + n, (int64_t)stride, L + 1, h.cnt.p);
  EXA_HIP(hipMemcpyAsync(tot.data(), h.cnt.p, (L + 1) * sizeof(int), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  sync_bounds(h);  // (arrived before the totals)
  int nc = 0;
  std::vector<int> off(L + 2, 0);
  for (int t = 0; t <= L; t++) {
    counts[t] = tot[t];
    off[t] = nc;
    nc += counts[t];
  }
  off[L + 1] = nc;
  h.ncells = nc;
  h.nlevels = L + 1;
  h.ie.reserve(nc + 1);  // positions, level after level
  LevelOffsets posoff{};
  {
    int o = 0;
    for (int t = 0; t <= L; t++) {
      posoff.v[t] = o;
      o += counts[t];
    }
  }
  hipLaunchKernelGGL(k_level_positions_all, dim3(div_up(n, T), L + 1), dim3(T), 0, h.stream, h.ic.p, h.id.p, n,
                     (int64_t)stride, posoff, h.ie.p);
  h.cX.resize(nc);
  h.cRb.resize(nc);
  h.cParent.resize(nc);
  h.cChild.resize(nc);
  h.cNChild.resize(nc);
  h.cBody.resize(nc);
  h.cNBody.resize(nc);
  h.cKey.resize(nc);
  h.cLevel.resize(nc);
  LevelArgsAll all{};
  int maxcount = 0;
  for (int t = 0; t <= L; t++) {
    LevelArgs &a = all.a[t];
    a.pos = h.ie.p + posoff.v[t];
    a.scan = h.id.p + stride * t;
    a.scan_up = t > 0 ? h.id.p + stride * (t - 1) : nullptr;
    a.scan_dn = t < L ? h.id.p + stride * (t + 1) : nullptr;
    a.count = counts[t];
    a.count_up = t > 0 ? counts[t - 1] : 0;
    a.count_dn = t < L ? counts[t + 1] : 0;
    a.off = off[t];
    a.off_up = t > 0 ? off[t - 1] : 0;
    a.off_dn = t < L ? off[t + 1] : 0;
    a.shift = 3 * (L - t);
    a.leaf = (t == L);
    a.level = t;
    maxcount = std::max(maxcount, counts[t]);
  }
  hipLaunchKernelGGL(k_level_cells_all, dim3(div_up(maxcount, T), L + 1), dim3(T), 0, h.stream, all, h.body4.p, key, n,
                     nc, h.box.p + 8, h.cX.p, h.cRb.p, h.cParent.p, h.cChild.p, h.cNChild.p,
                     h.cBody.p, h.cNBody.p, h.cKey.p, h.cLevel.p);
  // Level order: the layout is already level-contiguous.
  h.lvl_off.assign(off.begin(), off.begin() + L + 2);
  h.lvl_cells.resize(nc);
  hipLaunchKernelGGL(k_iota, dim3(div_up(nc, T)), dim3(T), 0, h.stream, h.lvl_cells.p, nc);
  EXA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Forest of uniform trees (batched evaluation of independent systems).
// System s owns bodies [off[s], off[s+1]) of the input and gets its own
// bounds, depth L_s and Morton keys exactly as a single-system build of its
// bodies would (the same float/int arithmetic); the bodies are sorted by
// (s, key_s), so systems stay contiguous and in input order.  Level t holds
// the level-t cells of every system with L_s >= t; a system's cells are the
// ones its single-system build makes, in the same relative order.

// Per-system min/max, one workgroup per system.
__global__ void __launch_bounds__(256)
    k_bounds_seg(const double4 *__restrict__ b, const int *__restrict__ off, double *__restrict__ out) {
  __shared__ double s[6][256];
  const int sys = blockIdx.x, i0 = off[sys], i1 = off[sys + 1];
  double mn[3] = {b[i0].x, b[i0].y, b[i0].z};
  double mx[3] = {mn[0], mn[1], mn[2]};
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const double4 v = b[i];
    mn[0] = v.x < mn[0] ? v.x : mn[0];
    mn[1] = v.y < mn[1] ? v.y : mn[1];
    mn[2] = v.z < mn[2] ? v.z : mn[2];
    mx[0] = v.x > mx[0] ? v.x : mx[0];
    mx[1] = v.y > mx[1] ? v.y : mx[1];
    mx[2] = v.z > mx[2] ? v.z : mx[2];
  }
  for (int d = 0; d < 3; d++) {
    s[d][threadIdx.x] = mn[d];
    s[3 + d][threadIdx.x] = mx[d];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) {
        const double a = s[d][threadIdx.x + w], c = s[3 + d][threadIdx.x + w];
        s[d][threadIdx.x] = a < s[d][threadIdx.x] ? a : s[d][threadIdx.x];
        s[3 + d][threadIdx.x] = c > s[3 + d][threadIdx.x] ? c : s[3 + d][threadIdx.x];
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[6 * sys + threadIdx.x] = s[threadIdx.x][0];
}

struct ForestSys {
  double x0, y0, z0;  // bounds minimum (the key origin)
  float d;            // leaf cell size, 2 R0 / 2^L
  int L;              // depth
  int begin, end;     // body range
};

// System of every body (input order), by binary search over the offsets.
__global__ void k_sys_of(const int *__restrict__ off, int nsys, int n, int *__restrict__ sys) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = nsys;  // largest s with off[s] <= i
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  sys[i] = lo;
}

// Sort key (s << 32) | key_s, key_s from getKey at the system's depth.
__global__ void k_morton_forest(const double4 *__restrict__ in, int n, const int *__restrict__ sys,
                                const ForestSys *__restrict__ fs, uint64_t *__restrict__ key, int *__restrict__ idx) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int s = sys[b];
  const ForestSys f = fs[s];
  const double4 v = in[b];
  const double dd = (double)f.d;
  int ix = (int)((v.x - f.x0) / dd);
  int iy = (int)((v.y - f.y0) / dd);
  int iz = (int)((v.z - f.z0) / dd);
  int id = 0;
  for (int l = 0; l != f.L; ++l) {
    id += (ix & 1) << (3 * l);
    id += (iy & 1) << (3 * l + 1);
    id += (iz & 1) << (3 * l + 2);
    ix >>= 1;
    iy >>= 1;
    iz >>= 1;
  }
  key[b] = ((uint64_t)s << 32) | (uint32_t)id;
  idx[b] = b;
}

__device__ inline bool forest_prefix(uint64_t k, const ForestSys *__restrict__ fs, int t, int &s, uint32_t &pre) {
  s = (int)(k >> 32);
  const int L = fs[s].L;
  if (t > L) return false;
  pre = (uint32_t)k >> (3 * (L - t));
  return true;
}

__global__ void k_level_heads_forest(const uint64_t *__restrict__ key, int n, int t, const ForestSys *__restrict__ fs,
                                     int *__restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  int f = 0;
  int s, s2;
  uint32_t p, p2;
  if (i < n && forest_prefix(key[i], fs, t, s, p))
    f = i == 0 || !forest_prefix(key[i - 1], fs, t, s2, p2) || s2 != s || p2 != p;
  flags[i] = f;
}

// As k_level_cells, with the system's origin and cell size, the cell's body
// range cut at its system's end, and leaves wherever no child heads follow.
__global__ void k_level_cells_forest(LevelArgs a, const double4 *__restrict__ body, const uint64_t *__restrict__ key,
                                     int n, int nc, const ForestSys *__restrict__ fs, double4 *__restrict__ cX,
                                     double *__restrict__ cRb, int *__restrict__ cParent, int *__restrict__ cChild,
                                     int *__restrict__ cNChild, int *__restrict__ cBody, int *__restrict__ cNBody,
                                     uint64_t *__restrict__ cKey, int *__restrict__ cLevel) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.count) return;
  const int start = a.pos[r];
  const int sys = (int)(key[start] >> 32);
  const ForestSys f = fs[sys];
  int end = (r + 1 < a.count) ? a.pos[r + 1] : n;
  end = end < f.end ? end : f.end;
  const int c = a.off + a.count - 1 - r;
  float dt = f.d;
  for (int l = 0; l < f.L - a.level; l++) dt *= 2;  // leafs2cells doubles d per level (:166)
  const double4 v = body[start];
  const double dd = (double)dt;
  const int ix = (int)((v.x - f.x0) / dd);
  const int iy = (int)((v.y - f.y0) / dd);
  const int iz = (int)((v.z - f.z0) / dd);
  double4 X;
  X.x = dd * (ix + .5) + f.x0;
  X.y = dd * (iy + .5) + f.y0;
  X.z = dd * (iz + .5) + f.z0;
  X.w = dd * .5;
  cX[c] = X;
  cRb[c] = X.w;
  cBody[c] = start;
  cNBody[c] = end - start;
  cKey[c] = (uint64_t)(int64_t)(int)((uint32_t)key[start] >> (3 * (f.L - a.level)));
  cLevel[c] = a.level;
  const int nch = a.scan_dn ? a.scan_dn[end] - a.scan_dn[start] : 0;
  if (nch == 0) {
    cNChild[c] = 0;
    cChild[c] = nc;
  } else {
    cNChild[c] = nch;
    cChild[c] = a.off_dn + a.count_dn - a.scan_dn[end];
  }
  if (a.level == 0) {
    cParent[c] = nc - 1;
  } else {
    const int rp = a.scan_up[start + 1] - 1;
    cParent[c] = a.off_up + a.count_up - 1 - rp;
  }
}

void run_build_forest(Handle &h) {
  const int n = h.n, S = h.nsys, ncrit = h.cfg.ncrit;
  const int T = 256;
  const std::vector<int> &off = h.sys_off;
  h.sys_offd.resize(S + 1);
  EXA_HIP(hipMemcpyAsync(h.sys_offd.p, off.data(), (S + 1) * sizeof(int), hipMemcpyHostToDevice, h.stream));
  h.red.reserve(6 * (size_t)S);
  hipLaunchKernelGGL(k_bounds_seg, dim3(S), dim3(256), 0, h.stream, h.in4.p, h.sys_offd.p, h.red.p);
  std::vector<double> bb(6 * (size_t)S);
  EXA_HIP(hipMemcpyAsync(bb.data(), h.red.p, bb.size() * sizeof(double), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  std::vector<ForestSys> fs(S);
  int Lmax = 0;
  for (int s = 0; s < S; s++) {
    const int ns = off[s + 1] - off[s];
    // build_tree_omp.h:218 and bounds2box (:21-33), per system.
    const int L = ns >= ncrit ? 1 + int(std::log((double)(ns / ncrit)) / M_LN2 / 3) : 0;
    if (L > 10) throw Error(EXAFMM_HIP_EUNSUP, "uniform tree deeper than 10 levels overflows the reference's int keys");
    const double *mn = &bb[6 * s], *mx = &bb[6 * s + 3];
    double Xc[3], R0 = 0;
    for (int d = 0; d < 3; d++) Xc[d] = (mx[d] + mn[d]) / 2;
    for (int d = 0; d < 3; d++) {
      R0 = std::max(Xc[d] - mn[d], R0);
      R0 = std::max(mx[d] - Xc[d], R0);
    }
    R0 *= 1.00001;
    fs[s] = ForestSys{mn[0], mn[1], mn[2], (float)(2 * R0 / (1 << L)), L, off[s], off[s + 1]};
    Lmax = std::max(Lmax, L);
  }
  h.sys_par.resize(S * sizeof(ForestSys) / sizeof(double) + 1);
  ForestSys *fsd = reinterpret_cast<ForestSys *>(h.sys_par.p);
  EXA_HIP(hipMemcpyAsync(fsd, fs.data(), S * sizeof(ForestSys), hipMemcpyHostToDevice, h.stream));
  h.k64a.reserve(n + 1);
  h.k64b.reserve(n + 1);
  h.ia.reserve(n + 1);
  h.ib.reserve(n + 1);
  h.ie.reserve(n + 1);
  h.body4.resize(n);
  h.bidx.resize(n);
  h.trg4.resize(n);
  hipLaunchKernelGGL(k_sys_of, dim3(div_up(n, T)), dim3(T), 0, h.stream, h.sys_offd.p, S, n, h.ie.p);
  hipLaunchKernelGGL(k_morton_forest, dim3(div_up(n, T)), dim3(T), 0, h.stream, h.in4.p, n, h.ie.p, fsd, h.k64a.p,
                     h.ia.p);
  int sbits = 1;
  while ((1 << sbits) < S) sbits++;
  bool alt = false;
  sort_pairs_u64_int(h, h.k64a.p, h.k64b.p, h.ia.p, h.ib.p, n, 32 + sbits, alt);
  uint64_t *key = alt ? h.k64b.p : h.k64a.p;
  int *order = alt ? h.ib.p : h.ia.p;
  hipLaunchKernelGGL(k_gather_bodies, dim3(div_up(n, T)), dim3(T), 0, h.stream, h.in4.p, order, n, h.body4.p, h.bidx.p);
  const int L = Lmax;
  const size_t stride = (size_t)n + 1;
  h.ic.reserve(stride * (L + 1));
  h.id.reserve(stride * (L + 1));
  for (int t = 0; t <= L; t++) {
    int *fl = h.ic.p + stride * t;
    int *sc = h.id.p + stride * t;
    hipLaunchKernelGGL(k_level_heads_forest, dim3(div_up(n + 1, T)), dim3(T), 0, h.stream, key, n, t, fsd, fl);
    exclusive_scan_int(h, fl, sc, n + 1);
  }
  std::vector<int> counts(L + 1), tot(L + 1);
  h.cnt.reserve(L + 1);
  hipLaunchKernelGGL(k_strided_gather, dim3(1), dim3(64), 0, h.stream, h.id.p + n, (int64_t)stride, L + 1, h.cnt.p);
  EXA_HIP(hipMemcpyAsync(tot.data(), h.cnt.p, (L + 1) * sizeof(int), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  int nc = 0;
  std::vector<int> loff(L + 2, 0);
  for (int t = 0; t <= L; t++) {
    counts[t] = tot[t];
    loff[t] = nc;
    nc += counts[t];
  }
  loff[L + 1] = nc;
  h.ncells = nc;
  h.nlevels = L + 1;
  h.ie.reserve(std::max((size_t)nc + 1, (size_t)n + 1));
  std::vector<int> posoff(L + 1);
  for (int t = 0, o = 0; t <= L; t++) {
    posoff[t] = o;
    o += counts[t];
  }
  for (int t = 0; t <= L; t++)
    hipLaunchKernelGGL(k_level_positions, dim3(div_up(n, T)), dim3(T), 0, h.stream, h.ic.p + stride * t,
                       h.id.p + stride * t, n, h.ie.p + posoff[t]);
  h.cX.resize(nc);
  h.cRb.resize(nc);
  h.cParent.resize(nc);
  h.cChild.resize(nc);
  h.cNChild.resize(nc);
  h.cBody.resize(nc);
  h.cNBody.resize(nc);
  h.cKey.resize(nc);
  h.cLevel.resize(nc);
  for (int t = 0; t <= L; t++) {
    if (counts[t] == 0) continue;
    LevelArgs a;
    a.pos = h.ie.p + posoff[t];
    a.scan = h.id.p + stride * t;
    a.scan_up = t > 0 ? h.id.p + stride * (t - 1) : nullptr;
    a.scan_dn = t < L ? h.id.p + stride * (t + 1) : nullptr;
    a.count = counts[t];
    a.count_up = t > 0 ? counts[t - 1] : 0;
    a.count_dn = t < L ? counts[t + 1] : 0;
    a.off = loff[t];
    a.off_up = t > 0 ? loff[t - 1] : 0;
    a.off_dn = t < L ? loff[t + 1] : 0;
    a.shift = 0;
    a.leaf = 0;
    a.level = t;
    hipLaunchKernelGGL(k_level_cells_forest, dim3(div_up(counts[t], T)), dim3(T), 0, h.stream, a, h.body4.p, key, n,
                       nc, fsd, h.cX.p, h.cRb.p, h.cParent.p, h.cChild.p, h.cNChild.p, h.cBody.p, h.cNBody.p, h.cKey.p,
                       h.cLevel.p);
  }
  h.lvl_off.assign(loff.begin(), loff.begin() + L + 2);
  h.lvl_cells.resize(nc);
  hipLaunchKernelGGL(k_iota, dim3(div_up(nc, T)), dim3(T), 0, h.stream, h.lvl_cells.p, nc);
  EXA_HIP(hipGetLastError());
}

}  // namespace exafmm
