// Dual tree traversal as a breadth-first frontier of cell pairs, emitting
// P2P and M2L interaction lists (reference: include/traversal.h:37-124,
// 188-264).  Compiled with -ffp-contract=off: the MAC must round exactly as
// the reference does so that both produce the same lists.
//
// Each frontier pass classifies every (Ci, Cj, image) pair with the
// reference's rule -- M2L when |Xi - Xj - Xp|^2 > (Ri + Rj)^2, P2P when both
// are leaves, otherwise the split of splitCell (:188-212) -- and compacts the
// three outputs with exclusive scans, so the lists are deterministic.  The
// frontier visits the same set of pairs as the reference's TraverseRange
// recursion (mutual or not); only the order differs, and the lists are
// regrouped by target cell (CSR) afterwards.
#include "internal.h"

namespace exafmm {

struct CellView {
  const double4 *X;
  const int *child, *nchild, *nbody;
};

__device__ inline void image_offset(int img, double cycle, double &px, double &py, double &pz) {
  px = (img / 9 - 1) * cycle;
  py = ((img / 3) % 3 - 1) * cycle;
  pz = (img % 3 - 1) * cycle;
}

// action: 0 M2L, 1 P2P, 2 split Ci, 3 split Cj, 4 split both.
__device__ inline int classify(const CellView &cv, int ci, int cj, int img, double cycle, int nspawn, int mutual) {
#pragma clang fp contract(off)
  const double4 Xi = cv.X[ci], Xj = cv.X[cj];
  double px, py, pz;
  image_offset(img, cycle, px, py, pz);
  const double dx = Xi.x - Xj.x - px;
  const double dy = Xi.y - Xj.y - py;
  const double dz = Xi.z - Xj.z - pz;
  double R2 = 0;
  R2 += dx * dx;
  R2 += dy * dy;
  R2 += dz * dz;
  const double Rs = Xi.w + Xj.w;
  if (R2 > Rs * Rs) return 0;
  const int li = cv.nchild[ci] == 0, lj = cv.nchild[cj] == 0;
  if (li && lj) return cv.nbody[cj] == 0 ? 0 : 1;
  if (lj) return 2;
  if (li) return 3;
  if (cv.nbody[ci] + cv.nbody[cj] >= nspawn || (mutual && ci == cj)) return 4;
  return Xi.w >= Xj.w ? 2 : 3;
}

// With mutual (the reference's -m 1), a visited pair (Ci, Cj), Ci != Cj,
// stands for both directions: M2L and P2P emit both directed entries, and a
// cell split against itself visits only the child pairs (a, b), a <= b, as
// TraverseRange skips the (latter, former) quadrant (traversal.h:108-120).
// Output counts of one frontier pair: (M2L entries, P2P entries, next pairs).
__device__ inline int4 pair_counts(const CellView &cv, int4 f, int a, int mutual) {
  const int both = (mutual && f.x != f.y) ? 2 : 1;
  int nx = 0;
  if (a == 2) {
    nx = cv.nchild[f.x];
  } else if (a == 3) {
    nx = cv.nchild[f.y];
  } else if (a == 4) {
    const int nci = cv.nchild[f.x];
    nx = (mutual && f.x == f.y) ? nci * (nci + 1) / 2 : nci * cv.nchild[f.y];
  }
  return make_int4(a == 0 ? both : 0, a == 1 ? both : 0, nx, 0);
}

__device__ inline int4 add4(int4 a, int4 b) { return make_int4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

constexpr int kTravBlock = 256;

// Exclusive prefix of v over the 256 threads of the block (wave scans, then
// the wave totals); *total receives the block sum.
__device__ inline int4 block_exscan4(int4 v, int4 *total) {
  __shared__ int4 wsum[kTravBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int4 inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    const int4 t = make_int4(__shfl_up(inc.x, d), __shfl_up(inc.y, d), __shfl_up(inc.z, d), 0);
    if (lane >= d) inc = add4(inc, t);
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int4 base = make_int4(0, 0, 0, 0), tot = make_int4(0, 0, 0, 0);
  for (int k = 0; k < kTravBlock / 64; k++) {
    if (k < w) base = add4(base, wsum[k]);
    tot = add4(tot, wsum[k]);
  }
  __syncthreads();
  *total = tot;
  return make_int4(base.x + inc.x - v.x, base.y + inc.y - v.y, base.z + inc.z - v.z, 0);
}

// Pass 1 over the frontier: action per pair and the block's output totals.
__global__ void __launch_bounds__(kTravBlock)
    k_classify(CellView cv, const int4 *__restrict__ fr, int nf, double cycle, int nspawn, int mutual,
               int *__restrict__ act, int4 *__restrict__ blk) {
  const int i = blockIdx.x * kTravBlock + threadIdx.x;
  int4 c = make_int4(0, 0, 0, 0);
  if (i < nf) {
    const int4 f = fr[i];
    const int a = classify(cv, f.x, f.y, f.z, cycle, nspawn, mutual);
    act[i] = a;
    c = pair_counts(cv, f, a, mutual);
  }
  int4 tot;
  block_exscan4(c, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
  if (blockIdx.x == 0 && threadIdx.x == 0) blk[gridDim.x] = make_int4(0, 0, 0, 0);  // the scan ends in the totals
}

// Pass 2: each pair's output offsets = its block's offset (scan of the block
// totals) + its prefix inside the block; the same order as one global
// exclusive scan, so the lists are deterministic.
__global__ void __launch_bounds__(kTravBlock)
    k_emit(CellView cv, const int4 *__restrict__ fr, int nf, int mutual, const int *__restrict__ act,
           const int4 *__restrict__ blkofs, int64_t base_m2l, int64_t base_p2p, int *__restrict__ m2l_i,
           int *__restrict__ m2l_j, unsigned char *__restrict__ m2l_img, int *__restrict__ p2p_i,
           int *__restrict__ p2p_j, unsigned char *__restrict__ p2p_img, int4 *__restrict__ next) {
  const int i = blockIdx.x * kTravBlock + threadIdx.x;
  int4 f = make_int4(0, 0, 0, 0), c = make_int4(0, 0, 0, 0);
  int a = -1;
  if (i < nf) {
    f = fr[i];
    a = act[i];
    c = pair_counts(cv, f, a, mutual);
  }
  int4 tot;
  const int4 ofs = add4(block_exscan4(c, &tot), blkofs[blockIdx.x]);
  if (i >= nf) return;
  const bool both = mutual && f.x != f.y;
  if (a == 0) {
    const int64_t o = base_m2l + ofs.x;
    m2l_i[o] = f.x;
    m2l_j[o] = f.y;
    m2l_img[o] = (unsigned char)f.z;
    if (both) {
      m2l_i[o + 1] = f.y;
      m2l_j[o + 1] = f.x;
      m2l_img[o + 1] = (unsigned char)f.z;
    }
  } else if (a == 1) {
    const int64_t o = base_p2p + ofs.y;
    p2p_i[o] = f.x;
    p2p_j[o] = f.y;
    p2p_img[o] = (unsigned char)f.z;
    if (both) {
      p2p_i[o + 1] = f.y;
      p2p_j[o + 1] = f.x;
      p2p_img[o + 1] = (unsigned char)f.z;
    }
  } else {
    int o = ofs.z;
    if (a == 2) {
      const int c0 = cv.child[f.x], nc = cv.nchild[f.x];
      for (int c = 0; c < nc; c++) next[o++] = make_int4(c0 + c, f.y, f.z, 0);
    } else if (a == 3) {
      const int c0 = cv.child[f.y], nc = cv.nchild[f.y];
      for (int c = 0; c < nc; c++) next[o++] = make_int4(f.x, c0 + c, f.z, 0);
    } else {
      const int ci0 = cv.child[f.x], nci = cv.nchild[f.x];
      const int cj0 = cv.child[f.y], ncj = cv.nchild[f.y];
      if (mutual && f.x == f.y) {
        for (int a1 = 0; a1 < nci; a1++)
          for (int b1 = a1; b1 < nci; b1++) next[o++] = make_int4(ci0 + a1, ci0 + b1, f.z, 0);
      } else {
        for (int a1 = 0; a1 < nci; a1++)
          for (int b1 = 0; b1 < ncj; b1++) next[o++] = make_int4(ci0 + a1, cj0 + b1, f.z, 0);
      }
    }
  }
}

__global__ void k_init_frontier(int4 *__restrict__ fr, int images) {
  const int i = threadIdx.x;
  if (images == 0) {
    if (i == 0) fr[0] = make_int4(0, 0, 13, 0);
  } else if (i < 27) {
    fr[i] = make_int4(0, 0, i, 0);
  }
}

// A forest (batched systems): every root against itself, no images; the
// traversal never pairs cells of two systems.
__global__ void k_init_frontier_roots(int4 *__restrict__ fr, const int *__restrict__ roots, int nroots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nroots) fr[i] = make_int4(roots[i], roots[i], 13, 0);
}

// CSR offsets from the target-sorted keys: off[c] = first index with key >= c.
__global__ void k_csr_offsets(const int *__restrict__ keys, int64_t n, int nc, int *__restrict__ off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > nc) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  off[c] = (int)lo;
}

__global__ void k_permute_pairs(const int *__restrict__ order, int64_t n, const int *__restrict__ src_in,
                                const unsigned char *__restrict__ img_in, int *__restrict__ src_out,
                                unsigned char *__restrict__ img_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = order[i];
  src_out[i] = src_in[s];
  img_out[i] = img_in[s];
}

__global__ void k_flag_nonzero(const int *__restrict__ off, int n, int *__restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) flag[i] = i < n ? (off[i + 1] > off[i]) : 0;
}

__global__ void k_compact(const int *__restrict__ flag, const int *__restrict__ scan, int n, int *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[scan[i]] = i;
}

__global__ void k_seq(int *__restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (int)i;
}

// Grow a buffer, keeping its first `keep` elements.
template <typename T>
static void grow_keep(Handle &h, DBuf<T> &b, size_t want, size_t keep) {
  if (want <= b.cap) return;
  DBuf<T> nb;
  nb.reserve(want * 2);
  if (keep) EXA_HIP(hipMemcpyAsync(nb.p, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  b.release();
  b.p = nb.p;
  b.cap = nb.cap;
  nb.p = nullptr;
  nb.cap = 0;
}

// Group (tgt[k], src[k], img[k]) by target: CSR offsets off[0..ncells], the
// sources and images in target order (stable w.r.t. emission order), and the
// compact list of targets with at least one entry; tgt_out (if given) gets
// the target of every CSR entry.
static void group_by_target(Handle &h, int64_t np, DBuf<int> &tgt, DBuf<int> &src, DBuf<unsigned char> &img,
                            DBuf<int> &off, DBuf<int> &src_out, DBuf<unsigned char> &img_out, DBuf<int> &tlist,
                            int &ntl, DBuf<int> *tgt_out = nullptr) {
  const int nc = h.ncells;
  const int T = 256;
  off.resize(nc + 1);
  src_out.resize(np);
  img_out.resize(np);
  h.ia.reserve(np + 1);
  h.ib.reserve(np + 1);
  h.ic.reserve(np + 1);
  h.id.reserve(np + 1);
  const int *keys = h.ic.p;
  if (np > 0) {
    hipLaunchKernelGGL(k_seq, dim3(div_up(np, T)), dim3(T), 0, h.stream, h.ia.p, np);
    int bits = 1;
    while ((1 << bits) < nc) bits++;
    bool alt = false;
    // Keys: a copy of the targets (sorted stably), values: emission index.
    EXA_HIP(hipMemcpyAsync(h.ic.p, tgt.p, np * sizeof(int), hipMemcpyDeviceToDevice, h.stream));
    sort_pairs_int_int(h, h.ic.p, h.id.p, h.ia.p, h.ib.p, (int)np, bits, alt);
    const int *order = alt ? h.ib.p : h.ia.p;
    keys = alt ? h.id.p : h.ic.p;
    hipLaunchKernelGGL(k_permute_pairs, dim3(div_up(np, T)), dim3(T), 0, h.stream, order, np, src.p, img.p, src_out.p,
                       img_out.p);
    if (tgt_out) {
      tgt_out->resize(np);
      EXA_HIP(hipMemcpyAsync(tgt_out->p, keys, np * sizeof(int), hipMemcpyDeviceToDevice, h.stream));
    }
  }
  hipLaunchKernelGGL(k_csr_offsets, dim3(div_up(nc + 1, T)), dim3(T), 0, h.stream, keys, np, nc, off.p);
  // Compact non-empty targets.
  h.ie.reserve(2 * (size_t)nc + 2);
  int *flag = h.ie.p, *scan = h.ie.p + nc + 1;
  hipLaunchKernelGGL(k_flag_nonzero, dim3(div_up(nc + 1, T)), dim3(T), 0, h.stream, off.p, nc, flag);
  exclusive_scan_int(h, flag, scan, nc + 1);
  tlist.resize(nc);
  hipLaunchKernelGGL(k_compact, dim3(div_up(nc, T)), dim3(T), 0, h.stream, flag, scan, nc, tlist.p);
  EXA_HIP(hipMemcpyAsync(&ntl, scan + nc, sizeof(int), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
}

void run_lists(Handle &h) {
  const int T = 256;
  CellView cv{h.cX.p, h.cChild.p, h.cNChild.p, h.cNBody.p};
  // Raw (emission-order) lists, kept on the handle across evaluations.
  DBuf<int> &mi = h.raw_mi, &mj = h.raw_mj, &pi = h.raw_pi, &pj = h.raw_pj;
  DBuf<unsigned char> &mimg = h.raw_mimg, &pimg = h.raw_pimg;
  int64_t nm = 0, np = 0;
  int nf = h.cfg.images == 0 ? 1 : 27;
  // Periodic images are traversed non-mutually (traversal.h:250-259).
  const int mutual = (h.cfg.mutual && h.cfg.images == 0) ? 1 : 0;
  if (h.nsys > 1) {
    nf = h.lvl_off[1] - h.lvl_off[0];  // the roots: level 0 of the forest
    h.fa.reserve(nf + 64);
    hipLaunchKernelGGL(k_init_frontier_roots, dim3(div_up(nf, 256)), dim3(256), 0, h.stream, h.fa.p,
                       h.lvl_cells.p + h.lvl_off[0], nf);
  } else {
    h.fa.reserve(64);
    hipLaunchKernelGGL(k_init_frontier, dim3(1), dim3(64), 0, h.stream, h.fa.p, h.cfg.images);
  }
  // Start the raw lists at a capacity that covers uniform trees without
  // regrowth (about 180 M2L and 65 P2P sources per cell at theta = 0.4).
  grow_keep(h, mi, (size_t)h.ncells * 128 + 1024, 0);
  grow_keep(h, mj, (size_t)h.ncells * 128 + 1024, 0);
  grow_keep(h, mimg, (size_t)h.ncells * 128 + 1024, 0);
  grow_keep(h, pi, (size_t)h.ncells * 48 + 1024, 0);
  grow_keep(h, pj, (size_t)h.ncells * 48 + 1024, 0);
  grow_keep(h, pimg, (size_t)h.ncells * 48 + 1024, 0);
  h.cnt.reserve(4);
  while (nf > 0) {
    const int nb = div_up(nf, kTravBlock);
    h.ia.reserve(nf + 1);  // action
    h.ib.reserve(4 * ((size_t)nb + 1));  // block totals (int4)
    h.ic.reserve(4 * ((size_t)nb + 1));  // block offsets (int4)
    int *act = h.ia.p;
    int4 *blk = reinterpret_cast<int4 *>(h.ib.p), *blkofs = reinterpret_cast<int4 *>(h.ic.p);
    hipLaunchKernelGGL(k_classify, dim3(nb), dim3(kTravBlock), 0, h.stream, cv, h.fa.p, nf, h.cfg.cycle,
                       h.cfg.nspawn, mutual, act, blk);
    exclusive_scan_int4(h, blk, blkofs, nb + 1);
    int4 tot4;
    EXA_HIP(hipMemcpyAsync(&tot4, blkofs + nb, sizeof(tot4), hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    const int tot[3] = {tot4.x, tot4.y, tot4.z};
    grow_keep(h, mi, nm + tot[0] + 1, nm);
    grow_keep(h, mj, nm + tot[0] + 1, nm);
    grow_keep(h, mimg, nm + tot[0] + 1, nm);
    grow_keep(h, pi, np + tot[1] + 1, np);
    grow_keep(h, pj, np + tot[1] + 1, np);
    grow_keep(h, pimg, np + tot[1] + 1, np);
    h.fb.reserve((size_t)tot[2] + 1);
    hipLaunchKernelGGL(k_emit, dim3(nb), dim3(kTravBlock), 0, h.stream, cv, h.fa.p, nf, mutual, act, blkofs, nm, np,
                       mi.p, mj.p, mimg.p, pi.p, pj.p, pimg.p, h.fb.p);
    nm += tot[0];
    np += tot[1];
    nf = tot[2];
    std::swap(h.fa.p, h.fb.p);
    std::swap(h.fa.cap, h.fb.cap);
  }
  h.nm2l = nm;
  h.np2p = np;
  group_by_target(h, nm, mi, mj, mimg, h.m2l_off, h.m2l_src, h.m2l_img, h.m2l_tgt, h.nm2l_tgt,
                  lattice_applies(h) ? &h.lat_ptgt : nullptr);
  group_by_target(h, np, pi, pj, pimg, h.p2p_off, h.p2p_src, h.p2p_img, h.p2p_tgt, h.np2p_tgt);
  EXA_HIP(hipGetLastError());
}

}  // namespace exafmm
