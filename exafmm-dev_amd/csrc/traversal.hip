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
#include <algorithm>
#include <climits>

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

// A frontier pair packed in 8 bytes: (Ci, Cj | image << 27); the host keeps
// cell indices below 2^27.
constexpr int kImgShift = 27;
__device__ inline int2 pack_pair(int ci, int cj, int img) {
  return make_int2(ci, (int)((unsigned)cj | ((unsigned)img << kImgShift)));
}
__device__ inline int4 unpack_pair(int2 p) {
  return make_int4(p.x, p.y & ((1 << kImgShift) - 1), (int)((unsigned)p.y >> kImgShift), 0);
}

// What the rule reads of the two cells of a pair, loaded up front (so that
// the loads of several pairs are in flight at once).
struct PairCells {
  double4 Xi, Xj;
  int nci, ncj, nbi, nbj;
};

__device__ inline PairCells load_cells(const CellView &cv, int ci, int cj) {
  return PairCells{cv.X[ci], cv.X[cj], cv.nchild[ci], cv.nchild[cj], cv.nbody[ci], cv.nbody[cj]};
}

// action: 0 M2L, 1 P2P, 2 split Ci, 3 split Cj, 4 split both.
__device__ inline int classify_cells(const PairCells &pc, int ci, int cj, int img, double cycle, int nspawn,
                                     int mutual) {
#pragma clang fp contract(off)
  const double4 Xi = pc.Xi, Xj = pc.Xj;
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
  const int li = pc.nci == 0, lj = pc.ncj == 0;
  if (li && lj) return pc.nbj == 0 ? 0 : 1;
  if (lj) return 2;
  if (li) return 3;
  if (pc.nbi + pc.nbj >= nspawn || (mutual && ci == cj)) return 4;
  return Xi.w >= Xj.w ? 2 : 3;
}

__device__ inline int classify(const CellView &cv, int ci, int cj, int img, double cycle, int nspawn, int mutual) {
  return classify_cells(load_cells(cv, ci, cj), ci, cj, img, cycle, nspawn, mutual);
}

// With mutual (the reference's -m 1), a visited pair (Ci, Cj), Ci != Cj,
// stands for both directions: M2L and P2P emit both directed entries, and a
// cell split against itself visits only the child pairs (a, b), a <= b, as
// TraverseRange skips the (latter, former) quadrant (traversal.h:108-120).
// Output counts of one frontier pair: (M2L entries, P2P entries, next pairs);
// nci, ncj: the child counts of its two cells.
__device__ inline int4 pair_counts_n(int4 f, int a, int mutual, int nci, int ncj) {
  const int both = (mutual && f.x != f.y) ? 2 : 1;
  int nx = 0;
  if (a == 2) {
    nx = nci;
  } else if (a == 3) {
    nx = ncj;
  } else if (a == 4) {
    nx = (mutual && f.x == f.y) ? nci * (nci + 1) / 2 : nci * ncj;
  }
  return make_int4(a == 0 ? both : 0, a == 1 ? both : 0, nx, 0);
}

__device__ inline int4 pair_counts(const CellView &cv, int4 f, int a, int mutual) {
  if (a < 2) return pair_counts_n(f, a, mutual, 0, 0);
  return pair_counts_n(f, a, mutual, cv.nchild[f.x], cv.nchild[f.y]);
}

__device__ inline int4 add4(int4 a, int4 b) { return make_int4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

constexpr int kTravBlock = 256;

// Exclusive prefix of v over the NT threads of the block (wave scans, then
// the wave totals); *total receives the block sum.
template <int NT = kTravBlock>
__device__ inline int4 block_exscan4(int4 v, int4 *total) {
  __shared__ int4 wsum[NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int4 inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    const int4 t = make_int4(__shfl_up(inc.x, d), __shfl_up(inc.y, d), __shfl_up(inc.z, d), 0);
    if (lane >= d) inc = add4(inc, t);
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int4 base = make_int4(0, 0, 0, 0), tot = make_int4(0, 0, 0, 0);
  if (NT <= 256) {
    for (int k = 0; k < NT / 64; k++) {
      if (k < w) base = add4(base, wsum[k]);
      tot = add4(tot, wsum[k]);
    }
  } else {  // many waves: the first wave scans the wave totals in place
    if (w == 0) {
      int4 x = lane < NT / 64 ? wsum[lane] : make_int4(0, 0, 0, 0);
      for (int d = 1; d < NT / 64; d <<= 1) {
        const int4 t = make_int4(__shfl_up(x.x, d), __shfl_up(x.y, d), __shfl_up(x.z, d), 0);
        if (lane >= d) x = add4(x, t);
      }
      if (lane < NT / 64) wsum[lane] = x;
    }
    __syncthreads();
    if (w > 0) base = wsum[w - 1];
    tot = wsum[NT / 64 - 1];
  }
  __syncthreads();
  *total = tot;
  return make_int4(base.x + inc.x - v.x, base.y + inc.y - v.y, base.z + inc.z - v.z, 0);
}

// Pass 1 over the frontier: action per pair and the block's output totals.
__global__ void __launch_bounds__(kTravBlock)
    k_classify(CellView cv, const int2 *__restrict__ fr, int nf, double cycle, int nspawn, int mutual,
               unsigned char *__restrict__ act, int4 *__restrict__ blk) {
  const int i = blockIdx.x * kTravBlock + threadIdx.x;
  int4 c = make_int4(0, 0, 0, 0);
  if (i < nf) {
    const int4 f = unpack_pair(fr[i]);
    const int a = classify(cv, f.x, f.y, f.z, cycle, nspawn, mutual);
    act[i] = (unsigned char)a;
    c = pair_counts(cv, f, a, mutual);
  }
  int4 tot;
  block_exscan4(c, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
  if (blockIdx.x == 0 && threadIdx.x == 0) blk[gridDim.x] = make_int4(0, 0, 0, 0);  // the scan ends in the totals
}

// The outputs of frontier pair f (action a) at offsets ofs (M2L, P2P, next
// frontier) past the level's list bases.
// lk: (first child, child count) of Ci, then of Cj (read for splits only).
__device__ inline void emit_pair_n(int4 f, int a, int4 lk, int mutual, int4 ofs, int64_t base_m2l,
                                   int64_t base_p2p, int *__restrict__ m2l_i, int *__restrict__ m2l_j,
                                   unsigned char *__restrict__ m2l_img, int *__restrict__ p2p_i,
                                   int *__restrict__ p2p_j, unsigned char *__restrict__ p2p_img,
                                   int2 *__restrict__ next) {
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
    const int ci0 = lk.x, nci = lk.y, cj0 = lk.z, ncj = lk.w;
    if (a == 2) {
      for (int c = 0; c < nci; c++) next[o++] = pack_pair(ci0 + c, f.y, f.z);
    } else if (a == 3) {
      for (int c = 0; c < ncj; c++) next[o++] = pack_pair(f.x, cj0 + c, f.z);
    } else {
      if (mutual && f.x == f.y) {
        for (int a1 = 0; a1 < nci; a1++)
          for (int b1 = a1; b1 < nci; b1++) next[o++] = pack_pair(ci0 + a1, ci0 + b1, f.z);
      } else {
        for (int a1 = 0; a1 < nci; a1++)
          for (int b1 = 0; b1 < ncj; b1++) next[o++] = pack_pair(ci0 + a1, cj0 + b1, f.z);
      }
    }
  }
}

__device__ inline int4 load_links(const CellView &cv, int4 f) {
  return make_int4(cv.child[f.x], cv.nchild[f.x], cv.child[f.y], cv.nchild[f.y]);
}

__device__ inline void emit_pair(const CellView &cv, int4 f, int a, int mutual, int4 ofs, int64_t base_m2l,
                                 int64_t base_p2p, int *__restrict__ m2l_i, int *__restrict__ m2l_j,
                                 unsigned char *__restrict__ m2l_img, int *__restrict__ p2p_i,
                                 int *__restrict__ p2p_j, unsigned char *__restrict__ p2p_img,
                                 int2 *__restrict__ next) {
  const int4 lk = a >= 2 ? load_links(cv, f) : make_int4(0, 0, 0, 0);
  emit_pair_n(f, a, lk, mutual, ofs, base_m2l, base_p2p, m2l_i, m2l_j, m2l_img, p2p_i, p2p_j, p2p_img, next);
}

// Pass 2: each pair's output offsets = its block's offset (scan of the block
// totals) + its prefix inside the block; the same order as one global
// exclusive scan, so the lists are deterministic.
__global__ void __launch_bounds__(kTravBlock)
    k_emit(CellView cv, const int2 *__restrict__ fr, int nf, int mutual, const unsigned char *__restrict__ act,
           const int4 *__restrict__ blkofs, int64_t base_m2l, int64_t base_p2p, int *__restrict__ m2l_i,
           int *__restrict__ m2l_j, unsigned char *__restrict__ m2l_img, int *__restrict__ p2p_i,
           int *__restrict__ p2p_j, unsigned char *__restrict__ p2p_img, int2 *__restrict__ next) {
  const int i = blockIdx.x * kTravBlock + threadIdx.x;
  int4 f = make_int4(0, 0, 0, 0), c = make_int4(0, 0, 0, 0);
  int a = -1;
  if (i < nf) {
    f = unpack_pair(fr[i]);
    a = act[i];
    c = pair_counts(cv, f, a, mutual);
  }
  int4 tot;
  const int4 ofs = add4(block_exscan4(c, &tot), blkofs[blockIdx.x]);
  if (i >= nf) return;
  emit_pair(cv, f, a, mutual, ofs, base_m2l, base_p2p, m2l_i, m2l_j, m2l_img, p2p_i, p2p_j, p2p_img, next);
}

// ---------------------------------------------------------------------------
// The same passes driven from the device: the frontier size, the list
// lengths and which frontier buffer is current live in a TravState on the
// device, so the host enqueues several levels without waiting for any of
// them.  A fixed grid of kTravGrid blocks takes contiguous chunks of the
// frontier (multiples of 64 pairs, in order), so the emission order -- and
// the lists -- are those of the per-level passes above.  Outputs that would
// not fit the buffers set `overflow` and stop the traversal (the host then
// runs the per-level passes, which grow the buffers).
constexpr int kTravGrid = 512;

struct TravState {
  long long nm, np;              // list lengths so far
  int nf, fsel;                  // frontier size, current buffer (0: fa, 1: fb)
  int overflow, levels;
  int levels_small;              // levels done by k_trav_small
  int done_blocks;               // k_emit_dev blocks finished (the last one advances the level)
};

constexpr int kSmallFrontier = 512;  // levels up to this size run in one workgroup
constexpr int kSmallBlock = 1024;

// The device-driven passes run NT threads per block, each taking K
// consecutive frontier pairs per step (pairs s + K * t + (0 .. K - 1): in
// order, so the block scan of the per-thread totals still gives the global
// order) and issuing all their loads before it uses any.  Chunks are whole
// multiples of 64 pairs (a step may end early), so that a small level with
// many children per pair still spreads over many blocks.  Measured on the
// 1M-body cube (six levels, up to 6 M pairs): classification with 1024
// threads and one pair each, emission with 512 threads and two pairs each
// (the fastest of 256/512/1024 threads with 1, 2 or 4 pairs).
constexpr int kClassifyBlock = 1024;
constexpr int kEmitBlock = 512, kEmitItems = 2;

__device__ inline void chunk_of(int nf, int &i0, int &i1) {
  const int c = ((nf + kTravGrid - 1) / kTravGrid + 63) / 64 * 64;
  i0 = min(nf, (int)blockIdx.x * c);
  i1 = min(nf, i0 + c);
}

__global__ void k_trav_init(TravState *st, int nf) {
  st->nm = st->np = 0;
  st->nf = nf;
  st->fsel = 0;
  st->overflow = 0;
  st->levels = 0;
  st->levels_small = 0;
  st->done_blocks = 0;
}

// The first levels, while 0 < nf <= kSmallFrontier, in one workgroup: both
// passes of a level back to back (the frontier in 1024-pair chunks, in
// order), no launch per level.
__global__ void __launch_bounds__(kSmallBlock)
    k_trav_small(CellView cv, int2 *__restrict__ fa, int2 *__restrict__ fb, TravState *st, double cycle,
                 int nspawn, int mutual, unsigned char *__restrict__ act, int *__restrict__ m2l_i,
                 int *__restrict__ m2l_j, unsigned char *__restrict__ m2l_img, int *__restrict__ p2p_i,
                 int *__restrict__ p2p_j, unsigned char *__restrict__ p2p_img, long long cap_m2l, long long cap_p2p,
                 int cap_f) {
  __shared__ int s_nf, s_fsel;
  __shared__ long long s_nm, s_np;
  if (threadIdx.x == 0) {
    s_nf = st->nf;
    s_fsel = st->fsel;
    s_nm = st->nm;
    s_np = st->np;
  }
  __syncthreads();
  int levels = 0;
  while (true) {
    const int nf = s_nf;
    if (nf == 0 || nf > kSmallFrontier) break;
    const int2 *fr = s_fsel ? fb : fa;
    int2 *next = s_fsel ? fa : fb;
    const long long nm = s_nm, np = s_np;
    int4 acc = make_int4(0, 0, 0, 0);
    for (int s0 = 0; s0 < nf; s0 += kSmallBlock) {
      const int i = s0 + threadIdx.x;
      int4 c = make_int4(0, 0, 0, 0);
      if (i < nf) {
        const int4 f = unpack_pair(fr[i]);
        const int a = classify(cv, f.x, f.y, f.z, cycle, nspawn, mutual);
        act[i] = (unsigned char)a;
        c = pair_counts(cv, f, a, mutual);
      }
      int4 tot;
      block_exscan4<kSmallBlock>(c, &tot);
      acc = add4(acc, tot);
    }
    if (!(nm + acc.x <= cap_m2l && np + acc.y <= cap_p2p && acc.z <= cap_f)) {
      if (threadIdx.x == 0) st->overflow = 1;
      break;
    }
    int4 run = make_int4(0, 0, 0, 0);
    for (int s0 = 0; s0 < nf; s0 += kSmallBlock) {
      const int i = s0 + threadIdx.x;
      int4 f = make_int4(0, 0, 0, 0), c = make_int4(0, 0, 0, 0);
      int a = -1;
      if (i < nf) {
        f = unpack_pair(fr[i]);
        a = act[i];
        c = pair_counts(cv, f, a, mutual);
      }
      int4 tot;
      const int4 ofs = add4(block_exscan4<kSmallBlock>(c, &tot), run);
      run = add4(run, tot);
      if (i < nf) emit_pair(cv, f, a, mutual, ofs, nm, np, m2l_i, m2l_j, m2l_img, p2p_i, p2p_j, p2p_img, next);
    }
    levels++;
    __syncthreads();
    if (threadIdx.x == 0) {
      s_nm = nm + acc.x;
      s_np = np + acc.y;
      s_nf = acc.z;
      s_fsel ^= 1;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (st->overflow) {
      st->nf = 0;
    } else {
      st->nf = s_nf;
      st->fsel = s_fsel;
      st->nm = s_nm;
      st->np = s_np;
      st->levels += levels;
      st->levels_small = levels;
    }
  }
}

template <int NT, int K>
__global__ void __launch_bounds__(NT, NT / 128)  // two blocks per CU
    k_classify_dev(CellView cv, const int2 *__restrict__ fa, const int2 *__restrict__ fb, const TravState *st,
                   double cycle, int nspawn, int mutual, unsigned char *__restrict__ act, int4 *__restrict__ blk) {
  const int nf = st->nf;
  const int2 *fr = st->fsel ? fb : fa;
  int i0, i1;
  chunk_of(nf, i0, i1);
  int4 acc = make_int4(0, 0, 0, 0);
  for (int s = i0; s < i1; s += NT * K) {
    const int ib = s + threadIdx.x * K;
    int4 f[K];
    PairCells pc[K];
#pragma unroll
    for (int k = 0; k < K; k++) f[k] = ib + k < i1 ? unpack_pair(fr[ib + k]) : make_int4(0, 0, 13, 0);
#pragma unroll
    for (int k = 0; k < K; k++) pc[k] = load_cells(cv, f[k].x, f[k].y);
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (ib + k < i1) {
        const int a = classify_cells(pc[k], f[k].x, f[k].y, f[k].z, cycle, nspawn, mutual);
        act[ib + k] = (unsigned char)a;
        acc = add4(acc, pair_counts_n(f[k], a, mutual, pc[k].nci, pc[k].ncj));
      }
    }
  }
  int4 tot;
  block_exscan4<NT>(acc, &tot);  // only the block total is kept
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// Second pass of a level.  Every block first scans the kTravGrid block
// totals of the first pass itself (8 KB, read by all blocks: no separate
// scan launch), which gives its output offsets, the level's totals and the
// capacity check (the same for every block); then it emits its chunk.  The
// last block to finish moves the state to the next level.
template <int NT, int K>
__global__ void __launch_bounds__(NT, NT / 128)  // two blocks per CU
    k_emit_dev(CellView cv, int2 *__restrict__ fa, int2 *__restrict__ fb, TravState *st, int mutual,
               const unsigned char *__restrict__ act, const int4 *__restrict__ blk, long long cap_m2l,
               long long cap_p2p, int cap_f, int *__restrict__ m2l_i, int *__restrict__ m2l_j,
               unsigned char *__restrict__ m2l_img, int *__restrict__ p2p_i, int *__restrict__ p2p_j,
               unsigned char *__restrict__ p2p_img) {
  constexpr int PER = (kTravGrid + NT - 1) / NT;
  __shared__ int s_last;
  __shared__ int4 s_run;
  const int nf = st->nf;
  const int fsel = st->fsel;
  const long long bm = st->nm, bp = st->np;
  const int2 *fr = fsel ? fb : fa;
  int2 *next = fsel ? fa : fb;
  int4 tot;
  {
    int4 v[PER];
    int4 sum = make_int4(0, 0, 0, 0);
    for (int k = 0; k < PER; k++) {
      const int b = threadIdx.x * PER + k;
      v[k] = b < kTravGrid ? blk[b] : make_int4(0, 0, 0, 0);
      sum = add4(sum, v[k]);
    }
    int4 r = block_exscan4<NT>(sum, &tot);
    for (int k = 0; k < PER; k++) {
      if (threadIdx.x * PER + k == (int)blockIdx.x) s_run = r;
      r = add4(r, v[k]);
    }
  }
  __syncthreads();
  const bool ok = bm + tot.x <= cap_m2l && bp + tot.y <= cap_p2p && tot.z <= cap_f;
  const bool emit = ok && nf > 0;
  int i0, i1;
  chunk_of(nf, i0, i1);
  if (!emit) i1 = i0;
  int4 run = s_run;
  for (int s = i0; s < i1; s += NT * K) {
    const int ib = s + threadIdx.x * K;
    int4 f[K], lk[K];
    int a[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const bool in = ib + k < i1;
      f[k] = in ? unpack_pair(fr[ib + k]) : make_int4(0, 0, 0, 0);
      a[k] = in ? act[ib + k] : -1;
    }
#pragma unroll
    for (int k = 0; k < K; k++) lk[k] = a[k] >= 2 ? load_links(cv, f[k]) : make_int4(0, 0, 0, 0);
    int4 c[K], sum = make_int4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < K; k++) {
      c[k] = pair_counts_n(f[k], a[k], mutual, lk[k].y, lk[k].w);
      sum = add4(sum, c[k]);
    }
    int4 t4;
    int4 ofs = add4(block_exscan4<NT>(sum, &t4), run);
    run = add4(run, t4);
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (a[k] >= 0)
        emit_pair_n(f[k], a[k], lk[k], mutual, ofs, bm, bp, m2l_i, m2l_j, m2l_img, p2p_i, p2p_j, p2p_img, next);
      ofs = add4(ofs, c[k]);
    }
  }
  // The last block to finish (every block has read the state by then) moves
  // the state to the next level; the next kernel sees it.
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&st->done_blocks, 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (s_last && threadIdx.x == 0) {
    st->done_blocks = 0;
    if (!ok) st->overflow = 1;
    if (!emit) {
      st->nf = 0;
    } else {
      st->nm = bm + tot.x;
      st->np = bp + tot.y;
      st->nf = tot.z;
      st->fsel = fsel ^ 1;
      st->levels++;
    }
  }
}

__global__ void k_init_frontier(int2 *__restrict__ fr, int images) {
  const int i = threadIdx.x;
  if (images == 0) {
    if (i == 0) fr[0] = pack_pair(0, 0, 13);
  } else if (i < 27) {
    fr[i] = pack_pair(0, 0, i);
  }
}

// A forest (batched systems): every root against itself, no images; the
// traversal never pairs cells of two systems.
__global__ void k_init_frontier_roots(int2 *__restrict__ fr, const int *__restrict__ roots, int nroots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nroots) fr[i] = pack_pair(roots[i], roots[i], 13);
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

// Scratch and stream of one list grouping (the M2L and P2P groupings run on
// two streams at once).
struct GroupCtx {
  DBuf<int> &a, &b, &c, &d, &e;
  DBuf<unsigned char> &tmp;
  hipStream_t st;
};

// Group (tgt[k], src[k], img[k]) by target: CSR offsets off[0..ncells], the
// sources and images in target order (stable w.r.t. emission order), and the
// compact list of targets with at least one entry, whose count is copied to
// *ntl_host (pinned; valid after g.st is synchronised) and left on the
// device at *ntl_dev (if given; valid until the scratch g.e is reused).
// Nothing waits for the device.
static void group_by_target(Handle &h, GroupCtx g, int64_t np, DBuf<int> &tgt, DBuf<int> &src,
                            DBuf<unsigned char> &img, DBuf<int> &off, DBuf<int> &src_out, DBuf<unsigned char> &img_out,
                            DBuf<int> &tlist, int *ntl_host, const int **ntl_dev = nullptr) {
  const int nc = h.ncells;
  const int T = 256;
  if (np > INT_MAX) throw Error(EXAFMM_HIP_EUNSUP, "list grouping: more than 2^31 - 1 entries");
  off.resize(nc + 1);
  src_out.resize(np);
  img_out.resize(np);
  g.a.reserve(np + 1);
  g.b.reserve(np + 1);
  g.c.reserve(np + 1);
  g.d.reserve(np + 1);
  const int *keys = g.c.p;
  if (np > 0) {
    hipLaunchKernelGGL(k_seq, dim3(div_up(np, T)), dim3(T), 0, g.st, g.a.p, np);
    int bits = 1;
    while ((1 << bits) < nc) bits++;
    bool alt = false;
    // Keys: a copy of the targets (sorted stably), values: emission index.
    EXA_HIP(hipMemcpyAsync(g.c.p, tgt.p, np * sizeof(int), hipMemcpyDeviceToDevice, g.st));
    sort_pairs_int_int(g.c.p, g.d.p, g.a.p, g.b.p, (int)np, bits, alt, g.st, g.tmp);
    const int *order = alt ? g.b.p : g.a.p;
    keys = alt ? g.d.p : g.c.p;
    hipLaunchKernelGGL(k_permute_pairs, dim3(div_up(np, T)), dim3(T), 0, g.st, order, np, src.p, img.p, src_out.p,
                       img_out.p);
  }
  hipLaunchKernelGGL(k_csr_offsets, dim3(div_up(nc + 1, T)), dim3(T), 0, g.st, keys, np, nc, off.p);
  // Compact non-empty targets.
  g.e.reserve(2 * (size_t)nc + 2);
  int *flag = g.e.p, *scan = g.e.p + nc + 1;
  hipLaunchKernelGGL(k_flag_nonzero, dim3(div_up(nc + 1, T)), dim3(T), 0, g.st, off.p, nc, flag);
  exclusive_scan_int(flag, scan, nc + 1, g.st, g.tmp);
  tlist.resize(nc);
  hipLaunchKernelGGL(k_compact, dim3(div_up(nc, T)), dim3(T), 0, g.st, flag, scan, nc, tlist.p);
  EXA_HIP(hipMemcpyAsync(ntl_host, scan + nc, sizeof(int), hipMemcpyDeviceToHost, g.st));
  if (ntl_dev) *ntl_dev = scan + nc;
}

// Levels of device-driven passes, enqueued without waiting; returns false
// (nothing usable emitted) when an output did not fit its buffer or the
// traversal did not end within kMaxLevels levels.  The frontier starts in fa.
static bool traverse_on_device(Handle &h, const CellView &cv, int nf0, int mutual, int64_t &nm, int64_t &np) {
  constexpr int kMaxLevels = 256;
  // (fa and fb were sized by the caller before the frontier was written.)
  h.ia.reserve(std::max(h.trav_fcap, nf0) + 1);
  h.ib.reserve(4 * (size_t)kTravGrid);
  h.trav_state.reserve(sizeof(TravState));
  TravState *st = reinterpret_cast<TravState *>(h.trav_state.p);
  const long long cap_m = (long long)std::min(std::min(h.raw_mi.cap, h.raw_mj.cap), h.raw_mimg.cap) - 1;
  const long long cap_p = (long long)std::min(std::min(h.raw_pi.cap, h.raw_pj.cap), h.raw_pimg.cap) - 1;
  const int cap_fr = (int)std::min(std::min(h.fa.cap, h.fb.cap), h.ia.cap) - 1;  // frontier and action slots
  int4 *blk = reinterpret_cast<int4 *>(h.ib.p);
  unsigned char *act = reinterpret_cast<unsigned char *>(h.ia.p);
  hipLaunchKernelGGL(k_trav_init, dim3(1), dim3(1), 0, h.stream, st, nf0);
  hipLaunchKernelGGL(k_trav_small, dim3(1), dim3(kSmallBlock), 0, h.stream, cv, h.fa.p, h.fb.p, st, h.cfg.cycle,
                     h.cfg.nspawn, mutual, act, h.raw_mi.p, h.raw_mj.p, h.raw_mimg.p, h.raw_pi.p, h.raw_pj.p,
                     h.raw_pimg.p, cap_m, cap_p, cap_fr);
  TravState hs{};
  int enq = 0;
  // The levels after the small ones: as many as the last traversal had (its
  // last level leaves an empty frontier), then more in batches if needed.
  int batch = h.trav_levels > 0 ? h.trav_levels : 8;
  while (true) {
    for (int l = 0; l < batch; l++) {
      hipLaunchKernelGGL((k_classify_dev<kClassifyBlock, 1>), dim3(kTravGrid), dim3(kClassifyBlock), 0, h.stream, cv,
                         h.fa.p, h.fb.p, st, h.cfg.cycle, h.cfg.nspawn, mutual, act, blk);
      hipLaunchKernelGGL((k_emit_dev<kEmitBlock, kEmitItems>), dim3(kTravGrid), dim3(kEmitBlock), 0, h.stream, cv,
                         h.fa.p, h.fb.p, st, mutual, act, blk, cap_m, cap_p, cap_fr, h.raw_mi.p, h.raw_mj.p,
                         h.raw_mimg.p, h.raw_pi.p, h.raw_pj.p, h.raw_pimg.p);
    }
    enq += batch;
    enqueue_multipoles(h);  // the second stream's work, while these levels run
    EXA_HIP(hipMemcpyAsync(&hs, st, sizeof(TravState), hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    if (hs.overflow) return false;
    if (hs.nf == 0) break;
    if (enq >= kMaxLevels) return false;
    batch = 4;
  }
  nm = hs.nm;
  np = hs.np;
  h.trav_levels = hs.levels - hs.levels_small;
  return true;
}

void run_lists(Handle &h) {
  const int T = 256;
  if (h.ncells >= (1 << kImgShift))
    throw Error(EXAFMM_HIP_EUNSUP, "traversal: more than 2^27 cells (frontier pairs hold 27-bit cell indices)");
  CellView cv{h.cX.p, h.cChild.p, h.cNChild.p, h.cNBody.p};
  // Raw (emission-order) lists, kept on the handle across evaluations.
  DBuf<int> &mi = h.raw_mi, &mj = h.raw_mj, &pi = h.raw_pi, &pj = h.raw_pj;
  DBuf<unsigned char> &mimg = h.raw_mimg, &pimg = h.raw_pimg;
  int64_t nm = 0, np = 0;
  int nf = h.cfg.images == 0 ? 1 : 27;
  // Periodic images are traversed non-mutually (traversal.h:250-259).
  const int mutual = (h.cfg.mutual && h.cfg.images == 0) ? 1 : 0;
  if (h.nsys > 1) nf = h.lvl_off[1] - h.lvl_off[0];  // the roots: level 0 of the forest
  // Frontier buffers sized before the first frontier is written (a reserve
  // that grows a buffer drops its contents).
  h.fa.reserve(std::max(nf + 64, h.trav_fcap));
  if (h.trav_fcap > 0) h.fb.reserve(std::max(nf + 64, h.trav_fcap));
  auto init_frontier = [&]() {
    if (h.nsys > 1)
      hipLaunchKernelGGL(k_init_frontier_roots, dim3(div_up(nf, 256)), dim3(256), 0, h.stream, h.fa.p,
                         h.lvl_cells.p + h.lvl_off[0], nf);
    else
      hipLaunchKernelGGL(k_init_frontier, dim3(1), dim3(64), 0, h.stream, h.fa.p, h.cfg.images);
  };
  init_frontier();
  // Start the raw lists at a capacity that covers uniform trees without
  // regrowth (about 180 M2L and 65 P2P sources per cell at theta = 0.4).
  grow_keep(h, mi, (size_t)h.ncells * 128 + 1024, 0);
  grow_keep(h, mj, (size_t)h.ncells * 128 + 1024, 0);
  grow_keep(h, mimg, (size_t)h.ncells * 128 + 1024, 0);
  grow_keep(h, pi, (size_t)h.ncells * 48 + 1024, 0);
  grow_keep(h, pj, (size_t)h.ncells * 48 + 1024, 0);
  grow_keep(h, pimg, (size_t)h.ncells * 48 + 1024, 0);
  h.cnt.reserve(4);
  const int nf0 = nf;
  bool done = false;
  if (h.trav_fcap > 0) {
    done = traverse_on_device(h, cv, nf0, mutual, nm, np);
    if (!done) init_frontier();  // the per-level passes start over
  }
  if (!done) h.trav_levels = 0;
  if (!done) enqueue_multipoles(h);
  int max_nf = nf0;
  while (!done && nf > 0) {
    const int nb = div_up(nf, kTravBlock);
    h.ia.reserve(nf + 1);  // action
    h.ib.reserve(4 * ((size_t)nb + 1));  // block totals (int4)
    h.ic.reserve(4 * ((size_t)nb + 1));  // block offsets (int4)
    unsigned char *act = reinterpret_cast<unsigned char *>(h.ia.p);
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
    max_nf = std::max(max_nf, nf);
    h.trav_levels++;
    std::swap(h.fa.p, h.fb.p);
    std::swap(h.fa.cap, h.fb.cap);
  }
  if (!done) h.trav_fcap = max_nf + max_nf / 4 + 64;  // the device-driven passes from the next call on
  h.nm2l = nm;
  h.np2p = np;
  // The lattice M2L works on the raw (emission-order) M2L list; the M2L
  // list is grouped by target only for the other M2L paths (and on request,
  // `ensure_m2l_grouped`).  With both grouped, the two groupings run at once:
  // M2L on the main stream, P2P on the second stream (its own scratch).
  // The P2P grouping's target count stays on the device (a resident P2P
  // grid reads it there; resolve_p2p_count reads it back when the host needs
  // it), so the host queues P2P and the M2L set-up without waiting for it.
  h.m2l_grouped = !lattice_applies(h);
  EXA_HIP(hipEventRecord(h.ev_lists, h.stream));
  EXA_HIP(hipStreamWaitEvent(h.stream2, h.ev_lists, 0));
  if (h.m2l_grouped)
    group_by_target(h, GroupCtx{h.ia, h.ib, h.ic, h.id, h.ie, h.tmp, h.stream}, nm, mi, mj, mimg, h.m2l_off,
                    h.m2l_src, h.m2l_img, h.m2l_tgt, h.hpin + 8);
  group_by_target(h, GroupCtx{h.g2a, h.g2b, h.g2c, h.g2d, h.g2e, h.tmp2, h.stream2}, np, pi, pj, pimg, h.p2p_off,
                  h.p2p_src, h.p2p_img, h.p2p_tgt, h.hpin + 9, &h.p2p_cnt_dev);
  EXA_HIP(hipEventRecord(h.ev_p2pcnt, h.stream2));
  h.np2p_tgt = -1;
  h.nm2l_tgt = 0;
  if (h.m2l_grouped) {
    EXA_HIP(hipStreamSynchronize(h.stream));
    h.nm2l_tgt = h.hpin[8];
  }
  EXA_HIP(hipGetLastError());
}

void ensure_m2l_grouped(Handle &h) {
  if (h.m2l_grouped) return;
  group_by_target(h, GroupCtx{h.ia, h.ib, h.ic, h.id, h.ie, h.tmp, h.stream}, h.nm2l, h.raw_mi, h.raw_mj, h.raw_mimg,
                  h.m2l_off, h.m2l_src, h.m2l_img, h.m2l_tgt, h.hpin + 8);
  EXA_HIP(hipStreamSynchronize(h.stream));
  h.nm2l_tgt = h.hpin[8];
  h.m2l_grouped = true;
}

void group_pairs(Handle &h, int64_t np, DBuf<int> &tgt, DBuf<int> &src, DBuf<unsigned char> &img, DBuf<int> &off,
                 DBuf<int> &src_out, DBuf<unsigned char> &img_out, DBuf<int> &tlist, int &ntl) {
  group_by_target(h, GroupCtx{h.ia, h.ib, h.ic, h.id, h.ie, h.tmp, h.stream}, np, tgt, src, img, off, src_out, img_out,
                  tlist, h.hpin + 10);
  EXA_HIP(hipStreamSynchronize(h.stream));
  ntl = h.hpin[10];
}

}  // namespace exafmm
