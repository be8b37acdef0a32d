// Adaptive octree (reference: include/build_tree_tbb.h:36-374), built
// level-synchronously on the device.  Compiled with -ffp-contract=off.
//
// Reference semantics reproduced:
//   * a node with more than ncrit bodies is split; its bodies are partitioned
//     stably by octant (x > Xc) + ((y > Yc) << 1) + ((z > Zc) << 2)
//     (CountBodies/MoveBodies, :52-116, a stable out-of-place scatter);
//   * child centres are X +- R0 / 2^(level+1) (:199-203), empty octants get
//     no node;
//   * cells are numbered as Nodes2cells does (:250-294): root 0, each node's
//     children form one block at the running cursor, and child i's subtree
//     block follows the blocks of children 0..i-1;
//   * R = R0 / 2^level, ICELL = level offset + Morton interleave of the
//     centre (:240-249; computed in 64 bits, the reference's int overflows
//     beyond level 10).
//
// Per level: the bodies of the nodes still being split are compacted, keyed
// by (node rank, octant), sorted stably and scattered back into the same
// positions, which is exactly a stable per-node octant partition; bodies of
// finished leaves never move.
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace exafmm {

namespace {

struct NodeLevel {
  // Nodes of one level, in the order they were created ((parent, octant)).
  int count = 0;
  int first = 0;  // index of the level's first node in the global node arrays
};

// Active bodies in compacted order: position a belongs to active node r,
// the last with aoff[r] <= a (aoff: exclusive scan of the active nodes' body
// counts; a node's bodies are the contiguous range starting at astart[r]).
// Key (node rank << 3 | octant) for the stable per-node octant partition.
__global__ void k_active_keys(int na, const int *__restrict__ aoff, int nact, const int *__restrict__ astart,
                              const double4 *__restrict__ nX, const double4 *__restrict__ body, int *__restrict__ apos,
                              int *__restrict__ key, int *__restrict__ val) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  int lo = 0, hi = nact;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (aoff[mid] <= a)
      lo = mid;
    else
      hi = mid;
  }
  const int b = astart[lo] + (a - aoff[lo]);
  const double4 x = body[b], c = nX[lo];
  const int oct = (x.x > c.x) + ((x.y > c.y) << 1) + ((x.z > c.z) << 2);
  apos[a] = b;
  key[a] = (lo << 3) | oct;  // node rank < 2^28
  val[a] = a;
}

// The active bodies in sorted order into scratch, then back into the active
// positions: a stable octant partition of every active node in place
// (bodies of finished leaves are not touched).
__global__ void k_gather_active(const int *__restrict__ apos, const int *__restrict__ order, int na,
                                const double4 *__restrict__ bin, const int *__restrict__ iin, double4 *__restrict__ bt,
                                int *__restrict__ it) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  const int src = apos[order[a]];
  bt[a] = bin[src];
  it[a] = iin[src];
}

__global__ void k_put_active(const int *__restrict__ apos, int na, const double4 *__restrict__ bt,
                             const int *__restrict__ it, double4 *__restrict__ bout, int *__restrict__ iout) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  const int dst = apos[a];
  bout[dst] = bt[a];
  iout[dst] = it[a];
}

// Child heads: position a starts a child when its (node, octant) key differs
// from position a-1's.
__global__ void k_child_heads(const int *__restrict__ skey, int na, int *__restrict__ head) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a > na) return;
  head[a] = a < na ? ((a == 0) || (skey[a] != skey[a - 1])) : 0;
}

__global__ void k_child_positions(const int *__restrict__ head, const int *__restrict__ hscan, int na,
                                  int *__restrict__ hpos) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < na && head[a]) hpos[hscan[a]] = a;
}

// Reset the child bookkeeping of the active (parent) nodes.
__global__ void k_reset_parents(const int *__restrict__ agid, int nact, int *__restrict__ first_child,
                                int *__restrict__ nchild) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nact) return;
  first_child[agid[r]] = 0x7fffffff;
  nchild[agid[r]] = 0;
}

struct ChildArgs {
  const int *skey;
  const int *apos;
  const int *hpos;        // sorted position of each child's first body
  const int *nch;         // children created at this level (device word)
  int na;
  const double4 *nX;      // centres of the active (parent) nodes
  const int *ngid;        // global node id of each active node
  double r;               // R0 / 2^(level+1)
  int out_first;          // global node id of the first child
  int ncrit;
  // outputs (global node arrays)
  double4 *X;
  int *start, *count, *parent, *leaf;
  int *first_child, *nchild;  // of the parents
};

__global__ void k_children(ChildArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nch = *a.nch;
  if (c >= nch) return;
  const int i = a.hpos[c];
  const int e = c + 1 < nch ? a.hpos[c + 1] : a.na;
  const int k = a.skey[i];
  const int pr = k >> 3, oct = k & 7;
  const double4 P = a.nX[pr];
  double4 X;
  X.x = P.x + a.r * ((oct & 1) * 2 - 1);
  X.y = P.y + a.r * (((oct >> 1) & 1) * 2 - 1);
  X.z = P.z + a.r * (((oct >> 2) & 1) * 2 - 1);
  X.w = 0;
  const int g = a.out_first + c;
  a.X[g] = X;
  a.start[g] = a.apos[i];
  a.count[g] = e - i;
  a.parent[g] = a.ngid[pr];
  a.leaf[g] = (e - i) <= a.ncrit;
  a.first_child[g] = 0x7fffffff;
  a.nchild[g] = 0;
  // Children of one parent are consecutive (sorted by (parent, octant)), so
  // the parent's block starts at the minimum child id; counts are integers.
  atomicMin(&a.first_child[a.ngid[pr]], g);
  atomicAdd(&a.nchild[a.ngid[pr]], 1);
}

__global__ void k_identity(int *__restrict__ v, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) v[i] = i;
}

// Gather the next level's active nodes (children that are not leaves) and
// count their bodies into *nbody (zeroed beforehand).  *cnt: the level's
// children (device word).
__global__ void k_next_active(const int *__restrict__ leaf, int first, const int *__restrict__ cnt,
                              const int *__restrict__ scan, const double4 *__restrict__ X,
                              const int *__restrict__ start, const int *__restrict__ count, double4 *__restrict__ aX,
                              int *__restrict__ astart, int *__restrict__ acount, int *__restrict__ agid,
                              int *__restrict__ nbody) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *cnt || leaf[first + i]) return;
  const int a = scan[i];
  aX[a] = X[first + i];
  astart[a] = start[first + i];
  acount[a] = count[first + i];
  agid[a] = first + i;
  atomicAdd(nbody, count[first + i]);
}

// out[i] = child i is not a leaf, for i < *cnt; 0 up to out[bound].
__global__ void k_not_leaf(const int *__restrict__ leaf, int first, const int *__restrict__ cnt, int bound,
                           int *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > bound) return;
  out[i] = i < *cnt ? !leaf[first + i] : 0;
}

// The level's three counts into w[0..2] for one read-back: children, next
// active nodes, their bodies.
__global__ void k_level_counts(const int *__restrict__ nch, const int *__restrict__ nnext,
                               const int *__restrict__ nbody, int *__restrict__ w) {
  w[0] = *nch;
  w[1] = *nnext;
  w[2] = *nbody;
}

// Subtree sizes, one level at a time from the deepest.
__global__ void k_nnode(int first, int cnt, const int *__restrict__ first_child, const int *__restrict__ nchild,
                        int *__restrict__ nnode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int g = first + i;
  int s = 1;
  for (int c = 0; c < nchild[g]; c++) s += nnode[first_child[g] + c];
  nnode[g] = s;
}

// Nodes2cells numbering, one level at a time from the root.
__global__ void k_number(int first, int cnt, const int *__restrict__ first_child, const int *__restrict__ nchild,
                         const int *__restrict__ nnode, int *__restrict__ cellid, int *__restrict__ cursor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const int g = first + i;
  const int nc = nchild[g];
  if (nc == 0) return;
  const int cur = cursor[g];
  int sub = cur + nc;
  for (int c = 0; c < nc; c++) {
    const int ch = first_child[g] + c;
    cellid[ch] = cur + c;
    cursor[ch] = sub;
    sub += nnode[ch] - 1;
  }
}

__global__ void k_emit_cells(int total, const int *__restrict__ cellid, const int *__restrict__ parent,
                             const int *__restrict__ first_child, const int *__restrict__ nchild,
                             const int *__restrict__ cursor, const int *__restrict__ start,
                             const int *__restrict__ count, const double4 *__restrict__ X,
                             const int *__restrict__ level, double R0, double x0, double y0, double z0,
                             double4 *__restrict__ cX, double *__restrict__ cRb, int *__restrict__ cParent,
                             int *__restrict__ cChild, int *__restrict__ cNChild, int *__restrict__ cBody,
                             int *__restrict__ cNBody, uint64_t *__restrict__ cKey, int *__restrict__ cLevel,
                             int *__restrict__ lvl_cells) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int c = cellid[g];
  if (c < 0 || c >= total) return;  // caught by k_validate
  const int lev = level[g];
  const double R = R0 / (1 << lev);
  double4 Xc = X[g];
  Xc.w = R;
  cX[c] = Xc;
  cRb[c] = R;
  cParent[c] = parent[g] >= 0 ? cellid[parent[g]] : 0;
  cChild[c] = nchild[g] > 0 ? cursor[g] : 0;
  cNChild[c] = nchild[g];
  cBody[c] = start[g];
  cNBody[c] = count[g];
  cLevel[c] = lev;
  // getKey (:240-249), 64-bit.
  const double diameter = 2 * R;
  int64_t ix = (int64_t)(int)((Xc.x - x0) / diameter);
  int64_t iy = (int64_t)(int)((Xc.y - y0) / diameter);
  int64_t iz = (int64_t)(int)((Xc.z - z0) / diameter);
  uint64_t index = ((1ULL << (3 * lev)) - 1) / 7;
  for (int l = 0; l < lev; l++) {
    index += (uint64_t)(ix & 1) << (3 * l);
    index += (uint64_t)(iy & 1) << (3 * l + 1);
    index += (uint64_t)(iz & 1) << (3 * l + 2);
    ix >>= 1;
    iy >>= 1;
    iz >>= 1;
  }
  cKey[c] = index;
  lvl_cells[g] = c;
}

__global__ void k_root(double4 *X, int *start, int *count, int *parent, int *leaf, int *level, int *first_child,
                       int *nchild, double x, double y, double z, int n, int ncrit, double4 *aX, int *astart,
                       int *acount, int *agid) {
  X[0] = make_double4(x, y, z, 0);
  first_child[0] = 0x7fffffff;
  nchild[0] = 0;
  start[0] = 0;
  count[0] = n;
  parent[0] = -1;
  leaf[0] = n <= ncrit;
  level[0] = 0;
  aX[0] = X[0];
  astart[0] = 0;
  acount[0] = n;
  agid[0] = 0;
}

// Structural check of a finished tree: every cell written, parent and child
// ranges inside the cell array, body ranges inside the body array, and the
// cell's children pointing back to it.  Violations are counted.
__global__ void k_validate(int nc, int n, const int *__restrict__ cParent, const int *__restrict__ cChild,
                           const int *__restrict__ cNChild, const int *__restrict__ cBody,
                           const int *__restrict__ cNBody, int *__restrict__ bad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int p = cParent[c], ch = cChild[c], nch = cNChild[c], b = cBody[c], nb = cNBody[c];
  bool ok = p >= 0 && p < nc && nch >= 0 && nch <= 8 && b >= 0 && nb >= 0 && b + nb <= n;
  if (ok && nch > 0) {
    ok = ch > c && ch + nch <= nc;
    for (int k = 0; ok && k < nch; k++) ok = cParent[ch + k] == c;
  }
  if (!ok) atomicAdd(bad, 1);
}

__global__ void k_fill_level(int *__restrict__ level, int first, const int *__restrict__ cnt, int lev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *cnt) level[first + i] = lev;
}

}  // namespace

void run_build_adaptive(Handle &h) {
  const int n = h.n, ncrit = h.cfg.ncrit;
  const int T = 256;
  sync_bounds(h);
  // bounds2box (build_tree_tbb.h:298-310).
  double Xc[3], R0 = 0;
  for (int d = 0; d < 3; d++) Xc[d] = (h.xmax[d] + h.xmin[d]) / 2;
  for (int d = 0; d < 3; d++) {
    R0 = std::max(Xc[d] - h.xmin[d], R0);
    R0 = std::max(h.xmax[d] - Xc[d], R0);
  }
  R0 *= 1.00001;
  if (!(R0 > 0)) throw Error(EXAFMM_HIP_EINVAL, "build_tree: degenerate bounds (box radius 0)");

  // Global node arrays (capacity grows per level, contents kept).
  size_t cap = 2 * (size_t)(n / std::max(ncrit, 1)) * 8 + 1024;
  DBuf<double4> &nX = h.at_X;
  DBuf<int> &nstart = h.at_start, &ncount = h.at_count, &nparent = h.at_parent, &nleaf = h.at_leaf,
            &nlevel = h.at_level, &nfirst = h.at_first, &nnch = h.at_nchild;
  // The node arrays persist in the handle and grow together: what they hold
  // from an earlier build is reused (no allocation, no synchronisation).
  size_t ncap = std::min({nX.cap, nstart.cap, ncount.cap, nparent.cap, nleaf.cap, nlevel.cap, nfirst.cap, nnch.cap});
  auto grow = [&](auto &b, size_t want, size_t keep) {
    using Tt = typename std::remove_reference<decltype(*b.p)>::type;
    DBuf<Tt> nb;
    nb.reserve(want);
    if (b.p && keep) EXA_HIP(hipMemcpyAsync(nb.p, b.p, keep * sizeof(Tt), hipMemcpyDeviceToDevice, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    b.release();
    b.p = nb.p;
    b.cap = nb.cap;
    nb.p = nullptr;
    nb.cap = 0;
  };
  auto reserve_nodes = [&](size_t want, size_t keep) {
    if (want <= ncap) return;
    const size_t w = want * 2;
    grow(nX, w, keep);
    grow(nstart, w, keep);
    grow(ncount, w, keep);
    grow(nparent, w, keep);
    grow(nleaf, w, keep);
    grow(nlevel, w, keep);
    grow(nfirst, w, keep);
    grow(nnch, w, keep);
    ncap = w;
  };
  reserve_nodes(cap, 0);

  // Bodies start in caller order.
  h.body4.resize(n);
  h.bidx.resize(n);
  h.trg4.resize(n);
  DBuf<double4> &b2 = h.at_body2;
  DBuf<int> &i2 = h.at_idx2;
  b2.reserve(n);
  i2.reserve(n);
  EXA_HIP(hipMemcpyAsync(h.body4.p, h.in4.p, n * sizeof(double4), hipMemcpyDeviceToDevice, h.stream));
  hipLaunchKernelGGL(k_identity, dim3(div_up(n, T)), dim3(T), 0, h.stream, h.bidx.p, n);

  // Active-node scratch.
  DBuf<double4> &aX = h.at_aX;
  DBuf<int> &astart = h.at_astart, &acount = h.at_acount, &agid = h.at_agid;
  aX.reserve(cap);
  astart.reserve(cap);
  acount.reserve(cap);
  agid.reserve(cap);
  hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, h.stream, nX.p, nstart.p, ncount.p, nparent.p, nleaf.p, nlevel.p,
                     nfirst.p, nnch.p, Xc[0], Xc[1], Xc[2], n, ncrit, aX.p, astart.p, acount.p, agid.p);
  std::vector<NodeLevel> levels;
  levels.push_back({1, 0});
  int total = 1;
  int nact = n > ncrit ? 1 : 0;
  int na = nact ? n : 0;  // bodies of the active nodes
  int lev = 0;
  // Device words: [0] bodies of the next active set, [1..3] the level's
  // counts gathered for one read-back per level (pinned words 16..18).
  h.cnt.reserve(4);
  int *const hcnt = h.hpin + 16;
  while (nact > 0) {
    if (lev >= 20)
      throw Error(EXAFMM_HIP_EUNSUP, "adaptive tree deeper than 20 levels (more than ncrit coincident bodies?)");
    // 1. The bodies of active nodes (na of them, known from the previous
    // level's read-back) in compacted order, keyed by (node rank, octant).
    h.ic.reserve((size_t)nact + 1);
    exclusive_scan_int(h, acount.p, h.ic.p, nact);  // aoff
    h.ie.reserve(na + 1);  // apos
    h.k64a.reserve(na + 1);
    h.ia.reserve(na + 1);
    h.ib.reserve(na + 1);
    // Keys as 32-bit ints: the two key buffers in k64a (a radix sort over
    // 64-bit keys takes several times as many passes).
    if (nact >= (1 << 28)) throw Error(EXAFMM_HIP_EUNSUP, "adaptive tree: too many nodes on one level");
    int *k32 = reinterpret_cast<int *>(h.k64a.p), *k32alt = k32 + (na + 1);
    hipLaunchKernelGGL(k_active_keys, dim3(div_up(na, T)), dim3(T), 0, h.stream, na, h.ic.p, nact, astart.p, aX.p,
                       h.body4.p, h.ie.p, k32, h.ia.p);
    // 2. Stable sort by (node rank, octant).
    int bits = 3;
    while ((1LL << (bits - 3)) < nact) bits++;
    bool alt = false;
    sort_pairs_int_int(h, k32, k32alt, h.ia.p, h.ib.p, na, bits, alt);
    const int *skey = alt ? k32alt : k32;
    const int *order = alt ? h.ib.p : h.ia.p;
    // 3. Move the active bodies into sorted order, in place.
    hipLaunchKernelGGL(k_gather_active, dim3(div_up(na, T)), dim3(T), 0, h.stream, h.ie.p, order, na, h.body4.p,
                       h.bidx.p, b2.p, i2.p);
    hipLaunchKernelGGL(k_put_active, dim3(div_up(na, T)), dim3(T), 0, h.stream, h.ie.p, na, b2.p, i2.p, h.body4.p,
                       h.bidx.p);
    // 4. Children: one per (node, octant) run; their count stays on the
    // device (hscan[na]) and launches are sized by its bound.
    h.id.reserve(3 * ((size_t)na + 1));
    int *head = h.id.p, *hscan = h.id.p + na + 1, *hpos = h.id.p + 2 * ((size_t)na + 1);
    hipLaunchKernelGGL(k_child_heads, dim3(div_up(na + 1, T)), dim3(T), 0, h.stream, skey, na, head);
    exclusive_scan_int(h, head, hscan, na + 1);
    const int *nch_dev = hscan + na;
    const int bound = (int)std::min<int64_t>(8LL * nact, na);  // children <= 8 per node, >= 1 body each
    hipLaunchKernelGGL(k_child_positions, dim3(div_up(na, T)), dim3(T), 0, h.stream, head, hscan, na, hpos);
    reserve_nodes((size_t)total + bound + 1, (size_t)total);
    hipLaunchKernelGGL(k_reset_parents, dim3(div_up(nact, T)), dim3(T), 0, h.stream, agid.p, nact, nfirst.p, nnch.p);
    ChildArgs ca;
    ca.skey = skey;
    ca.apos = h.ie.p;
    ca.hpos = hpos;
    ca.nch = nch_dev;
    ca.na = na;
    ca.nX = aX.p;
    ca.ngid = agid.p;
    ca.r = R0 / (1 << (lev + 1));
    ca.out_first = total;
    ca.ncrit = ncrit;
    ca.X = nX.p;
    ca.start = nstart.p;
    ca.count = ncount.p;
    ca.parent = nparent.p;
    ca.leaf = nleaf.p;
    ca.first_child = nfirst.p;
    ca.nchild = nnch.p;
    hipLaunchKernelGGL(k_children, dim3(div_up(bound, T)), dim3(T), 0, h.stream, ca);
    hipLaunchKernelGGL(k_fill_level, dim3(div_up(bound, T)), dim3(T), 0, h.stream, nlevel.p, total, nch_dev,
                       lev + 1);
    // 5. Next active set: the children that are not leaves, and their bodies.
    DBuf<int> &nlb = h.at_nl;
    nlb.reserve(2 * ((size_t)bound + 1));
    int *nl = nlb.p, *nls = nlb.p + bound + 1;
    hipLaunchKernelGGL(k_not_leaf, dim3(div_up(bound + 1, T)), dim3(T), 0, h.stream, nleaf.p, total, nch_dev, bound,
                       nl);
    exclusive_scan_int(h, nl, nls, bound + 1);
    if ((size_t)bound + 1 > aX.cap) {  // the current active set is read up to here: let it finish first
      EXA_HIP(hipStreamSynchronize(h.stream));
      aX.reserve(bound + 1);
      astart.reserve(bound + 1);
      acount.reserve(bound + 1);
      agid.reserve(bound + 1);
    }
    zero_device(h.cnt.p, 2 * sizeof(int), h.stream);
    hipLaunchKernelGGL(k_next_active, dim3(div_up(bound, T)), dim3(T), 0, h.stream, nleaf.p, total, nch_dev, nls,
                       nX.p, nstart.p, ncount.p, aX.p, astart.p, acount.p, agid.p, h.cnt.p);
    // One read-back per level: children, next active nodes, their bodies.
    hipLaunchKernelGGL(k_level_counts, dim3(1), dim3(1), 0, h.stream, nch_dev, nls + bound, h.cnt.p, h.cnt.p + 1);
    EXA_HIP(hipMemcpyAsync(hcnt, h.cnt.p + 1, 3 * sizeof(int), hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    const int nchildren = hcnt[0];
    levels.push_back({nchildren, total});
    total += nchildren;
    nact = hcnt[1];
    na = hcnt[2];
    lev++;
  }
  // Subtree sizes (bottom-up), then cell numbering (top-down).
  DBuf<int> &nnode = h.at_nnode, &cellid = h.at_cellid, &cursor = h.at_cursor;
  nnode.reserve(total);
  cellid.reserve(total);
  cursor.reserve(total);
  for (int l = (int)levels.size() - 1; l >= 0; l--)
    hipLaunchKernelGGL(k_nnode, dim3(div_up(levels[l].count, T)), dim3(T), 0, h.stream, levels[l].first,
                       levels[l].count, nfirst.p, nnch.p, nnode.p);
  const int zero = 0, one = 1;
  EXA_HIP(hipMemcpyAsync(cellid.p, &zero, sizeof(int), hipMemcpyHostToDevice, h.stream));
  EXA_HIP(hipMemcpyAsync(cursor.p, &one, sizeof(int), hipMemcpyHostToDevice, h.stream));
  for (size_t l = 0; l < levels.size(); l++)
    hipLaunchKernelGGL(k_number, dim3(div_up(levels[l].count, T)), dim3(T), 0, h.stream, levels[l].first,
                       levels[l].count, nfirst.p, nnch.p, nnode.p, cellid.p, cursor.p);
  h.ncells = total;
  h.nlevels = (int)levels.size();
  h.cX.resize(total);
  h.cRb.resize(total);
  h.cParent.resize(total);
  h.cChild.resize(total);
  h.cNChild.resize(total);
  h.cBody.resize(total);
  h.cNBody.resize(total);
  h.cKey.resize(total);
  h.cLevel.resize(total);
  h.lvl_cells.resize(total);
  EXA_HIP(hipMemsetAsync(h.cNChild.p, 0xff, total * sizeof(int), h.stream));  // -1: not written
  const double Xmin[3] = {Xc[0] - R0, Xc[1] - R0, Xc[2] - R0};
  hipLaunchKernelGGL(k_emit_cells, dim3(div_up(total, T)), dim3(T), 0, h.stream, total, cellid.p, nparent.p, nfirst.p,
                     nnch.p, cursor.p, nstart.p, ncount.p, nX.p, nlevel.p, R0, Xmin[0], Xmin[1], Xmin[2], h.cX.p,
                     h.cRb.p, h.cParent.p, h.cChild.p, h.cNChild.p, h.cBody.p, h.cNBody.p, h.cKey.p, h.cLevel.p,
                     h.lvl_cells.p);
  h.lvl_off.assign(levels.size() + 1, 0);
  for (size_t l = 0; l < levels.size(); l++) h.lvl_off[l + 1] = h.lvl_off[l] + levels[l].count;
  EXA_HIP(hipGetLastError());
  int bad = 0;
  EXA_HIP(hipMemsetAsync(h.cnt.p, 0, sizeof(int), h.stream));
  hipLaunchKernelGGL(k_validate, dim3(div_up(total, T)), dim3(T), 0, h.stream, total, n, h.cParent.p, h.cChild.p,
                     h.cNChild.p, h.cBody.p, h.cNBody.p, h.cnt.p);
  EXA_HIP(hipMemcpyAsync(&bad, h.cnt.p, sizeof(int), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  if (bad) throw Error(EXAFMM_HIP_EDEVICE, "adaptive tree failed its structural check (" + std::to_string(bad) + " cells)");
}

}  // namespace exafmm
