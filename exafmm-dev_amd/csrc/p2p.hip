// P2P over the CSR interaction list: a 64-lane workgroup works on one target
// leaf at a time, taking the next leaf from a work counter when done (the
// grid is one wave per leaf, or a resident grid of a few waves per CU so that
// the lattice M2L kernels of the other stream fit beside it on every CU).
//
// Reference: kernels/LaplaceP2PCPU.cxx:141-169 (pair) and :172-331 (self).
// Per pair: dX = x_i - x_j - X_periodic, R2 = |dX|^2, skipped when R2 == 0;
// pot += q_j / r, grad -= q_j dX / r^3.  The reference multiplies by q_i
// here and divides it out in L2P/normalize (LaplaceSphericalCPU.cxx:235,
// traversal.h:333-337); the GPU accumulates the normalised quantities
// directly.
//
// Lane layout: a leaf's targets (64 at a time) are split into two passes:
// pass A takes TA = the largest power of two <= n (at least 8) targets, each
// replicated over GA = 64 / TA lane groups; pass B takes the remaining r
// targets over GB = 64 / r groups (any whole number), lane g r + i holding
// target i of group g, so at most r - 1 lanes idle.  Group g takes every
// G-th source body.  Both passes run over every staged source tile, so 37
// targets cost 32 + 5 lanes' worth of work (64 / 12 groups) instead of 64.
// (The mixed kernel keeps pass B at a power of two, at least 8 targets.)
// Source bodies are staged through LDS (consecutive list entries, each
// shifted by its periodic image); the G partial sums of a target are
// combined in a fixed order at the end (deterministic, no atomics).
#include <algorithm>
#include <type_traits>

#include "internal.h"

namespace exafmm {

__device__ inline void image_shift(int img, double cycle, double &px, double &py, double &pz) {
  px = (img / 9 - 1) * cycle;
  py = ((img / 3) % 3 - 1) * cycle;
  pz = (img % 3 - 1) * cycle;
}

constexpr int kP2PStage = 256;  // source bodies staged in LDS per fill
// FP64 pairs in flight per lane, side by side: written one pair after the
// other, the compiler ran each pair's dependent chain (LDS read, wait, 18
// dependent operations) alone; two side by side measured fastest (four need
// 156 VGPRs, and three such waves per SIMD leave no room for the kernels of
// the other streams).
constexpr int kP2PUnroll = 2;
static_assert(kP2PStage % (8 * kP2PUnroll) == 0, "stages are padded to multiples of kP2PUnroll G, G <= 8");
constexpr double kFar = 1e100;  // position of the zero-charge padding bodies

// 1/sqrt(x) for finite x > 0: the hardware estimate (v_rsq_f64, measured
// on MI355X at most 2^-24.2 relative error over 2^22 arguments spanning 60
// binades) refined by one Newton step y (1 + e/2), e = 1 - x y^2, whose
// error is -1.5 times the square of the estimate's: below 4.4e-15 relative.
// (No special-case fix-ups: the caller selects 0 for x == 0.)
__device__ inline double rsqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, 0.5, y);
}

// Walker over a target's source list (CSR entries [e, e1)).  The
// descriptors of up to 64 consecutive entries -- first body, body count,
// image -- and the exclusive prefix of their counts are loaded in one
// parallel step.  A stage fill maps its slots to entries through a slot map:
// every entry that starts inside the stage marks its first slot, and a
// running maximum over the slots (four per lane, then across the wave) gives
// each slot its entry; the body loads of a lane's slots are then issued
// together.  The workgroup is one wave; every call is uniform across it.
struct SrcWalker {
  int e, sbase, dbase, dcount;
};

constexpr int kFillMax = 256;  // largest stage (slots per fill): four map slots per lane

struct SrcDesc {
  int b0[64];
  int pre[65];
  unsigned char img[64];
  int map[kFillMax];  // slot -> entry (index into b0 / pre / img)
};

// skip: a cell whose unshifted entry (image 13) is left out (-1: none).
// put(k, v, im) stores body v (x, y, z, q) with image im into slot k; NB
// body loads are in flight at a time (registers against latency).
template <int STAGE, int NB = 4, typename Put>
__device__ inline int fill_stage(SrcWalker &w, int e1, int lane, const int *__restrict__ src,
                                 const unsigned char *__restrict__ img, const int *__restrict__ cBody,
                                 const int *__restrict__ cNBody, const double4 *__restrict__ body, SrcDesc &sd,
                                 Put put, int skip = -1) {
  static_assert(STAGE <= kFillMax && STAGE % 64 == 0, "the slot map covers a stage, four slots per lane");
  static_assert((STAGE / 64) % NB == 0, "whole batches of loads");
  __syncthreads();  // the previous stage (and descriptor batch) are consumed
  if (w.e >= w.dbase + w.dcount) {
    w.dbase = w.e;
    w.dcount = min(64, e1 - w.e);
    int n = 0;
    if (lane < w.dcount) {
      const int s = src[w.e + lane];
      const unsigned char im = img[w.e + lane];
      sd.b0[lane] = cBody[s];
      n = (s == skip && im == 13) ? 0 : cNBody[s];
      sd.img[lane] = im;
    }
    int incl = n;
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    sd.pre[lane] = incl - n;
    if (lane == 63) sd.pre[64] = incl;
    __syncthreads();
  }
  const int first = w.e - w.dbase;
  const int q0 = sd.pre[first] + w.sbase;
  const int qend = sd.pre[64];
  const int cnt = min(STAGE, qend - q0);
  // Slot map: entry `first` from slot 0, each later entry with bodies from
  // its first slot in the stage; a running maximum fills the rest.
  int4 *m4 = reinterpret_cast<int4 *>(sd.map);
  m4[lane] = make_int4(lane == 0 ? first : -1, -1, -1, -1);
  {
    const int e = first + 1 + lane;
    if (e < w.dcount) {
      const int s0 = sd.pre[e] - q0;
      if (s0 < cnt && sd.pre[e + 1] > sd.pre[e]) sd.map[s0] = e;
    }
  }
  __syncthreads();
  int4 mv = m4[lane];
  mv.y = max(mv.y, mv.x);
  mv.z = max(mv.z, mv.y);
  mv.w = max(mv.w, mv.z);
  int run = mv.w;
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(run, d);
    if (lane >= d) run = max(run, t);
  }
  int before = __shfl_up(run, 1);
  if (lane == 0) before = -1;
  mv.x = max(mv.x, before);
  mv.y = max(mv.y, before);
  mv.z = max(mv.z, before);
  mv.w = max(mv.w, before);
  m4[lane] = mv;
  __syncthreads();
  // The lane's slots k = lane + 64 u, NB at a time: their body loads first
  // (slots past the stage's end load a body of the stage and store nothing),
  // then the stores.
#pragma unroll 1
  for (int u0 = 0; u0 < STAGE / 64; u0 += NB) {
    double4 v[NB];
    int ims[NB];
#pragma unroll
    for (int u = 0; u < NB; u++) {
      const int k = lane + 64 * (u0 + u);
      const int e = k < cnt ? sd.map[k] : first;
      const int b = k < cnt ? sd.b0[e] + (q0 + k - sd.pre[e]) : sd.b0[e] + (q0 - sd.pre[e]);
      ims[u] = sd.img[e];
      v[u] = body[b];
    }
#pragma unroll
    for (int u = 0; u < NB; u++)
      if (lane + 64 * (u0 + u) < cnt) put(lane + 64 * (u0 + u), v[u], ims[u]);
  }
  // Advance to the first unconsumed body.
  const int qn = q0 + cnt;
  if (qn >= qend) {
    w.e = w.dbase + w.dcount;
    w.sbase = 0;
  } else {
    int lo = first, hi = w.dcount;  // largest i in [first, dcount) with pre[i] <= qn
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sd.pre[mid] <= qn)
        lo = mid;
      else
        hi = mid;
    }
    w.e = w.dbase + lo;
    w.sbase = qn - sd.pre[lo];
  }
  __syncthreads();  // the stage is filled
  return cnt;
}

// The next target leaf of a one-wave workgroup: counter + grid size (the
// counter starts at 0 and the grid's own first leaves are 0 .. grid - 1).
__device__ inline int next_leaf(int *next, int lane) {
  int v = 0;
  if (lane == 0) v = atomicAdd(next, 1);
  return __shfl(v, 0) + (int)gridDim.x;
}

__global__ void k_zero_int(int *p) { *p = 0; }

// Mixed kernel: both passes of power-of-two sizes, TB = r rounded up to a
// power of two (at least 8), so that every group stride is a compile-time
// constant (a run-time stride cost the mixed kernel ~20 more VGPRs, which
// crowded out the lattice set-up kernels running beside it).
struct TargetPasses2 {
  int TA, TB, r;  // r = targets of pass B (0: no pass B)
  __device__ explicit TargetPasses2(int n) {
    TA = 8;
    while (TA * 2 <= n) TA *= 2;
    r = n > TA ? n - TA : 0;
    TB = 8;
    while (TB < r) TB *= 2;
  }
};

// The two target passes of a block of n <= 64 targets (see above).
struct TargetPasses {
  int TA, r, GB;
  __device__ explicit TargetPasses(int n) {
    TA = 8;
    while (TA * 2 <= n) TA *= 2;
    r = n > TA ? n - TA : 0;
    GB = r > 0 ? 64 / r : 1;
  }
};

__global__ void __launch_bounds__(64)
    k_p2p_f64(const int *__restrict__ tgt_list, int ntl_host, const int *__restrict__ ntl_dev,
              int *__restrict__ next, const int *__restrict__ off, const int *__restrict__ src,
              const unsigned char *__restrict__ img, const int *__restrict__ cBody,
              const int *__restrict__ cNBody, const double4 *__restrict__ body, double4 *__restrict__ trg,
              double cycle) {
  __shared__ double4 sb[kP2PStage];
  // The group partials are combined in the stage buffer (every stage is
  // consumed by then): 8.6 KB of LDS per wave, which leaves more of each CU
  // to the other streams' kernels beside twelve resident waves.
  double4 *const sacc = sb;
  __shared__ SrcDesc sd;
  const int lane = threadIdx.x;
  // Target leaves: blockIdx.x first, then the next unclaimed one of the
  // work counter (leaves differ in cost; the grid may be smaller than ntl).
  // ntl_dev (if set): the count as the list grouping left it on the device.
  const int ntl = ntl_dev ? *ntl_dev : ntl_host;
  for (int bi = blockIdx.x; bi < ntl; bi = next_leaf(next, lane)) {
    const int c = tgt_list[bi];
    const int tb0 = cBody[c], tn = cNBody[c];
    const int e0 = off[c], e1 = off[c + 1];
    // Sum the G group partials (lanes gg T + i) of the pass's targets in
    // group order.
    auto combine = [&](const double4 &acc, int T, int G, int n, int tfirst) {
      __syncthreads();
      sacc[lane] = acc;
      __syncthreads();
      if (lane < n) {
        double4 s = sacc[lane];
        for (int gg = 1; gg < G; gg++) {
          const double4 o = sacc[gg * T + lane];
          s.x += o.x;
          s.y += o.y;
          s.z += o.z;
          s.w += o.w;
        }
        double4 tv = trg[tb0 + tfirst + lane];
        tv.x += s.x;
        tv.y -= s.y;
        tv.z -= s.z;
        tv.w -= s.w;
        trg[tb0 + tfirst + lane] = tv;
      }
    };
    for (int tbase = 0; tbase < tn; tbase += 64) {
      const int nt = min(64, tn - tbase);
      const TargetPasses tp(nt);
      const int GA = 64 / tp.TA, GB = tp.GB;
      const int ga = lane / tp.TA, ia = lane % tp.TA;
      const int rr = tp.r > 0 ? tp.r : 1, gb = min(lane / rr, GB - 1), ib = lane % rr;
      const double4 xa = ia < nt ? body[tb0 + tbase + ia] : make_double4(0, 0, 0, 0);
      const double4 xb = tp.r > 0 ? body[tb0 + tbase + tp.TA + ib] : make_double4(0, 0, 0, 0);
      double4 acca = make_double4(0, 0, 0, 0), accb = make_double4(0, 0, 0, 0);
      // One pass over a staged tile with G (1, 2, 4 or 8) a compile-time
      // constant and a wave-uniform trip count: the source addresses of a
      // group are one base and immediate offsets.  MASK: the tile may hold a
      // body at the target's own position (its own leaf, or a periodic
      // image), so R2 == 0 is skipped by a select; other tiles have R2 > 0
      // for every pair (equal positions share a Morton key and hence a leaf).
      auto pass_g = [&](auto mask, auto gc, int ncp, const double4 &x, int g,
                        double4 &acc) __attribute__((always_inline)) {
        constexpr int G = decltype(gc)::value;
        constexpr bool MASK = decltype(mask)::value;
        const double4 *sg = sb + g;
        for (int j = 0; j < ncp; j += kP2PUnroll * G) {
          // The kP2PUnroll pairs side by side (loads, then each step of the
          // pair arithmetic for all of them), summed in source order.
          double4 xj[kP2PUnroll];
          double dx[kP2PUnroll], dy[kP2PUnroll], dz[kP2PUnroll], r2[kP2PUnroll], rs[kP2PUnroll];
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) xj[u] = sg[j + u * G];
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) {
            dx[u] = x.x - xj[u].x;
            dy[u] = x.y - xj[u].y;
            dz[u] = x.z - xj[u].z;
            r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
          }
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) rs[u] = rsqrt_pos(r2[u]);
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) {
            const double invr = MASK ? (r2[u] > 0 ? rs[u] : 0.0) : rs[u];
            const double qinvr = xj[u].w * invr;
            const double qinvr3 = qinvr * invr * invr;
            acc.x += qinvr;
            acc.y += dx[u] * qinvr3;
            acc.z += dy[u] * qinvr3;
            acc.w += dz[u] * qinvr3;
          }
        }
      };
      auto pass = [&](auto mask, int G, int ncp, const double4 &x, int g,
                      double4 &acc) __attribute__((always_inline)) {
        if (G == 1)
          pass_g(mask, std::integral_constant<int, 1>{}, ncp, x, g, acc);
        else if (G == 2)
          pass_g(mask, std::integral_constant<int, 2>{}, ncp, x, g, acc);
        else if (G == 4)
          pass_g(mask, std::integral_constant<int, 4>{}, ncp, x, g, acc);
        else
          pass_g(mask, std::integral_constant<int, 8>{}, ncp, x, g, acc);
      };
      // Pass B: G = GB groups at run time.  Whole steps of kP2PUnroll G
      // sources run as in pass_g; the last, partial step reads only slots
      // below ncp (a slot past it counts as a zero charge far away).
      auto pass_b = [&](auto mask, int ncp) __attribute__((always_inline)) {
        constexpr bool MASK = decltype(mask)::value;
        const int G = __builtin_amdgcn_readfirstlane(GB), step = kP2PUnroll * G;
        const int nfull = ncp / step * step;
        const double4 *sg = sb + gb;
        auto body_step = [&](int j, bool tail) __attribute__((always_inline)) {
          double4 xj[kP2PUnroll];
          double dx[kP2PUnroll], dy[kP2PUnroll], dz[kP2PUnroll], r2[kP2PUnroll], rs[kP2PUnroll];
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) {
            const int k = j + u * G;
            if (tail) {
              xj[u] = sg[min(k, ncp - 1 - gb)];
              if (gb + k >= ncp) xj[u] = make_double4(kFar, kFar, kFar, 0);
            } else {
              xj[u] = sg[k];
            }
          }
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) {
            dx[u] = xb.x - xj[u].x;
            dy[u] = xb.y - xj[u].y;
            dz[u] = xb.z - xj[u].z;
            r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
          }
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) rs[u] = rsqrt_pos(r2[u]);
#pragma unroll
          for (int u = 0; u < kP2PUnroll; u++) {
            const double invr = MASK ? (r2[u] > 0 ? rs[u] : 0.0) : rs[u];
            const double qinvr = xj[u].w * invr;
            const double qinvr3 = qinvr * invr * invr;
            accb.x += qinvr;
            accb.y += dx[u] * qinvr3;
            accb.z += dy[u] * qinvr3;
            accb.w += dz[u] * qinvr3;
          }
        };
        for (int j = 0; j < nfull; j += step) body_step(j, false);
        if (nfull < ncp) body_step(nfull, true);
      };
      // Pad a filled tile of cnt sources to a multiple of kP2PUnroll GA
      // (GA at most 8; the stage size is a multiple of 64) with zero charges
      // far away, which add exactly zero, then run both passes.
      auto run = [&](auto mask, int cnt) __attribute__((always_inline)) {
        const int g4 = kP2PUnroll * GA;
        const int cntp = (cnt + g4 - 1) / g4 * g4;
        for (int k = cnt + lane; k < cntp; k += 64) sb[k] = make_double4(kFar, kFar, kFar, 0);
        __syncthreads();
        const int ncp = __builtin_amdgcn_readfirstlane(cntp);
        pass(mask, GA, ncp, xa, ga, acca);
        if (tp.r > 0) pass_b(mask, ncp);
      };
      // The target leaf's own bodies first (masked), then the rest of the
      // list with the own entry left out (the traversal lists every leaf
      // against itself once).
      for (int s0 = 0; s0 < tn; s0 += kP2PStage) {
        const int cnt = min(kP2PStage, tn - s0);
        __syncthreads();
        for (int k = lane; k < cnt; k += 64) sb[k] = body[tb0 + s0 + k];
        run(std::true_type{}, cnt);
      }
      SrcWalker w{e0, 0, e0, 0};
      while (w.e < e1) {
        // Fill: consecutive source leaves (or parts of one) up to kP2PStage
        // bodies, each shifted by its periodic image.
        bool imaged = false;
        const int cnt = fill_stage<kP2PStage, kP2PStage / 64>(
            w, e1, lane, src, img, cBody, cNBody, body, sd,
            [&](int k, const double4 &vb, int im) {
              double4 v = vb;
              if (im != 13) {
                double px, py, pz;
                image_shift(im, cycle, px, py, pz);
                v.x += px;
                v.y += py;
                v.z += pz;
                imaged = true;
              }
              sb[k] = v;
            },
            c);
        if (cnt == 0) continue;
        if (__any(imaged))
          run(std::true_type{}, cnt);
        else
          run(std::false_type{}, cnt);
      }
      combine(acca, tp.TA, GA, min(nt, tp.TA), tbase);
      if (tp.r > 0) combine(accb, tp.r, GB, tp.r, tbase + tp.TA);
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 P2P, register-blocked: every lane holds K targets, so each source
// body read from LDS serves K pairs (the one-target layout above reads a
// source per pair, which kept the LDS return path close to saturation), and
// the next source is read into registers while the current one is evaluated
// (the loop never waits on a read it has just issued).
//
// Lane layout: lane = 8 g + i (g = 0..7 groups, i = 0..7) holds targets
// i, i + 8, ..., i + 8 (K - 1) of a block of n <= 8 KMAX targets,
// K = ceil(n / 8) (a compile-time constant per block), and takes the staged
// sources g, g + 8, g + 16, ...; the 8 group partials of a target are summed
// in group order at the end of the block (deterministic).  With ~32 bodies
// per leaf this wastes ~10% of the pair slots, as the two-pass layout does.
//
// The target leaf's own bodies form a masked stage sequence (R2 == 0 is the
// self pair, skipped by a select); the list walk leaves the own entry out, so
// every other stage runs without the select (distinct leaves never hold two
// bodies at one position: equal positions share a Morton key and a leaf),
// except stages with periodic images, which keep it.
constexpr int kRbStage = 256;  // sources staged per fill
constexpr int kRbG = 8;        // source groups (lanes per group: 8)

template <bool MASK>
__device__ inline void rb_pair(double xi, double yi, double zi, const double4 &xj, double4 &acc) {
  const double dx = xi - xj.x, dy = yi - xj.y, dz = zi - xj.z;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double rs = rsqrt_pos(r2);
  const double invr = MASK ? (r2 > 0 ? rs : 0.0) : rs;
  const double qinvr = xj.w * invr;
  const double qinvr3 = qinvr * invr * invr;
  acc.x += qinvr;
  acc.y += dx * qinvr3;
  acc.z += dy * qinvr3;
  acc.w += dz * qinvr3;
}

// One tile: sources sg[0], sg[8], ..., sg[ncp - 8] (ncp a multiple of 8);
// the read-ahead touches sg[ncp], which the caller pads.
template <bool MASK, int K>
__device__ inline void rb_tile(const double4 *__restrict__ sg, int ncp, const double (&x)[K], const double (&y)[K],
                               const double (&z)[K], double4 (&acc)[K]) {
  double4 cur = sg[0];
  for (int j = 0; j < ncp; j += kRbG) {
    const double4 nxt = sg[j + kRbG];
#pragma unroll
    for (int k = 0; k < K; k++) rb_pair<MASK>(x[k], y[k], z[k], cur, acc[k]);
    cur = nxt;
  }
}

template <int K>
__device__ void rb_block(int lane, int c, int tb0, int tbase, int nt, int e0, int e1, const int *__restrict__ src,
                         const unsigned char *__restrict__ img, const int *__restrict__ cBody,
                         const int *__restrict__ cNBody, const double4 *__restrict__ body, double4 *__restrict__ trg,
                         double cycle, int tn, double4 *sb, SrcDesc &sd) {
  const int g = lane / kRbG, i = lane % kRbG;
  double x[K], y[K], z[K];
  double4 acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int t = i + kRbG * k;
    // Absent targets sit far from everything (their results are dropped).
    const double4 v = t < nt ? body[tb0 + tbase + t] : make_double4(-kFar, -kFar, -kFar, 0);
    x[k] = v.x;
    y[k] = v.y;
    z[k] = v.z;
    acc[k] = make_double4(0, 0, 0, 0);
  }
  // Pad the tile to a multiple of 8, plus 8 for the read-ahead, with zero
  // charges far away (they add exactly zero), then run it.
  auto run = [&](int cnt, bool masked) __attribute__((always_inline)) {
    const int cntp = (cnt + kRbG - 1) / kRbG * kRbG;
    for (int k = cnt + lane; k < cntp + kRbG; k += 64) sb[k] = make_double4(kFar, kFar, kFar, 0);
    __syncthreads();
    const int ncp = __builtin_amdgcn_readfirstlane(cntp);
    if (masked)
      rb_tile<true, K>(sb + g, ncp, x, y, z, acc);
    else
      rb_tile<false, K>(sb + g, ncp, x, y, z, acc);
  };
  // The own leaf's bodies, masked.
  for (int s0 = 0; s0 < tn; s0 += kRbStage) {
    const int cnt = min(kRbStage, tn - s0);
    __syncthreads();
    for (int k = lane; k < cnt; k += 64) sb[k] = body[tb0 + s0 + k];
    run(cnt, true);
  }
  // The rest of the list, the own entry left out.
  SrcWalker w{e0, 0, e0, 0};
  while (w.e < e1) {
    bool imaged = false;
    const int cnt = fill_stage<kRbStage>(
        w, e1, lane, src, img, cBody, cNBody, body, sd,
        [&](int k, const double4 &vb, int im) {
          double px, py, pz;
          image_shift(im, cycle, px, py, pz);
          double4 v = vb;
          v.x += px;
          v.y += py;
          v.z += pz;
          sb[k] = v;
          imaged |= im != 13;
        },
        c);
    if (cnt > 0) run(cnt, __any(imaged));
  }
  // Group partials of target i + 8 k in slots 8 g + i: summed in group order.
#pragma unroll
  for (int k = 0; k < K; k++) {
    __syncthreads();
    sb[lane] = acc[k];
    __syncthreads();
    const int t = lane + kRbG * k;
    if (lane < kRbG && t < nt) {
      double4 s = sb[lane];
      for (int gg = 1; gg < 64 / kRbG; gg++) {
        const double4 o = sb[gg * kRbG + lane];
        s.x += o.x;
        s.y += o.y;
        s.z += o.z;
        s.w += o.w;
      }
      double4 tv = trg[tb0 + tbase + t];
      tv.x += s.x;
      tv.y -= s.y;
      tv.z -= s.z;
      tv.w -= s.w;
      trg[tb0 + tbase + t] = tv;
    }
  }
}

// KMAX: the most targets per lane; a leaf of more than 8 KMAX bodies runs as
// several blocks (each walks its list once).
template <int KMAX>
__global__ void __launch_bounds__(64)
    k_p2p_f64_rb(const int *__restrict__ tgt_list, int ntl_host, const int *__restrict__ ntl_dev,
                 int *__restrict__ next, const int *__restrict__ off, const int *__restrict__ src,
                 const unsigned char *__restrict__ img, const int *__restrict__ cBody,
                 const int *__restrict__ cNBody, const double4 *__restrict__ body, double4 *__restrict__ trg,
                 double cycle) {
  __shared__ double4 sb[kRbStage + 2 * kRbG];
  __shared__ SrcDesc sd;
  const int lane = threadIdx.x;
  const int ntl = ntl_dev ? *ntl_dev : ntl_host;
  for (int bi = blockIdx.x; bi < ntl; bi = next_leaf(next, lane)) {
    const int c = tgt_list[bi];
    const int tb0 = cBody[c], tn = cNBody[c];
    const int e0 = off[c], e1 = off[c + 1];
    for (int tbase = 0; tbase < tn; tbase += kRbG * KMAX) {
      const int nt = min(kRbG * KMAX, tn - tbase);
      const int K = (nt + kRbG - 1) / kRbG;
#define EXA_RB_BLOCK(k) rb_block<k>(lane, c, tb0, tbase, nt, e0, e1, src, img, cBody, cNBody, body, trg, cycle, tn, sb, sd)
      if (K <= 1)
        EXA_RB_BLOCK(1);
      else if (K == 2)
        EXA_RB_BLOCK(2);
      else if (K == 3 || KMAX == 3)
        EXA_RB_BLOCK(3 <= KMAX ? 3 : KMAX);
      else if (K == 4 || KMAX == 4)
        EXA_RB_BLOCK(4 <= KMAX ? 4 : KMAX);
      else
        EXA_RB_BLOCK(KMAX);
#undef EXA_RB_BLOCK
    }
  }
}

// Mixed precision (cfg.precision = 1): FP32 pair arithmetic on positions
// taken relative to the target leaf's centre in FP64 (the source's image
// shift folded in) and split into FP32 hi + lo parts, so the FP32 values are
// cell-sized and x_i - x_j = (hi_i - hi_j) + (lo_i - lo_j) keeps ~FP32
// relative accuracy even for bodies much closer to each other than the cell
// size (clustered distributions; a plain FP32 difference there would carry
// an error of one FP32 ulp of the cell size).  Each lane sums runs of kSeg
// sources in FP32 and adds each run's sum into a compensated (Kahan)
// accumulator: the per-target compensated summation of the reference's
// kahan<T> (kahan.h:35-41), applied per run instead of per term.  The
// compensated value is read out as s - c.
//
// The lo parts matter only for close pairs.  A target lies inside its leaf
// (|x_i - X| <= sqrt(3) w / 2 for leaf width w = 2 Rb), so a source within r
// of it has |x_j - X| <= sqrt(3) w / 2 + r, and the hi-only difference errs
// by at most 2^-24 (sqrt(3) w + r) per component: a relative error below
// 2^-24 (sqrt(3) w / r + 1), i.e. 1.7e-6, for r >= w / 16.  Stages that
// cannot hold the target's own position therefore take the hi parts alone
// and add the lo parts only in the (wave-uniform, rare) case that a lane's
// hi-only R2 falls below (w / 16)^2 = Rb^2 / 64; stages that may hold the
// target itself (its own leaf, periodic images) always take both.
constexpr int kP2PStageF = 256;
static_assert(kP2PStageF % 64 == 0, "stage tiles are padded to multiples of 64");
constexpr int kSeg = 8;  // FP32 pair terms summed before each compensated add
constexpr int kSegRun = 2;  // of which side by side (as kP2PUnroll in k_p2p_f64)
constexpr float kFarF = 1e18f;  // padding position: R2 = 3e36 stays finite in FP32

struct Kahan32 {
  float s = 0, c = 0;
  __device__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ double value() const { return (double)s - (double)c; }
};

__global__ void __launch_bounds__(64)
    k_p2p_mixed(const int *__restrict__ tgt_list, int ntl_host, const int *__restrict__ ntl_dev,
                int *__restrict__ next, const int *__restrict__ off, const int *__restrict__ src,
                const unsigned char *__restrict__ img, const int *__restrict__ cBody,
                const int *__restrict__ cNBody, const double4 *__restrict__ cX, const double *__restrict__ cRb,
                const double4 *__restrict__ body, double4 *__restrict__ trg, double cycle) {
  __shared__ __attribute__((aligned(32))) float4 sb[kP2PStageF];  // hi parts and charge
  __shared__ float4 sbl[kP2PStageF];                               // lo parts
  double4 *const sacc = reinterpret_cast<double4 *>(sb);           // (as in k_p2p_f64)
  __shared__ SrcDesc sd;
  const int lane = threadIdx.x;
  // Target leaves: blockIdx.x first, then the next unclaimed one of the
  // work counter (leaves differ in cost; the grid may be smaller than ntl).
  // ntl_dev (if set): the count as the list grouping left it on the device.
  const int ntl = ntl_dev ? *ntl_dev : ntl_host;
  for (int bi = blockIdx.x; bi < ntl; bi = next_leaf(next, lane)) {
    const int c = tgt_list[bi];
    const int tb0 = cBody[c], tn = cNBody[c];
    const double4 Xc = cX[c];
    const float thr2 = (float)(cRb[c] * cRb[c] * (1.0 / 64));  // (w / 16)^2: pairs that need the lo parts
    const int e0 = off[c], e1 = off[c + 1];
    // A target's position relative to the leaf centre, FP32 hi + lo.
    struct Tgt {
      float x, y, z, xl, yl, zl;
    };
    auto load_tgt = [&](bool active, int b) {
      const double4 xd = active ? body[b] : Xc;
      const double rx = xd.x - Xc.x, ry = xd.y - Xc.y, rz = xd.z - Xc.z;
      Tgt t;
      t.x = (float)rx;
      t.y = (float)ry;
      t.z = (float)rz;
      t.xl = (float)(rx - t.x);
      t.yl = (float)(ry - t.y);
      t.zl = (float)(rz - t.z);
      return t;
    };
    struct Acc {
      Kahan32 p, x, y, z;
    };
    // kSeg sources k0, k0 + G, ... in FP32, then one compensated add each.
    // MASK as in k_p2p_f64: only stages that may hold the target's own
    // position skip R2 == 0 by a select.
    auto segment = [&](auto mask, auto gc, const Tgt &t, int k0,
                       Acc &a) __attribute__((always_inline)) {
      constexpr int G = decltype(gc)::value;
      const float4 *sg = sb + k0, *sgl = sbl + k0;
      float pot = 0, ax = 0, ay = 0, az = 0;
      // The kSeg pairs in runs of kSegRun side by side (loads, then each
      // step of the pair arithmetic for all of them), summed in source order.
#pragma unroll
      for (int u0 = 0; u0 < kSeg; u0 += kSegRun) {
        float4 xj[kSegRun];
        float dx[kSegRun], dy[kSegRun], dz[kSegRun], r2[kSegRun], rs[kSegRun];
#pragma unroll
        for (int u = 0; u < kSegRun; u++) xj[u] = sg[(u0 + u) * G];
        bool close = decltype(mask)::value;
#pragma unroll
        for (int u = 0; u < kSegRun; u++) {
          dx[u] = t.x - xj[u].x;
          dy[u] = t.y - xj[u].y;
          dz[u] = t.z - xj[u].z;
          r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
          if (!decltype(mask)::value) close |= r2[u] < thr2;
        }
        // The lo parts: always where the stage may hold the target itself,
        // else only if some lane has a close pair (uniform branch).
        if (decltype(mask)::value || __any(close)) {
#pragma unroll
          for (int u = 0; u < kSegRun; u++) {
            const float4 xl = sgl[(u0 + u) * G];
            dx[u] = dx[u] + (t.xl - xl.x);
            dy[u] = dy[u] + (t.yl - xl.y);
            dz[u] = dz[u] + (t.zl - xl.z);
            r2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
          }
        }
#pragma unroll
        for (int u = 0; u < kSegRun; u++) rs[u] = __builtin_amdgcn_rsqf(r2[u]);
#pragma unroll
        for (int u = 0; u < kSegRun; u++) {
          const float invr = decltype(mask)::value ? (r2[u] > 0.0f ? rs[u] : 0.0f) : rs[u];
          const float qinvr = xj[u].w * invr;
          const float qinvr3 = qinvr * invr * invr;
          pot += qinvr;
          ax += dx[u] * qinvr3;
          ay += dy[u] * qinvr3;
          az += dz[u] * qinvr3;
        }
      }
      a.p.add(pot);
      a.x.add(ax);
      a.y.add(ay);
      a.z.add(az);
    };
    auto combine = [&](const Acc &a, int T, int n, int tfirst) {
      __syncthreads();
      sacc[lane] = make_double4(a.p.value(), a.x.value(), a.y.value(), a.z.value());
      __syncthreads();
      if (lane < n) {
        double4 s = sacc[lane];
        for (int gg = 1; gg < 64 / T; gg++) {
          const double4 o = sacc[gg * T + lane];
          s.x += o.x;
          s.y += o.y;
          s.z += o.z;
          s.w += o.w;
        }
        double4 tv = trg[tb0 + tfirst + lane];
        tv.x += s.x;
        tv.y -= s.y;
        tv.z -= s.z;
        tv.w -= s.w;
        trg[tb0 + tfirst + lane] = tv;
      }
    };
    for (int tbase = 0; tbase < tn; tbase += 64) {
      const int nt = min(64, tn - tbase);
      const TargetPasses2 tp(nt);
      const int GA = 64 / tp.TA, GB = 64 / tp.TB;
      const int ga = lane / tp.TA, ia = lane % tp.TA, gb = lane / tp.TB, ib = lane % tp.TB;
      const Tgt ta = load_tgt(ia < nt, tb0 + tbase + ia);
      const Tgt tb = load_tgt(ib < tp.r, tb0 + tbase + tp.TA + ib);
      Acc acca, accb;
      // Passes with G (1, 2, 4 or 8) a compile-time constant and a
      // wave-uniform trip count, as pass A of k_p2p_f64.
      auto pass_g = [&](auto mask, auto gc, int ncp, const Tgt &t, int g,
                        Acc &a) __attribute__((always_inline)) {
        constexpr int G = decltype(gc)::value;
        for (int j = 0; j < ncp; j += kSeg * G) segment(mask, gc, t, j + g, a);
      };
      auto pass = [&](auto mask, int G, int ncp, const Tgt &t, int g,
                      Acc &a) __attribute__((always_inline)) {
        if (G == 1)
          pass_g(mask, std::integral_constant<int, 1>{}, ncp, t, g, a);
        else if (G == 2)
          pass_g(mask, std::integral_constant<int, 2>{}, ncp, t, g, a);
        else if (G == 4)
          pass_g(mask, std::integral_constant<int, 4>{}, ncp, t, g, a);
        else
          pass_g(mask, std::integral_constant<int, 8>{}, ncp, t, g, a);
      };
      // Pad a filled tile to a multiple of kSeg G (G of either pass, at
      // most 8; the stage size is a multiple of 64) with zero charges far
      // away, which add exactly 0, then run both passes.
      auto run = [&](auto mask, int cnt) __attribute__((always_inline)) {
        const int g8 = kSeg * (tp.r > 0 ? GB : GA);
        const int cntp = (cnt + g8 - 1) / g8 * g8;
        for (int k = cnt + lane; k < cntp; k += 64) {
          sb[k] = make_float4(kFarF, kFarF, kFarF, 0);
          sbl[k] = make_float4(0, 0, 0, 0);
        }
        __syncthreads();
        const int ncp = __builtin_amdgcn_readfirstlane(cntp);
        pass(mask, GA, ncp, ta, ga, acca);
        if (tp.r > 0) pass(mask, GB, ncp, tb, gb, accb);
      };
      // Source body v shifted by (px, py, pz), relative to the leaf centre,
      // into slot k as FP32 hi (with the charge) and lo parts.
      auto put = [&](int k, const double4 &v, double px, double py, double pz) __attribute__((always_inline)) {
        const double rx = (v.x + px) - Xc.x, ry = (v.y + py) - Xc.y, rz = (v.z + pz) - Xc.z;
        const float hx = (float)rx, hy = (float)ry, hz = (float)rz;
        sb[k] = make_float4(hx, hy, hz, (float)v.w);
        sbl[k] = make_float4((float)(rx - hx), (float)(ry - hy), (float)(rz - hz), 0.0f);
      };
      // The target leaf's own bodies first (masked), then the rest of the
      // list with the own entry left out (as in k_p2p_f64).
      for (int s0 = 0; s0 < tn; s0 += kP2PStageF) {
        const int cnt = min(kP2PStageF, tn - s0);
        __syncthreads();
        for (int k = lane; k < cnt; k += 64) put(k, body[tb0 + s0 + k], 0, 0, 0);
        run(std::true_type{}, cnt);
      }
      SrcWalker w{e0, 0, e0, 0};
      while (w.e < e1) {
        bool imaged = false;
        const int cnt = fill_stage<kP2PStageF, 1>(
            w, e1, lane, src, img, cBody, cNBody, body, sd,
            [&](int k, const double4 &vb, int im) {
              double px = 0, py = 0, pz = 0;
              if (im != 13) {
                image_shift(im, cycle, px, py, pz);
                imaged = true;
              }
              put(k, vb, px, py, pz);
            },
            c);
        if (cnt == 0) continue;
        if (__any(imaged))
          run(std::true_type{}, cnt);
        else
          run(std::false_type{}, cnt);
      }
      combine(acca, tp.TA, min(nt, tp.TA), tbase);
      if (tp.r > 0) combine(accb, tp.TB, tp.r, tbase + tp.TA);
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed precision, register-blocked: the lane layout and stage walk of
// k_p2p_f64_rb (K targets per lane, own leaf masked, read-ahead) with the
// arithmetic of k_p2p_mixed: leaf-centred FP32 hi + lo positions, runs of
// kSeg FP32 terms per target, each run added into a compensated
// accumulator (kahan.h:35-41), read out as s - c.  Tiles are padded to whole
// runs of every group (kSeg 8 = 64 sources).
struct RbTgt32 {
  float x, y, z, xl, yl, zl;
};

// The pair terms of K targets against one source s (hi parts and charge);
// sl: the source's lo parts, read only where needed (see k_p2p_mixed).
template <bool MASK, int K>
__device__ inline void rb_pairs32(const RbTgt32 (&t)[K], const float4 &s, const float4 *__restrict__ sl, float thr2,
                                  float4 (&a)[K]) {
  float dx[K], dy[K], dz[K], r2[K];
  bool close = MASK;
#pragma unroll
  for (int k = 0; k < K; k++) {
    dx[k] = t[k].x - s.x;
    dy[k] = t[k].y - s.y;
    dz[k] = t[k].z - s.z;
    r2[k] = dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k];
    if (!MASK) close |= r2[k] < thr2;
  }
  if (MASK || __any(close)) {
    const float4 l = *sl;
#pragma unroll
    for (int k = 0; k < K; k++) {
      dx[k] = dx[k] + (t[k].xl - l.x);
      dy[k] = dy[k] + (t[k].yl - l.y);
      dz[k] = dz[k] + (t[k].zl - l.z);
      r2[k] = dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k];
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const float rs = __builtin_amdgcn_rsqf(r2[k]);
    const float invr = MASK ? (r2[k] > 0.0f ? rs : 0.0f) : rs;
    const float qinvr = s.w * invr;
    const float qinvr3 = qinvr * invr * invr;
    a[k].x += qinvr;
    a[k].y += dx[k] * qinvr3;
    a[k].z += dy[k] * qinvr3;
    a[k].w += dz[k] * qinvr3;
  }
}

template <bool MASK, int K>
__device__ inline void rb_tile32(const float4 *__restrict__ sg, const float4 *__restrict__ sgl, int ncp,
                                 const RbTgt32 (&t)[K], float thr2, Kahan32 (&kp)[K][4]) {
  float4 cur = sg[0];
  for (int j = 0; j < ncp; j += kRbG * kSeg) {
    float4 run[K];
#pragma unroll
    for (int k = 0; k < K; k++) run[k] = make_float4(0, 0, 0, 0);
#pragma unroll
    for (int u = 0; u < kSeg; u++) {
      const int jc = j + u * kRbG;
      const float4 nxt = sg[jc + kRbG];
      rb_pairs32<MASK, K>(t, cur, sgl + jc, thr2, run);
      cur = nxt;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      kp[k][0].add(run[k].x);
      kp[k][1].add(run[k].y);
      kp[k][2].add(run[k].z);
      kp[k][3].add(run[k].w);
    }
  }
}

template <int K>
__device__ void rb_block32(int lane, int c, int tb0, int tbase, int nt, int e0, int e1, const int *__restrict__ src,
                           const unsigned char *__restrict__ img, const int *__restrict__ cBody,
                           const int *__restrict__ cNBody, const double4 &Xc, float thr2,
                           const double4 *__restrict__ body, double4 *__restrict__ trg, double cycle, int tn, float4 *sb,
                           float4 *sbl, SrcDesc &sd) {
  const int g = lane / kRbG, i = lane % kRbG;
  RbTgt32 t[K];
  Kahan32 kp[K][4];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int ti = i + kRbG * k;
    // Absent targets sit at the leaf centre (finite; their results are dropped).
    const double4 v = ti < nt ? body[tb0 + tbase + ti] : Xc;
    const double rx = v.x - Xc.x, ry = v.y - Xc.y, rz = v.z - Xc.z;
    t[k].x = (float)rx;
    t[k].y = (float)ry;
    t[k].z = (float)rz;
    t[k].xl = (float)(rx - t[k].x);
    t[k].yl = (float)(ry - t[k].y);
    t[k].zl = (float)(rz - t[k].z);
  }
  // Source body b shifted by image im, relative to the leaf centre, into
  // slot k as FP32 hi (with the charge) and lo parts.
  auto put = [&](int k, const double4 &v, double px, double py, double pz) __attribute__((always_inline)) {
    const double rx = (v.x + px) - Xc.x, ry = (v.y + py) - Xc.y, rz = (v.z + pz) - Xc.z;
    const float hx = (float)rx, hy = (float)ry, hz = (float)rz;
    sb[k] = make_float4(hx, hy, hz, (float)v.w);
    sbl[k] = make_float4((float)(rx - hx), (float)(ry - hy), (float)(rz - hz), 0.0f);
  };
  auto run = [&](int cnt, bool masked) __attribute__((always_inline)) {
    constexpr int step = kRbG * kSeg;
    const int cntp = (cnt + step - 1) / step * step;
    for (int k = cnt + lane; k < cntp + kRbG; k += 64) {
      sb[k] = make_float4(kFarF, kFarF, kFarF, 0);
      sbl[k] = make_float4(0, 0, 0, 0);
    }
    __syncthreads();
    const int ncp = __builtin_amdgcn_readfirstlane(cntp);
    if (masked)
      rb_tile32<true, K>(sb + g, sbl + g, ncp, t, thr2, kp);
    else
      rb_tile32<false, K>(sb + g, sbl + g, ncp, t, thr2, kp);
  };
  for (int s0 = 0; s0 < tn; s0 += kRbStage) {
    const int cnt = min(kRbStage, tn - s0);
    __syncthreads();
    for (int k = lane; k < cnt; k += 64) put(k, body[tb0 + s0 + k], 0, 0, 0);
    run(cnt, true);
  }
  SrcWalker w{e0, 0, e0, 0};
  while (w.e < e1) {
    bool imaged = false;
    const int cnt = fill_stage<kRbStage>(
        w, e1, lane, src, img, cBody, cNBody, body, sd,
        [&](int k, const double4 &vb, int im) {
          double px, py, pz;
          image_shift(im, cycle, px, py, pz);
          put(k, vb, px, py, pz);
          imaged |= im != 13;
        },
        c);
    if (cnt > 0) run(cnt, __any(imaged));
  }
  double4 *const sacc = reinterpret_cast<double4 *>(sb);
#pragma unroll
  for (int k = 0; k < K; k++) {
    __syncthreads();
    sacc[lane] = make_double4(kp[k][0].value(), kp[k][1].value(), kp[k][2].value(), kp[k][3].value());
    __syncthreads();
    const int ti = lane + kRbG * k;
    if (lane < kRbG && ti < nt) {
      double4 s = sacc[lane];
      for (int gg = 1; gg < 64 / kRbG; gg++) {
        const double4 o = sacc[gg * kRbG + lane];
        s.x += o.x;
        s.y += o.y;
        s.z += o.z;
        s.w += o.w;
      }
      double4 tv = trg[tb0 + tbase + ti];
      tv.x += s.x;
      tv.y -= s.y;
      tv.z -= s.z;
      tv.w -= s.w;
      trg[tb0 + tbase + ti] = tv;
    }
  }
}

template <int KMAX>
__global__ void __launch_bounds__(64)
    k_p2p_mixed_rb(const int *__restrict__ tgt_list, int ntl_host, const int *__restrict__ ntl_dev,
                   int *__restrict__ next, const int *__restrict__ off, const int *__restrict__ src,
                   const unsigned char *__restrict__ img, const int *__restrict__ cBody,
                   const int *__restrict__ cNBody, const double4 *__restrict__ cX, const double *__restrict__ cRb,
                   const double4 *__restrict__ body, double4 *__restrict__ trg, double cycle) {
  // Stage (hi parts and charge, then lo parts), padded to whole runs plus
  // the read-ahead; the block's partials (64 double4) reuse the hi array.
  constexpr int kCap = kRbStage + kRbG * kSeg + kRbG;
  __shared__ __attribute__((aligned(32))) float4 sb[kCap];
  __shared__ float4 sbl[kCap];
  __shared__ SrcDesc sd;
  const int lane = threadIdx.x;
  const int ntl = ntl_dev ? *ntl_dev : ntl_host;
  for (int bi = blockIdx.x; bi < ntl; bi = next_leaf(next, lane)) {
    const int c = tgt_list[bi];
    const int tb0 = cBody[c], tn = cNBody[c];
    const double4 Xc = cX[c];
    const float thr2 = (float)(cRb[c] * cRb[c] * (1.0 / 64));  // as in k_p2p_mixed
    const int e0 = off[c], e1 = off[c + 1];
    for (int tbase = 0; tbase < tn; tbase += kRbG * KMAX) {
      const int nt = min(kRbG * KMAX, tn - tbase);
      const int K = (nt + kRbG - 1) / kRbG;
#define EXA_RB_BLOCK32(k) \
  rb_block32<k>(lane, c, tb0, tbase, nt, e0, e1, src, img, cBody, cNBody, Xc, thr2, body, trg, cycle, tn, sb, sbl, sd)
      if (K <= 1)
        EXA_RB_BLOCK32(1);
      else if (K == 2)
        EXA_RB_BLOCK32(2);
      else if (K == 3 || KMAX == 3)
        EXA_RB_BLOCK32(3 <= KMAX ? 3 : KMAX);
      else if (K == 4 || KMAX == 4)
        EXA_RB_BLOCK32(4 <= KMAX ? 4 : KMAX);
      else
        EXA_RB_BLOCK32(KMAX);
#undef EXA_RB_BLOCK32
    }
  }
}

// Target-source body pairs of the P2P list (for the interactions/s metric).
__global__ void k_count_body_pairs(const int *__restrict__ tgt_list, int ntl, const int *__restrict__ off,
                                   const int *__restrict__ src, const int *__restrict__ cNBody,
                                   unsigned long long *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntl) return;
  const int c = tgt_list[i];
  unsigned long long s = 0;
  for (int e = off[c]; e < off[c + 1]; e++) s += (unsigned long long)cNBody[src[e]];
  atomicAdd(out, s * (unsigned long long)cNBody[c]);
}

void run_p2p(Handle &h, hipStream_t st) {
  zero_device(h.trg4.p, (size_t)h.n * sizeof(double4), st);
  if (h.np2p_tgt == 0) return;
  // A resident grid of p2p_wpc waves per CU that claims target leaves from a
  // counter (0: one wave per leaf).  A resident grid leaves the rest of each
  // CU to the other stream's kernels (the lattice M2L set-up runs beside P2P
  // instead of queueing behind one wave per leaf); 12 waves (10.6 KB of LDS
  // and at most 128 VGPRs each) leave a third of the LDS and of the VGPRs.
  // The target count is read on the device when the host has not read it
  // back yet and the grid does not depend on it (a resident grid).
  const int *ntl_dev = nullptr;
  int grid = h.p2p_wpc * h.ncu;
  if (h.np2p_tgt < 0 && h.p2p_wpc > 0) {
    ntl_dev = h.p2p_cnt_dev;
  } else {
    resolve_p2p_count(h);
    if (h.np2p_tgt == 0) return;
    grid = h.p2p_wpc > 0 ? std::min(h.np2p_tgt, grid) : h.np2p_tgt;
  }
  h.p2p_next.reserve(1);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, h.p2p_next.p);
  if (h.cfg.precision == 1 && h.p2p_kmax == 1)
    hipLaunchKernelGGL(k_p2p_mixed, dim3(grid), dim3(64), 0, st, h.p2p_tgt.p, h.np2p_tgt, ntl_dev, h.p2p_next.p,
                       h.p2p_off.p, h.p2p_src.p, h.p2p_img.p, h.cBody.p, h.cNBody.p, h.cX.p, h.cRb.p, h.body4.p,
                       h.trg4.p, h.cfg.cycle);
  else if (h.cfg.precision == 1) {
    auto *kern = h.p2p_kmax == 3 ? k_p2p_mixed_rb<3> : h.p2p_kmax == 5 ? k_p2p_mixed_rb<5> : k_p2p_mixed_rb<4>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, st, h.p2p_tgt.p, h.np2p_tgt, ntl_dev, h.p2p_next.p,
                       h.p2p_off.p, h.p2p_src.p, h.p2p_img.p, h.cBody.p, h.cNBody.p, h.cX.p, h.cRb.p, h.body4.p,
                       h.trg4.p, h.cfg.cycle);
  }
  else if (h.p2p_kmax == 1)
    hipLaunchKernelGGL(k_p2p_f64, dim3(grid), dim3(64), 0, st, h.p2p_tgt.p, h.np2p_tgt, ntl_dev, h.p2p_next.p,
                       h.p2p_off.p, h.p2p_src.p, h.p2p_img.p, h.cBody.p, h.cNBody.p, h.body4.p, h.trg4.p, h.cfg.cycle);
  else {
    auto *kern = h.p2p_kmax == 3 ? k_p2p_f64_rb<3> : h.p2p_kmax == 5 ? k_p2p_f64_rb<5> : k_p2p_f64_rb<4>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), 0, st, h.p2p_tgt.p, h.np2p_tgt, ntl_dev, h.p2p_next.p,
                       h.p2p_off.p, h.p2p_src.p, h.p2p_img.p, h.cBody.p, h.cNBody.p, h.body4.p, h.trg4.p, h.cfg.cycle);
  }
  EXA_HIP(hipGetLastError());
}

int64_t count_p2p_body_pairs(Handle &h) {
  resolve_p2p_count(h);
  if (h.np2p_tgt == 0) return 0;
  h.red.reserve(2);
  EXA_HIP(hipMemsetAsync(h.red.p, 0, sizeof(unsigned long long), h.stream));
  hipLaunchKernelGGL(k_count_body_pairs, dim3(div_up(h.np2p_tgt, 256)), dim3(256), 0, h.stream, h.p2p_tgt.p,
                     h.np2p_tgt, h.p2p_off.p, h.p2p_src.p, h.cNBody.p, (unsigned long long *)h.red.p);
  unsigned long long v = 0;
  EXA_HIP(hipMemcpyAsync(&v, h.red.p, sizeof(v), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  return (int64_t)v;
}

}  // namespace exafmm
