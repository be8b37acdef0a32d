// Upward pass (P2M, M2M, setRmax, R /= theta) and downward pass (L2L, L2P),
// level-synchronous: one 64-lane workgroup per cell, one launch per level.
//
// Reference: include/up_down_pass.h:45-148 and
// kernels/LaplaceSphericalCPU.cxx:103-147,193-256.  Lanes own expansion
// coefficients (P2M, M2M, L2L) or bodies (L2P).  Bodies are summed in body
// order and children in child order, as the reference does.
#include <cmath>

#include "harmonics.h"
#include "internal.h"

namespace exafmm {

constexpr int kP2MChunk = 32;  // bodies per P2M table chunk

// Row t (real index) of one body's P2M table, column b: q conj(R_n^m(g)) for
// every n, m >= 0, from the recurrence of regular_nm (harmonics.h) run once
// per order m: h_m^m = (r sin(alpha)/2)^m / m!,
// h_n^m = -((2n-1) r cos(alpha) h_{n-1}^m + r^2 h_{n-2}^m) / (n^2 - m^2),
// R_n^m = h_n^m e^{i m beta}.
template <int P>
__device__ inline void p2m_rows(const Geom &g, double q, int b, double (*sR)[kP2MChunk + 1]) {
  const double rs = 0.5 * g.r * g.st, cr = g.ct * g.r, r2 = g.r * g.r;
  double pm = q, cm = 1, sm = 0;  // q (rs)^m / m!, e^{i m beta}
#pragma unroll
  for (int m = 0; m < P; m++) {
    double hp = 0, h = pm;
#pragma unroll
    for (int n = m; n < P; n++) {
      if (n > m) {
        const double hn = -((2 * n - 1) * cr * h + r2 * hp) * (1.0 / (n * n - m * m));
        hp = h;
        h = hn;
      }
      const int t = n * (n + 1) / 2 + m;
      sR[2 * t][b] = h * cm;
      sR[2 * t + 1][b] = -h * sm;
    }
    pm *= rs * (1.0 / (m + 1));
    const double c1 = cm * g.cp - sm * g.sp;
    sm = sm * g.cp + cm * g.sp;
    cm = c1;
  }
}

template <int P>
__global__ void __launch_bounds__(64)
    k_upward(const int *__restrict__ cells, const double4 *__restrict__ body, double4 *__restrict__ cX,
             const double *__restrict__ cRb, const int *__restrict__ cChild, const int *__restrict__ cNChild,
             const int *__restrict__ cBody, const int *__restrict__ cNBody, cplx *__restrict__ M, double theta,
             int useRmax, int m2m, int what) {
  // what: 1 = multipoles (P2M; M2M unless it runs as an MFMA product),
  //       2 = radii (setRmax, R /= theta), 3 = both.
  constexpr int NT = Expn<P>::NT;
  constexpr int NO = (NT + 63) / 64;  // coefficients per lane
  __shared__ double sR[2 * NT][kP2MChunk + 1];  // P2M rows (padded against bank conflicts)
  __shared__ cplx sY[NT];
  __shared__ cplx sM[NT];
  __shared__ double sred[64];
  const int c = cells[blockIdx.x];
  const int lane = threadIdx.x;
  const double4 Xc = cX[c];
  const int nchild = cNChild[c];
  double acc_re[NO], acc_im[NO];
  int tn[NO], tm[NO];
  for (int o = 0; o < NO; o++) {
    acc_re[o] = acc_im[o] = 0;
    const int t = lane + 64 * o;
    coef_nm(t < NT ? t : 0, tn[o], tm[o]);
  }
  double rmax = 0;
  if (nchild == 0) {
    const int b0 = cBody[c], nb = cNBody[c];
    if (!(what & 1)) {
      for (int i = lane; i < nb; i += 64) {
        const double4 v = body[b0 + i];
        const double r = exact_dist(Xc.x, Xc.y, Xc.z, v.x, v.y, v.z);
        rmax = r > rmax ? r : rmax;
      }
    } else {
      // P2M: M_n^m += q conj(R_n^m(x_b - X_c)) over the cell's bodies, in
      // order.  Chunks of kP2MChunk bodies: each body's lane writes its
      // charge-weighted table (all n, m by recurrence) into LDS, then lanes
      // own coefficients and add the chunk's rows in body order.
      for (int base = 0; base < nb; base += kP2MChunk) {
        const int cnt = min(kP2MChunk, nb - base);
        __syncthreads();  // the previous chunk's rows are consumed
        if (lane < cnt) {
          const double4 v = body[b0 + base + lane];
          if (what & 2) {
            const double r = exact_dist(Xc.x, Xc.y, Xc.z, v.x, v.y, v.z);
            rmax = r > rmax ? r : rmax;
          }
          p2m_rows<P>(make_geom(v.x - Xc.x, v.y - Xc.y, v.z - Xc.z), v.w, lane, sR);
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < NO; o++) {
          const int t = lane + 64 * o;
          if (t < NT)
            for (int k = 0; k < cnt; k++) {
              acc_re[o] += sR[2 * t][k];
              acc_im[o] += sR[2 * t + 1][k];
            }
        }
      }
    }
  } else if (!m2m || !(what & 1)) {
    // M2M runs as a batched MFMA product (or is not asked for): Rmax only.
    const int ch0 = cChild[c];
    if (lane == 0)
      for (int ci = 0; ci < nchild; ci++) {
        const double4 Xj = cX[ch0 + ci];
        const double r = exact_dist(Xc.x, Xc.y, Xc.z, Xj.x, Xj.y, Xj.z) + Xj.w;
        rmax = r > rmax ? r : rmax;
      }
  } else {
    // M2M (LaplaceSphericalCPU.cxx:120-147): child multipoles shifted to X_c.
    const int ch0 = cChild[c];
    for (int ci = 0; ci < nchild; ci++) {
      const int cc = ch0 + ci;
      const double4 Xj = cX[cc];
      const Geom g = make_geom(Xc.x - Xj.x, Xc.y - Xj.y, Xc.z - Xj.z);
      for (int t = lane; t < NT; t += 64) {
        int n, m;
        coef_nm(t, n, m);
        double re, im;
        regular_nm(n, m, g, re, im);
        sY[t] = {re, im};
        sM[t] = M[(size_t)cc * NT + t];
      }
      if (lane == 0) {
        const double r = exact_dist(Xc.x, Xc.y, Xc.z, Xj.x, Xj.y, Xj.z) + Xj.w;
        rmax = r > rmax ? r : rmax;
      }
      __syncthreads();
#pragma unroll
      for (int o = 0; o < NO; o++) {
        if (lane + 64 * o >= NT) continue;
        double lr = 0, li = 0;
        m2m_lane(tn[o], tm[o], sM, sY, lr, li);
        acc_re[o] += lr;
        acc_im[o] += li;
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int o = 0; o < NO; o++) {
    const int t = lane + 64 * o;
    if (t < NT && (what & 1) && (nchild == 0 || m2m)) M[(size_t)c * NT + t] = {acc_re[o], acc_im[o]};
  }
  if (!(what & 2)) return;
  // setRmax (up_down_pass.h:51-67) and R /= theta (:82).
  sred[lane] = rmax;
  __syncthreads();
  if (lane == 0) {
    double rm = 0;
    for (int i = 0; i < 64; i++) rm = sred[i] > rm ? sred[i] : rm;
    double R = cRb[c];
    if (useRmax) R = R < rm ? R : rm;
    R /= theta;
    cX[c].w = R;
  }
}

// Radii alone (what = 2 of k_upward: setRmax, up_down_pass.h:51-67, and
// R /= theta, :82), one wave per cell and four cells per workgroup, without
// k_upward's P2M tables in LDS (which held it to two waves per CU: the leaf
// level took 85 us).  The maximum is exact in any order.
__global__ void __launch_bounds__(256)
    k_radii(const int *__restrict__ cells, int ncell, const double4 *__restrict__ body, double4 *__restrict__ cX,
            const double *__restrict__ cRb, const int *__restrict__ cChild, const int *__restrict__ cNChild,
            const int *__restrict__ cBody, const int *__restrict__ cNBody, double theta, int useRmax) {
  const int w = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= ncell) return;  // wave-uniform
  const int c = cells[w];
  const double4 Xc = cX[c];
  const int nchild = cNChild[c];
  double rmax = 0;
  if (nchild == 0) {
    const int b0 = cBody[c], nb = cNBody[c];
    for (int i = lane; i < nb; i += 64) {
      const double4 v = body[b0 + i];
      const double r = exact_dist(Xc.x, Xc.y, Xc.z, v.x, v.y, v.z);
      rmax = r > rmax ? r : rmax;
    }
  } else if (lane < nchild) {
    const double4 Xj = cX[cChild[c] + lane];
    rmax = exact_dist(Xc.x, Xc.y, Xc.z, Xj.x, Xj.y, Xj.z) + Xj.w;
  }
  for (int d = 32; d >= 1; d >>= 1) {
    const double o = __shfl_xor(rmax, d);
    rmax = o > rmax ? o : rmax;
  }
  if (lane == 0) {
    double R = cRb[c];
    if (useRmax) R = R < rmax ? R : rmax;
    R /= theta;
    cX[c].w = R;
  }
}

// L2L (LaplaceSphericalCPU.cxx:193-224) for every cell of one level, then
// L2P (:226-256) for the leaves among them.
//
// PAIRS: L2P only, of the leaves in `cells` (all of them hold bodies), two
// leaves per 64-lane workgroup: lanes 32 h .. 32 h + 31 take leaf
// 2 blockIdx.x + h (a leaf holds about ncrit / 2 bodies, so a wave per leaf
// would idle half of its lanes); used when the L2L ran level by level on the
// matrix cores beforehand.  ncells: the number of entries of `cells`.
template <int P, bool PAIRS>
__global__ void __launch_bounds__(64)
    k_downward(const int *__restrict__ cells, const double4 *__restrict__ body, const double4 *__restrict__ cX,
               const int *__restrict__ cParent, const int *__restrict__ cNChild, const int *__restrict__ cBody,
               const int *__restrict__ cNBody, const cplx *__restrict__ M, cplx *__restrict__ L,
               double4 *__restrict__ trg, int do_l2l, int mass, int ncells) {
  constexpr int NT = Expn<P>::NT;
  constexpr int LW = PAIRS ? 32 : 64;  // lanes per cell
  __shared__ cplx sY[PAIRS ? 1 : NT];
  __shared__ cplx sL[PAIRS ? 1 : NT];
  __shared__ cplx sLcs[PAIRS ? 2 : 1][NT];
  const int half = PAIRS ? (int)threadIdx.x / 32 : 0;
  const int lane = PAIRS ? (int)threadIdx.x % 32 : (int)threadIdx.x;
  const int ci = PAIRS ? 2 * (int)blockIdx.x + half : (int)blockIdx.x;
  if (PAIRS && 2 * (int)blockIdx.x + 1 >= ncells && half == 1) return;  // no second leaf (wave-uniform per half)
  cplx *sLc = sLcs[half];
  const int c = cells[ci];
  const double4 Xc = cX[c];
  if (PAIRS) {
    for (int t = lane; t < NT; t += LW) sLc[t] = L[(size_t)c * NT + t];
  } else if (do_l2l) {
    const int p = cParent[c];
    const double4 Xp = cX[p];
    const Geom g = make_geom(Xc.x - Xp.x, Xc.y - Xp.y, Xc.z - Xp.z);
    for (int t = lane; t < NT; t += 64) {
      int n, m;
      coef_nm(t, n, m);
      double re, im;
      regular_nm(n, m, g, re, im);
      sY[t] = {re, im};
      sL[t] = L[(size_t)p * NT + t];
    }
    double m0 = 1;
    if (mass) m0 = M[(size_t)c * NT].re;  // L /= M[0], M[0] is real
    __syncthreads();
    for (int t = lane; t < NT; t += 64) {
      int j, k;
      coef_nm(t, j, k);
      double lr = 0, li = 0;
      l2l_lane<P>(j, k, sL, sY, lr, li);
      cplx own = L[(size_t)c * NT + t];
      if (mass) {
        own.re /= m0;
        own.im /= m0;
      }
      own.re += lr;
      own.im += li;
      L[(size_t)c * NT + t] = own;
      sLc[t] = own;
    }
  } else if (!PAIRS) {
    for (int t = lane; t < NT; t += 64) sLc[t] = L[(size_t)c * NT + t];
  }
  if (!PAIRS && cNChild[c] != 0) return;
  if (!PAIRS) __syncthreads();
  // (PAIRS: each half's own lanes wrote its expansion; LDS stores and loads
  // of one wave are ordered, so no barrier is needed between the halves.)
  // L2P: lanes own bodies.  The potential is sum_n sum_m w Re(L_n^m R_n^m)
  // (w = 1 for m = 0, else 2), as in LaplaceSphericalCPU.cxx:236-250.  The
  // gradient is taken in Cartesian form from the ladder relations of the
  // regular harmonics,
  //   d/dz R_n^m = -R_{n-1}^m,  (d/dx + i d/dy) R_n^m = -R_{n-1}^{m+1},
  //   (d/dx - i d/dy) R_n^m = R_{n-1}^{m-1}  (m >= 1; -conj R_{n-1}^1 at m = 0),
  // so every R_n'^m' (n' <= P-2) feeds the degree-(n'+1) terms.  This is the
  // same function as the reference's spherical-coordinate form without its
  // 1/r and 1/sin(theta) factors, so it stays finite at the expansion centre
  // and on the polar axis, where cart2sph/sph2cart give NaN.
  const int b0 = cBody[c], nb = cNBody[c];
  for (int bi = lane; bi < nb; bi += LW) {
    const double4 v = body[b0 + bi];
    const Geom g = make_geom(v.x - Xc.x, v.y - Xc.y, v.z - Xc.z);
    const double rho = g.r, x = g.ct, y = g.st;
    double T0 = 0, gx = 0, gy = 0, gz = 0;
    double fact = 1, pn = 1, rhom = 1;
    double er = 1, ei = 0;  // e^{i m beta}
    auto accumulate = [&](int n, int m, double Rr, double Ri) {
      const cplx Lv = sLc[tri(n) + m];
      T0 += (m == 0 ? 1.0 : 2.0) * (Lv.re * Rr - Lv.im * Ri);
      if (n + 1 >= P) return;
      const int n1 = n + 1;
      if (m >= 1) {  // D+ of (n+1, m-1)
        const cplx A = sLc[tri(n1) + m - 1];
        const double w = (m == 1) ? 0.5 : 1.0;
        gx -= w * (A.re * Rr - A.im * Ri);
        gy -= w * (A.re * Ri + A.im * Rr);
      }
      {  // D- of (n+1, m+1)
        const cplx A = sLc[tri(n1) + m + 1];
        gx += A.re * Rr - A.im * Ri;
        gy -= A.re * Ri + A.im * Rr;
      }
      if (m == 1) {  // D- of (n+1, 0): -conj(R_n^1)
        const cplx A = sLc[tri(n1)];
        gx -= 0.5 * (A.re * Rr + A.im * Ri);
        gy += 0.5 * (A.im * Rr - A.re * Ri);
      }
      {  // d/dz of (n+1, m)
        const cplx A = sLc[tri(n1) + m];
        gz -= (m == 0 ? 1.0 : 2.0) * (A.re * Rr - A.im * Ri);
      }
    };
#pragma unroll
    for (int m = 0; m < P; m++) {
      // One order m at a time: without the barrier the scheduler issued the
      // LDS reads of every term up front (256 VGPRs, one wave per SIMD).
      __builtin_amdgcn_sched_barrier(0);
      double p = pn;
      double p1 = p;
      p = x * (2 * m + 1) * p1;
      accumulate(m, m, rhom * p1 * er, rhom * p1 * ei);
      rhom *= rho;
      double rhon = rhom;
#pragma unroll
      for (int n = m + 1; n < P; n++) {
        rhon *= -1.0 / (n + m);  // compile-time reciprocals after unrolling
        accumulate(n, m, rhon * p * er, rhon * p * ei);
        const double p2 = p1;
        p1 = p;
        p = (x * (2 * n + 1) * p1 - (n + m) * p2) * (1.0 / (n - m + 1));
        rhon *= rho;
      }
      rhom *= -1.0 / ((2 * m + 2) * (2 * m + 1));
      pn = -pn * fact * y;
      fact += 2;
      const double t = er * g.cp - ei * g.sp;
      ei = er * g.sp + ei * g.cp;
      er = t;
    }
    double4 tv = trg[b0 + bi];
    tv.x += T0;
    tv.y += gx;
    tv.z += gy;
    tv.w += gz;
    trg[b0 + bi] = tv;
  }
}

template <int P>
static void upward_impl(Handle &h, int what, hipStream_t st) {
  const bool mf = updown_mfma_enabled(h);
  for (int t = h.nlevels - 1; t >= 0; t--) {
    const int cnt = h.lvl_off[t + 1] - h.lvl_off[t];
    if (cnt == 0) continue;
    const int *cells = h.lvl_cells.p + h.lvl_off[t];
    if (what == 2) {
      hipLaunchKernelGGL(k_radii, dim3(div_up(cnt, 4)), dim3(256), 0, st, cells, cnt, h.body4.p, h.cX.p, h.cRb.p,
                         h.cChild.p, h.cNChild.p, h.cBody.p, h.cNBody.p, h.cfg.theta, h.cfg.useRmax);
      continue;
    }
    hipLaunchKernelGGL(k_upward<P>, dim3(cnt), dim3(64), 0, st, cells, h.body4.p, h.cX.p, h.cRb.p, h.cChild.p,
                       h.cNChild.p, h.cBody.p, h.cNBody.p, h.M.p, h.cfg.theta, h.cfg.useRmax, mf ? 0 : 1, what);
    if ((what & 1) && mf && t + 1 < h.nlevels) run_m2m_level_mfma(h, cells, cnt, st);
  }
  EXA_HIP(hipGetLastError());
}

template <int P>
static void downward_impl(Handle &h) {
  // MASS mode divides each child's own L by its M_0 before the L2L
  // (LaplaceSphericalCPU.cxx MASS branch): kept in the per-cell kernel.
  const bool mf = updown_mfma_enabled(h) && !h.cfg.mass;
  if (mf) {
    // L2L level by level on the matrix cores, then L2P of every leaf with
    // bodies (the P2P target list holds exactly those: a leaf is always in
    // its own P2P list).
    for (int t = 1; t < h.nlevels; t++) {
      const int pc = h.lvl_off[t] - h.lvl_off[t - 1];
      if (pc > 0 && h.lvl_off[t + 1] > h.lvl_off[t]) run_l2l_level_mfma(h, h.lvl_cells.p + h.lvl_off[t - 1], pc);
    }
    resolve_p2p_count(h);
    if (h.np2p_tgt > 0)
      hipLaunchKernelGGL((k_downward<P, true>), dim3(div_up(h.np2p_tgt, 2)), dim3(64), 0, h.stream, h.p2p_tgt.p,
                         h.body4.p, h.cX.p, h.cParent.p, h.cNChild.p, h.cBody.p, h.cNBody.p, h.M.p, h.L.p, h.trg4.p,
                         0, 0, h.np2p_tgt);
    EXA_HIP(hipGetLastError());
    return;
  }
  for (int t = 0; t < h.nlevels; t++) {
    const int cnt = h.lvl_off[t + 1] - h.lvl_off[t];
    if (cnt == 0) continue;
    // Level 0 is the root: no L2L; L2P only when the root is a leaf
    // (up_down_pass.h:139).
    hipLaunchKernelGGL((k_downward<P, false>), dim3(cnt), dim3(64), 0, h.stream, h.lvl_cells.p + h.lvl_off[t],
                       h.body4.p, h.cX.p, h.cParent.p, h.cNChild.p, h.cBody.p, h.cNBody.p, h.M.p, h.L.p, h.trg4.p,
                       t > 0 ? 1 : 0, h.cfg.mass, cnt);
  }
  EXA_HIP(hipGetLastError());
}

#define EXA_DISPATCH_P(fn, h)                                                        \
  switch ((h).cfg.P) {                                                               \
    case 4: fn<4>(h); break;                                                         \
    case 6: fn<6>(h); break;                                                         \
    case 8: fn<8>(h); break;                                                         \
    case 10: fn<10>(h); break;                                                       \
    case 12: fn<12>(h); break;                                                       \
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");    \
  }

#define EXA_DISPATCH_P3(fn, h, a, b)                                                 \
  switch ((h).cfg.P) {                                                               \
    case 4: fn<4>(h, a, b); break;                                                   \
    case 6: fn<6>(h, a, b); break;                                                   \
    case 8: fn<8>(h, a, b); break;                                                   \
    case 10: fn<10>(h, a, b); break;                                                 \
    case 12: fn<12>(h, a, b); break;                                                 \
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");    \
  }

void run_upward(Handle &h) {
  h.M.resize((size_t)h.ncells * h.nterm);
  h.L.resize((size_t)h.ncells * h.nterm);
  EXA_HIP(hipMemsetAsync(h.L.p, 0, h.L.bytes(), h.stream));
  if (h.cfg.basis == 1) return run_upward_cart(h);
  EXA_DISPATCH_P3(upward_impl, h, 3, h.stream);
}

// The upward pass split over two streams: radii (needed by the traversal)
// on the main stream, multipoles (needed only by M2L) on the third stream
// after the radii, joined by the main stream through h.ev_mult before M is
// read (join_multipoles).  run_upward_mult is enqueued by the traversal once its
// own first kernels are queued (its launches would otherwise hold the host
// while the main stream idles).
void run_upward_radii(Handle &h) {
  h.M.resize((size_t)h.ncells * h.nterm);
  h.L.resize((size_t)h.ncells * h.nterm);
  EXA_HIP(hipMemsetAsync(h.L.p, 0, h.L.bytes(), h.stream));
  EXA_DISPATCH_P3(upward_impl, h, 2, h.stream);
  EXA_HIP(hipEventRecord(h.ev[16], h.stream));
}

void run_upward_mult(Handle &h) {
  EXA_HIP(hipStreamWaitEvent(h.stream3, h.ev[16], 0));
  EXA_HIP(hipEventRecord(h.ev[17], h.stream3));
  EXA_DISPATCH_P3(upward_impl, h, 1, h.stream3);
  run_lattice_mscale(h, h.stream3);  // the lattice M2L's input
  EXA_HIP(hipEventRecord(h.ev_mult, h.stream3));
}

void run_multipoles(Handle &h, hipStream_t st) { EXA_DISPATCH_P3(upward_impl, h, 1, st); }

void run_downward(Handle &h) {
  if (h.cfg.basis == 1) return run_downward_cart(h);
  EXA_DISPATCH_P(downward_impl, h);
}

// The pass kernels on a small cell array built by the caller (the per-cell
// operators of the C-ABI): multipoles (P2M at leaves, M2M per cell from its
// children) of `cells`, and L2L (do_l2l) then L2P at leaves of `cells`.
template <int P>
static void upward_cells_impl(Handle &h, const CellSetView &v) {
  hipLaunchKernelGGL(k_upward<P>, dim3(v.ncell), dim3(64), 0, h.stream, v.cells, v.body, v.cX, v.cRb, v.cChild,
                     v.cNChild, v.cBody, v.cNBody, v.M, h.cfg.theta, h.cfg.useRmax, 1, 1);
}

template <int P>
static void downward_cells_impl(Handle &h, const CellSetView &v) {
  hipLaunchKernelGGL((k_downward<P, false>), dim3(v.ncell), dim3(64), 0, h.stream, v.cells, v.body, v.cX,
                     v.cParent, v.cNChild, v.cBody, v.cNBody, v.M, v.L, v.trg, v.do_l2l, h.cfg.mass, v.ncell);
}

void launch_upward_cells(Handle &h, const CellSetView &v) {
  if (h.cfg.basis == 1) return launch_upward_cart_cells(h, v);
  switch (h.cfg.P) {
    case 4: upward_cells_impl<4>(h, v); break;
    case 6: upward_cells_impl<6>(h, v); break;
    case 8: upward_cells_impl<8>(h, v); break;
    case 10: upward_cells_impl<10>(h, v); break;
    case 12: upward_cells_impl<12>(h, v); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
  EXA_HIP(hipGetLastError());
}

void launch_downward_cells(Handle &h, const CellSetView &v) {
  if (h.cfg.basis == 1) return launch_downward_cart_cells(h, v);
  switch (h.cfg.P) {
    case 4: downward_cells_impl<4>(h, v); break;
    case 6: downward_cells_impl<6>(h, v); break;
    case 8: downward_cells_impl<8>(h, v); break;
    case 10: downward_cells_impl<10>(h, v); break;
    case 12: downward_cells_impl<12>(h, v); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
  EXA_HIP(hipGetLastError());
}

}  // namespace exafmm
