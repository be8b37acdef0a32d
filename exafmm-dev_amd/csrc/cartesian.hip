// Cartesian (Taylor-series) basis: P2M, M2M, M2L, L2L, L2P and the periodic
// far field, for P = 4, 6, 8 (NTERM = P (P+1) (P+2) / 6 real coefficients).
//
// Reference: kernels/LaplaceCartesianCPU.cxx (Index :3-24, power/derivative/
// scale and the M2M/M2L/L2L/L2P sums :26-252, getCoef/sumM2L :254-588, the
// kernel entry points :590-704).  Conventions:
//   M_k = sum_b q_b (X_c - x_b)^k / k!          (P2M, M2M)
//   C_k = d^k (1/r) at dX = X_i - X_j - X_p     (raw derivatives; MASS: times
//                                                M0_i M0_j)
//   L_a += M_0 C_a + sum_{b} M_b C_{a+b}         (MASS: C_a for the first term)
//   phi(x) = sum_k L_k (x - X_c)^k / k!          (L2L, L2P)
// with k a multi-index, |k| <= P-1.  Coefficients are stored in the real parts
// of the handle's complex M and L arrays (imaginary parts 0), at the index
// order of Index<kx,ky,kz>::I: by degree, then kx descending, then ky
// descending.
//
// The upward pass follows the reference's operation order without FMA
// contraction (multipoles bit-identical to the reference).  M2L pairs are
// batched per target cell: each lane builds one pair's derivative tensor
// (Taylor recurrence) in LDS, then the lanes run a host-built, load-balanced
// schedule of (a, b) terms over the whole batch.
#include <algorithm>
#include <vector>

#include "harmonics.h"
#include "internal.h"

namespace exafmm {

template <int P>
struct Cart {
  static constexpr int NT = P * (P + 1) * (P + 2) / 6;
};

// Index of (a, b, c): n(n+1)(n+2)/6 + m(m+1)/2 + (m - b), n = a+b+c, m = b+c.
__host__ __device__ inline int cidx(int a, int b, int c) {
  const int n = a + b + c, m = b + c;
  return n * (n + 1) * (n + 2) / 6 + m * (m + 1) / 2 + (m - b);
}

__host__ __device__ inline void cmi(int t, int &a, int &b, int &c) {
  int n = 0;
  while ((n + 1) * (n + 2) * (n + 3) / 6 <= t) n++;
  const int r = t - n * (n + 1) * (n + 2) / 6;
  int m = 0;
  while ((m + 1) * (m + 2) / 2 <= r) m++;
  b = m - (r - m * (m + 1) / 2);
  c = m - b;
  a = n - m;
}

// One coefficient of Kernels::power (:183-186,213-216,243-246): the chain
// C_(a,0,0) = C_(a-1,0,0) dx / a, C_(a,b,0) = C_(a,b-1,0) dy / b,
// C_(a,b,c) = C_(a,b,c-1) dz / c, started from v.
__device__ inline double cpow(double v, int a, int b, int c, double dx, double dy, double dz) {
#pragma clang fp contract(off)
  for (int i = 1; i <= a; i++) v = v * dx / i;
  for (int i = 1; i <= b; i++) v = v * dy / i;
  for (int i = 1; i <= c; i++) v = v * dz / i;
  return v;
}

// Derivative tensor C_k = d^k (1/r) of one offset into row[0, NT): the
// Taylor coefficients a_k = ((1-2n) sum_d x_d a_{k-e_d} + (1-n) sum_d
// a_{k-2e_d}) / n / r^2 (Kernels::derivative, :187-191), then k! (scale).
template <int P>
__device__ inline void cart_coef(double *row, double dx, double dy, double dz, double invR2, double invR) {
  constexpr int NT = Cart<P>::NT;
  row[0] = invR;
  int t = 1;
  for (int n = 1; n < P; n++)
    for (int a = n; a >= 0; a--)
      for (int b = n - a; b >= 0; b--, t++) {
        const int c = n - a - b;
        double s1 = 0, s2 = 0;
        if (a >= 1) s1 += dx * row[cidx(a - 1, b, c)];
        if (b >= 1) s1 += dy * row[cidx(a, b - 1, c)];
        if (c >= 1) s1 += dz * row[cidx(a, b, c - 1)];
        if (a >= 2) s2 += row[cidx(a - 2, b, c)];
        if (b >= 2) s2 += row[cidx(a, b - 2, c)];
        if (c >= 2) s2 += row[cidx(a, b, c - 2)];
        row[t] = ((1 - 2 * n) * s1 + (1 - n) * s2) / n * invR2;
      }
  t = 1;
  for (int n = 1; n < P; n++)
    for (int a = n; a >= 0; a--)
      for (int b = n - a; b >= 0; b--, t++) {
        const int c = n - a - b;
        double f = 1;
        for (int i = 2; i <= a; i++) f *= i;
        for (int i = 2; i <= b; i++) f *= i;
        for (int i = 2; i <= c; i++) f *= i;
        row[t] *= f;
      }
  (void)NT;
}

// Terms (a, b) of sumM2L at order P (:373-588): |a| + |b| <= P-1, where the
// reference's unrolled forms for P <= 7 leave out |b| = P-1 at a = 0.
__host__ __device__ inline bool m2l_term(int P, int da, int db) {
  return da + db <= P - 1 && !(P <= 7 && da == 0 && db == P - 1);
}

// ---------------------------------------------------------------------------
// Upward pass: P2M (:590-598) at leaves, M2M (:600-612) at parents, then
// setRmax and R /= theta as in the spherical pass.
template <int P>
__global__ void __launch_bounds__(64)
    k_upward_cart(const int *__restrict__ cells, const double4 *__restrict__ body, double4 *__restrict__ cX,
                  const double *__restrict__ cRb, const int *__restrict__ cChild, const int *__restrict__ cNChild,
                  const int *__restrict__ cBody, const int *__restrict__ cNBody, cplx *__restrict__ M, double theta,
                  int useRmax) {
#pragma clang fp contract(off)
  constexpr int NT = Cart<P>::NT;
  constexpr int NO = (NT + 63) / 64;
  __shared__ double4 sb[64];
  __shared__ double sC[NT];
  __shared__ double sM[NT];
  __shared__ double sred[64];
  const int c = cells[blockIdx.x];
  const int lane = threadIdx.x;
  const double4 Xc = cX[c];
  const int nchild = cNChild[c];
  int ka[NO], kb[NO], kc[NO];
  double acc[NO];
#pragma unroll
  for (int o = 0; o < NO; o++) {
    const int t = lane + 64 * o;
    cmi(t < NT ? t : 0, ka[o], kb[o], kc[o]);
    acc[o] = 0;
  }
  double rmax = 0;
  if (nchild == 0) {
    const int b0 = cBody[c], nb = cNBody[c];
    for (int base = 0; base < nb; base += 64) {
      const int cnt = min(64, nb - base);
      if (lane < cnt) {
        const double4 v = body[b0 + base + lane];
        sb[lane] = make_double4(Xc.x - v.x, Xc.y - v.y, Xc.z - v.z, v.w);
        const double r = exact_dist(Xc.x, Xc.y, Xc.z, v.x, v.y, v.z);
        rmax = r > rmax ? r : rmax;
      }
      __syncthreads();
      for (int k = 0; k < cnt; k++) {
        const double4 d = sb[k];
#pragma unroll
        for (int o = 0; o < NO; o++)
          if (lane + 64 * o < NT) acc[o] += cpow(d.w, ka[o], kb[o], kc[o], d.x, d.y, d.z);
      }
      __syncthreads();
    }
  } else {
    const int ch0 = cChild[c];
    for (int ci = 0; ci < nchild; ci++) {
      const int cc = ch0 + ci;
      const double4 Xj = cX[cc];
      const double dx = Xc.x - Xj.x, dy = Xc.y - Xj.y, dz = Xc.z - Xj.z;
      for (int t = lane; t < NT; t += 64) {
        int a, b, cz;
        cmi(t, a, b, cz);
        sC[t] = cpow(1.0, a, b, cz, dx, dy, dz);
        sM[t] = M[(size_t)cc * NT + t].re;
      }
      if (lane == 0) {
        const double r = exact_dist(Xc.x, Xc.y, Xc.z, Xj.x, Xj.y, Xj.z) + Xj.w;
        rmax = r > rmax ? r : rmax;
      }
      __syncthreads();
#pragma unroll
      for (int o = 0; o < NO; o++) {
        const int t = lane + 64 * o;
        if (t >= NT) continue;
        acc[o] += sC[t] * sM[0];
        if (t == 0) continue;
        // MultipoleSum: 0 <= k <= n, k != 0, lexicographic (kx, ky, kz).
        double sum = 0;
        for (int a = 0; a <= ka[o]; a++)
          for (int b = 0; b <= kb[o]; b++)
            for (int cz = 0; cz <= kc[o]; cz++) {
              if (a == 0 && b == 0 && cz == 0) continue;
              sum += sC[cidx(ka[o] - a, kb[o] - b, kc[o] - cz)] * sM[cidx(a, b, cz)];
            }
        acc[o] += sum;
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int o = 0; o < NO; o++) {
    const int t = lane + 64 * o;
    if (t < NT) M[(size_t)c * NT + t] = {acc[o], 0.0};
  }
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

// ---------------------------------------------------------------------------
// Downward pass: L2L (:631-645) for the cells of one level, then L2P
// (:647-664) at the leaves among them.
template <int P>
__global__ void __launch_bounds__(64)
    k_downward_cart(const int *__restrict__ cells, const double4 *__restrict__ body, const double4 *__restrict__ cX,
                    const int *__restrict__ cParent, const int *__restrict__ cNChild, const int *__restrict__ cBody,
                    const int *__restrict__ cNBody, const cplx *__restrict__ M, cplx *__restrict__ L,
                    double4 *__restrict__ trg, int do_l2l, int mass) {
#pragma clang fp contract(off)
  constexpr int NT = Cart<P>::NT;
  __shared__ double sC[NT];
  __shared__ double sL[NT];
  __shared__ double sLc[NT];
  const int c = cells[blockIdx.x];
  const int lane = threadIdx.x;
  const double4 Xc = cX[c];
  if (do_l2l) {
    const int p = cParent[c];
    const double4 Xp = cX[p];
    const double dx = Xc.x - Xp.x, dy = Xc.y - Xp.y, dz = Xc.z - Xp.z;
    for (int t = lane; t < NT; t += 64) {
      int a, b, cz;
      cmi(t, a, b, cz);
      sC[t] = cpow(1.0, a, b, cz, dx, dy, dz);
      sL[t] = L[(size_t)p * NT + t].re;
    }
    const double m0 = mass ? M[(size_t)c * NT].re : 1.0;
    __syncthreads();
    for (int t = lane; t < NT; t += 64) {
      double own = L[(size_t)c * NT + t].re;
      if (mass) own /= m0;
      own += sL[t];
      if (t == 0) {
        for (int i = 1; i < NT; i++) own += sC[i] * sL[i];
      } else {
        int a, b, cz;
        cmi(t, a, b, cz);
        const int dn = a + b + cz;
        double sum = 0;
        int k = 1;
        for (int n = 1; n <= P - 1 - dn; n++)
          for (int ka = n; ka >= 0; ka--)
            for (int kb = n - ka; kb >= 0; kb--, k++)
              sum += sC[k] * sL[cidx(a + ka, b + kb, cz + n - ka - kb)];
        own += sum;
      }
      L[(size_t)c * NT + t] = {own, 0.0};
      sLc[t] = own;
    }
  } else {
    for (int t = lane; t < NT; t += 64) sLc[t] = L[(size_t)c * NT + t].re;
  }
  if (cNChild[c] != 0) return;
  __syncthreads();
  const int b0 = cBody[c], nb = cNBody[c];
  for (int bi = lane; bi < nb; bi += 64) {
    const double4 v = body[b0 + bi];
    const double dx = v.x - Xc.x, dy = v.y - Xc.y, dz = v.z - Xc.z;
    double4 tv = trg[b0 + bi];
    tv.x += sLc[0];
    tv.y += sLc[1];
    tv.z += sLc[2];
    tv.w += sLc[3];
    double gx = 0, gy = 0, gz = 0;
    int k = 1;
    for (int n = 1; n < P; n++)
      for (int ka = n; ka >= 0; ka--)
        for (int kb = n - ka; kb >= 0; kb--, k++) {
          const int kc = n - ka - kb;
          const double Ck = cpow(1.0, ka, kb, kc, dx, dy, dz);
          tv.x += Ck * sLc[k];
          if (n <= P - 2) {
            gx += Ck * sLc[cidx(ka + 1, kb, kc)];
            gy += Ck * sLc[cidx(ka, kb + 1, kc)];
            gz += Ck * sLc[cidx(ka, kb, kc + 1)];
          }
        }
    tv.y += gx;
    tv.z += gy;
    tv.w += gz;
    trg[b0 + bi] = tv;
  }
}

// ---------------------------------------------------------------------------
// M2L over the CSR interaction list, one 64-lane workgroup per target cell.
// Schedule (host): the terms (a, b) of every output a, b in index order, cut
// into chunks of at most kCartChunk terms and packed onto the 64 lanes by
// longest-first.  Descriptor: bits [0,10) b, [10,20) index of a+b, bit 20
// end of chunk.  Padding items use b = NT (a zero column) and a+b = 0.
constexpr int kCartChunk = 8;

struct CartSchedule {
  std::vector<uint32_t> desc;  // ipl x 64, item-major
  std::vector<int> slot0;      // 64
  std::vector<int> outslot;    // NT x maxc (-1 padded)
  int ipl = 0, maxc = 0, nslots = 0;
};

static CartSchedule build_cart_schedule(int P) {
  const int NT = P * (P + 1) * (P + 2) / 6;
  struct Chunk {
    int a;
    std::vector<uint32_t> items;
  };
  std::vector<Chunk> chunks;
  std::vector<int> nchunk(NT, 0);
  for (int a = 0; a < NT; a++) {
    int xa, ya, za;
    cmi(a, xa, ya, za);
    const int da = xa + ya + za;
    Chunk cur{a, {}};
    for (int b = 0; b < NT; b++) {
      int xb, yb, zb;
      cmi(b, xb, yb, zb);
      if (!m2l_term(P, da, xb + yb + zb)) continue;
      cur.items.push_back((uint32_t)b | ((uint32_t)cidx(xa + xb, ya + yb, za + zb) << 10));
      if ((int)cur.items.size() == kCartChunk) {
        chunks.push_back(cur);
        nchunk[a]++;
        cur.items.clear();
      }
    }
    if (!cur.items.empty()) {
      chunks.push_back(cur);
      nchunk[a]++;
    }
  }
  std::vector<int> order(chunks.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return chunks[x].items.size() > chunks[y].items.size(); });
  std::vector<std::vector<int>> lanes(64);
  std::vector<int> load(64, 0);
  for (int ch : order) {
    int best = 0;
    for (int l = 1; l < 64; l++)
      if (load[l] < load[best]) best = l;
    lanes[best].push_back(ch);
    load[best] += (int)chunks[ch].items.size();
  }
  CartSchedule s;
  s.ipl = *std::max_element(load.begin(), load.end());
  s.maxc = *std::max_element(nchunk.begin(), nchunk.end());
  s.desc.assign((size_t)s.ipl * 64, (uint32_t)NT);
  s.slot0.assign(64, 0);
  s.outslot.assign((size_t)NT * s.maxc, -1);
  std::vector<int> fill(NT, 0);
  // Slots are numbered lane-major; outslot lists each output's chunks in
  // term order (chunk creation order).
  std::vector<int> chunk_slot(chunks.size());
  int slot = 0;
  for (int l = 0; l < 64; l++) {
    s.slot0[l] = slot;
    int it = 0;
    for (int ch : lanes[l]) {
      const auto &items = chunks[ch].items;
      for (size_t i = 0; i < items.size(); i++) {
        uint32_t d = items[i];
        if (i + 1 == items.size()) d |= 1u << 20;
        s.desc[(size_t)it * 64 + l] = d;
        it++;
      }
      chunk_slot[ch] = slot++;
    }
  }
  s.nslots = slot;
  for (size_t ch = 0; ch < chunks.size(); ch++) {
    const int a = chunks[ch].a;
    s.outslot[(size_t)a * s.maxc + fill[a]++] = chunk_slot[ch];
  }
  return s;
}

template <int P>
struct CartM2LShape {
  static constexpr int NT = Cart<P>::NT;
  static constexpr int RS = NT + 1;               // LDS row stride (zero column at NT)
  static constexpr int B = NT <= 64 ? 64 : 32;     // pairs per batch
  static constexpr int NCHUNK = 3 * NT + 256;      // slot capacity (checked on the host)
};

template <int P, int IPL>
__global__ void __launch_bounds__(64)
    k_m2l_cart(const int *__restrict__ tgt_list, const int *__restrict__ off, const int *__restrict__ src,
               const unsigned char *__restrict__ img, const double4 *__restrict__ cX, const cplx *__restrict__ M,
               cplx *__restrict__ L, const uint32_t *__restrict__ sched, const int *__restrict__ slot0,
               const int *__restrict__ outslot, int maxc, double cycle, int mass) {
  using S = CartM2LShape<P>;
  constexpr int NT = S::NT, RS = S::RS, B = S::B;
  __shared__ double sC[B * RS];
  __shared__ double sM[B * RS];
  __shared__ double slots[S::NCHUNK];
  __shared__ int ssrc[B];
  const int c = tgt_list[blockIdx.x];
  const int lane = threadIdx.x;
  uint32_t desc[IPL];
#pragma unroll
  for (int it = 0; it < IPL; it++) desc[it] = sched[it * 64 + lane];
  const int myslot0 = slot0[lane];
  for (int t = lane; t < S::NCHUNK; t += 64) slots[t] = 0;
  const double4 Xi = cX[c];
  const double mi0 = M[(size_t)c * NT].re;
  const int e0 = off[c], e1 = off[c + 1];
  for (int base = e0; base < e1; base += B) {
    const int cnt = min(B, e1 - base);
    __syncthreads();  // the previous batch is done with the tables
    if (lane < cnt) {
      const int s = src[base + lane];
      const int im = img[base + lane];
      ssrc[lane] = s;
      const double4 Xj = cX[s];
      const double px = (im / 9 - 1) * cycle, py = ((im / 3) % 3 - 1) * cycle, pz = (im % 3 - 1) * cycle;
      const double dx = Xi.x - Xj.x - px, dy = Xi.y - Xj.y - py, dz = Xi.z - Xj.z - pz;
      double r2 = 0;
      r2 += dx * dx;
      r2 += dy * dy;
      r2 += dz * dz;
      const double invR2 = 1 / r2;
      const double invR = mass ? mi0 * M[(size_t)s * NT].re * sqrt(invR2) : sqrt(invR2);
      cart_coef<P>(sC + lane * RS, dx, dy, dz, invR2, invR);
      sC[lane * RS + NT] = 0;
    }
    __syncthreads();
    for (int p = 0; p < cnt; p++) {
      const int s = ssrc[p];
      for (int t = lane; t < NT; t += 64) sM[p * RS + t] = (t == 0 && mass) ? 1.0 : M[(size_t)s * NT + t].re;
      if (lane == 0) sM[p * RS + NT] = 0;
    }
    __syncthreads();
    double acc = 0;
    int slot = myslot0;
#pragma unroll
    for (int it = 0; it < IPL; it++) {
      const uint32_t d = desc[it];
      const int b = d & 0x3FF, ab = (d >> 10) & 0x3FF;
      for (int p = 0; p < cnt; p++) acc += sM[p * RS + b] * sC[p * RS + ab];
      if (d & (1u << 20)) {
        slots[slot] += acc;
        slot++;
        acc = 0;
      }
    }
  }
  __syncthreads();
  for (int a = lane; a < NT; a += 64) {
    double sum = 0;
    for (int k = 0; k < maxc; k++) {
      const int sl = outslot[a * maxc + k];
      if (sl >= 0) sum += slots[sl];
    }
    cplx o = L[(size_t)c * NT + a];
    o.re += sum;
    L[(size_t)c * NT + a] = o;
  }
}

template <int P, int IPL>
static void m2l_cart_launch(Handle &h, const M2LList &lst, const double4 *cX, const cplx *M, cplx *L,
                            const uint32_t *desc, const int *slot0, const int *outslot, int maxc) {
  hipLaunchKernelGGL((k_m2l_cart<P, IPL>), dim3(lst.ntgt), dim3(64), 0, h.stream, lst.tgt, lst.off, lst.src, lst.img,
                     cX, M, L, desc, slot0, outslot, maxc, h.cfg.cycle, h.cfg.mass);
}

// Items per lane, rounded up to the compiled register budgets.
template <int P>
struct CartIPL;
template <>
struct CartIPL<4> {
  static constexpr int V = 8;
};
template <>
struct CartIPL<6> {
  static constexpr int V = 16;
};
template <>
struct CartIPL<8> {
  static constexpr int V = 40;
};

template <int P>
static void m2l_cart_arrays(Handle &h, const M2LList &lst, const double4 *cX, const cplx *M, cplx *L) {
  using S = CartM2LShape<P>;
  const int key = 1000 + P;
  if (h.m2l_sched_P != key) {
    const CartSchedule sc = build_cart_schedule(P);
    if (sc.ipl > CartIPL<P>::V || sc.nslots > S::NCHUNK)
      throw Error(EXAFMM_HIP_EUNSUP, "Cartesian M2L schedule exceeds its compiled budget");
    std::vector<uint32_t> blob(sc.desc);
    blob.resize((size_t)CartIPL<P>::V * 64, (uint32_t)S::NT);  // pad to the compiled item count
    for (int v : sc.slot0) blob.push_back((uint32_t)v);
    for (int v : sc.outslot) blob.push_back((uint32_t)v);
    h.m2l_sched.resize(blob.size());
    EXA_HIP(hipMemcpyAsync(h.m2l_sched.p, blob.data(), blob.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    h.m2l_sched_P = key;
    h.cart_maxc = sc.maxc;
  }
  const uint32_t *desc = h.m2l_sched.p;
  const int *slot0 = reinterpret_cast<const int *>(desc + 64 * CartIPL<P>::V);
  const int *outslot = slot0 + 64;
  if (lst.ntgt > 0) m2l_cart_launch<P, CartIPL<P>::V>(h, lst, cX, M, L, desc, slot0, outslot, h.cart_maxc);
  EXA_HIP(hipGetLastError());
}

template <int P>
static void m2l_cart_impl(Handle &h) {
  m2l_cart_arrays<P>(h, M2LList{h.m2l_tgt.p, h.nm2l_tgt, h.m2l_off.p, h.m2l_src.p, h.m2l_img.p}, h.cX.p, h.M.p,
                     h.L.p);
}

// ---------------------------------------------------------------------------
// traversePeriodic (traversal.h:127-185) in the Cartesian basis.
template <int P>
__global__ void __launch_bounds__(64)
    k_periodic_cart_m2l(const double4 *__restrict__ cX, const cplx *__restrict__ M, const cplx *__restrict__ Mp,
                        double cycle, int mass, cplx *__restrict__ part) {
  constexpr int NT = Cart<P>::NT;
  constexpr int NO = (NT + 63) / 64;
  __shared__ double sC[NT];
  __shared__ double sM[NT];
  const int lane = threadIdx.x;
  const int r = blockIdx.x < 13 ? blockIdx.x : blockIdx.x + 1;
  const int ix = r / 9 - 1, iy = (r / 3) % 3 - 1, iz = r % 3 - 1;
  for (int t = lane; t < NT; t += 64) sM[t] = (t == 0 && mass) ? 1.0 : Mp[t].re;
  const double mi0 = M[0].re, mj0 = Mp[0].re;
  double acc[NO];
  for (int o = 0; o < NO; o++) acc[o] = 0;
  const double4 X = cX[0];
  for (int cx = -1; cx <= 1; cx++)
    for (int cy = -1; cy <= 1; cy++)
      for (int cz = -1; cz <= 1; cz++) {
        const double dx = X.x - X.x - (ix * 3 + cx) * cycle, dy = X.y - X.y - (iy * 3 + cy) * cycle,
                     dz = X.z - X.z - (iz * 3 + cz) * cycle;
        __syncthreads();
        if (lane == 0) {
          double r2 = 0;
          r2 += dx * dx;
          r2 += dy * dy;
          r2 += dz * dz;
          const double invR2 = 1 / r2;
          const double invR = mass ? mi0 * mj0 * sqrt(invR2) : sqrt(invR2);
          cart_coef<P>(sC, dx, dy, dz, invR2, invR);
        }
        __syncthreads();
        for (int o = 0; o < NO; o++) {
          const int a = lane + 64 * o;
          if (a >= NT) continue;
          int xa, ya, za;
          cmi(a, xa, ya, za);
          const int da = xa + ya + za;
          double sum = 0;
          for (int b = 0; b < NT; b++) {
            int xb, yb, zb;
            cmi(b, xb, yb, zb);
            if (m2l_term(P, da, xb + yb + zb)) sum += sM[b] * sC[cidx(xa + xb, ya + yb, za + zb)];
          }
          acc[o] += sum;
        }
      }
  for (int o = 0; o < NO; o++) {
    const int a = lane + 64 * o;
    if (a < NT) part[(size_t)blockIdx.x * NT + a] = {acc[o], 0.0};
  }
}

template <int P>
__global__ void __launch_bounds__(64)
    k_periodic_cart_update(const double4 *__restrict__ cX, cplx *__restrict__ L, const cplx *__restrict__ part,
                           cplx *__restrict__ Mp, double cycle, int mass, int last) {
#pragma clang fp contract(off)
  constexpr int NT = Cart<P>::NT;
  constexpr int NO = (NT + 63) / 64;
  __shared__ double sM[NT];
  __shared__ double sC[NT];
  __shared__ double sN[NT];
  const int lane = threadIdx.x;
  for (int t = lane; t < NT; t += 64) {
    cplx l = L[t];
    for (int b = 0; b < 26; b++) l.re += part[(size_t)b * NT + t].re;
    L[t] = l;
  }
  if (last) return;
  const double m0 = Mp[0].re;
  for (int t = lane; t < NT; t += 64) sM[t] = (mass && t > 0) ? Mp[t].re * m0 : Mp[t].re;
  double acc[NO];
  int ka[NO], kb[NO], kc[NO];
  for (int o = 0; o < NO; o++) {
    acc[o] = 0;
    const int t = lane + 64 * o;
    cmi(t < NT ? t : 0, ka[o], kb[o], kc[o]);
  }
  const double4 X = cX[0];
  for (int ix = -1; ix <= 1; ix++)
    for (int iy = -1; iy <= 1; iy++)
      for (int iz = -1; iz <= 1; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        const double dx = X.x - (X.x + ix * cycle), dy = X.y - (X.y + iy * cycle), dz = X.z - (X.z + iz * cycle);
        __syncthreads();
        for (int t = lane; t < NT; t += 64) {
          int a, b, cz;
          cmi(t, a, b, cz);
          sC[t] = cpow(1.0, a, b, cz, dx, dy, dz);
        }
        __syncthreads();
        for (int o = 0; o < NO; o++) {
          const int t = lane + 64 * o;
          if (t >= NT) continue;
          acc[o] += sC[t] * sM[0];
          if (t == 0) continue;
          double sum = 0;
          for (int a = 0; a <= ka[o]; a++)
            for (int b = 0; b <= kb[o]; b++)
              for (int cz = 0; cz <= kc[o]; cz++) {
                if (a == 0 && b == 0 && cz == 0) continue;
                sum += sC[cidx(ka[o] - a, kb[o] - b, kc[o] - cz)] * sM[cidx(a, b, cz)];
              }
          acc[o] += sum;
        }
      }
  for (int o = 0; o < NO; o++) {
    const int t = lane + 64 * o;
    if (t < NT) sN[t] = acc[o];
  }
  __syncthreads();
  const double n0 = sN[0];
  for (int t = lane; t < NT; t += 64) Mp[t] = {(mass && t > 0) ? sN[t] / n0 : sN[t], 0.0};
}

template <int P>
static void periodic_cart_impl(Handle &h) {
  constexpr int NT = Cart<P>::NT;
  h.per_M.resize(NT);
  h.per_part.resize(26 * NT);
  EXA_HIP(hipMemcpyAsync(h.per_M.p, h.M.p, NT * sizeof(cplx), hipMemcpyDeviceToDevice, h.stream));
  double cycle = h.cfg.cycle;
  for (int lev = 0; lev < h.cfg.images - 1; lev++) {
    hipLaunchKernelGGL((k_periodic_cart_m2l<P>), dim3(26), dim3(64), 0, h.stream, h.cX.p, h.M.p, h.per_M.p, cycle,
                       h.cfg.mass, h.per_part.p);
    hipLaunchKernelGGL((k_periodic_cart_update<P>), dim3(1), dim3(64), 0, h.stream, h.cX.p, h.L.p, h.per_part.p,
                       h.per_M.p, cycle, h.cfg.mass, lev == h.cfg.images - 2 ? 1 : 0);
    cycle *= 3;
  }
  EXA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
template <int P>
static void upward_cart_impl(Handle &h) {
  for (int t = h.nlevels - 1; t >= 0; t--) {
    const int cnt = h.lvl_off[t + 1] - h.lvl_off[t];
    if (cnt == 0) continue;
    hipLaunchKernelGGL(k_upward_cart<P>, dim3(cnt), dim3(64), 0, h.stream, h.lvl_cells.p + h.lvl_off[t], h.body4.p,
                       h.cX.p, h.cRb.p, h.cChild.p, h.cNChild.p, h.cBody.p, h.cNBody.p, h.M.p, h.cfg.theta,
                       h.cfg.useRmax);
  }
  EXA_HIP(hipGetLastError());
}

template <int P>
static void downward_cart_impl(Handle &h) {
  for (int t = 0; t < h.nlevels; t++) {
    const int cnt = h.lvl_off[t + 1] - h.lvl_off[t];
    if (cnt == 0) continue;
    hipLaunchKernelGGL(k_downward_cart<P>, dim3(cnt), dim3(64), 0, h.stream, h.lvl_cells.p + h.lvl_off[t],
                       h.body4.p, h.cX.p, h.cParent.p, h.cNChild.p, h.cBody.p, h.cNBody.p, h.M.p, h.L.p, h.trg4.p,
                       t > 0 ? 1 : 0, h.cfg.mass);
  }
  EXA_HIP(hipGetLastError());
}

#define EXA_DISPATCH_CART(fn, h)                                                  \
  switch ((h).cfg.P) {                                                            \
    case 4: fn<4>(h); break;                                                      \
    case 6: fn<6>(h); break;                                                      \
    case 8: fn<8>(h); break;                                                      \
    default: throw Error(EXAFMM_HIP_EUNSUP, "Cartesian P must be one of 4, 6, 8"); \
  }

void run_upward_cart(Handle &h) { EXA_DISPATCH_CART(upward_cart_impl, h); }
void run_downward_cart(Handle &h) { EXA_DISPATCH_CART(downward_cart_impl, h); }
void run_m2l_cart(Handle &h) {
  if (h.nm2l_tgt == 0) return;
  EXA_DISPATCH_CART(m2l_cart_impl, h);
}
void run_periodic_cart(Handle &h) {
  if (h.cfg.images <= 0 || h.ncells == 0) return;
  EXA_DISPATCH_CART(periodic_cart_impl, h);
}

// The Cartesian pass kernels on a caller-built cell set (per-cell operators).
template <int P>
static void upward_cart_cells_impl(Handle &h, const CellSetView &v) {
  hipLaunchKernelGGL(k_upward_cart<P>, dim3(v.ncell), dim3(64), 0, h.stream, v.cells, v.body, v.cX, v.cRb, v.cChild,
                     v.cNChild, v.cBody, v.cNBody, v.M, h.cfg.theta, h.cfg.useRmax);
}

template <int P>
static void downward_cart_cells_impl(Handle &h, const CellSetView &v) {
  hipLaunchKernelGGL(k_downward_cart<P>, dim3(v.ncell), dim3(64), 0, h.stream, v.cells, v.body, v.cX, v.cParent,
                     v.cNChild, v.cBody, v.cNBody, v.M, v.L, v.trg, v.do_l2l, h.cfg.mass);
}

#define EXA_DISPATCH_CART2(fn, h, a)                                              \
  switch ((h).cfg.P) {                                                            \
    case 4: fn<4>(h, a); break;                                                   \
    case 6: fn<6>(h, a); break;                                                   \
    case 8: fn<8>(h, a); break;                                                   \
    default: throw Error(EXAFMM_HIP_EUNSUP, "Cartesian P must be one of 4, 6, 8"); \
  }

void launch_upward_cart_cells(Handle &h, const CellSetView &v) {
  EXA_DISPATCH_CART2(upward_cart_cells_impl, h, v);
  EXA_HIP(hipGetLastError());
}
void launch_downward_cart_cells(Handle &h, const CellSetView &v) {
  EXA_DISPATCH_CART2(downward_cart_cells_impl, h, v);
  EXA_HIP(hipGetLastError());
}
void launch_m2l_cart_arrays(Handle &h, const M2LList &lst, const double4 *cX, const cplx *M, cplx *L) {
  switch (h.cfg.P) {
    case 4: m2l_cart_arrays<4>(h, lst, cX, M, L); break;
    case 6: m2l_cart_arrays<6>(h, lst, cX, M, L); break;
    case 8: m2l_cart_arrays<8>(h, lst, cX, M, L); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "Cartesian P must be one of 4, 6, 8");
  }
}

}  // namespace exafmm
