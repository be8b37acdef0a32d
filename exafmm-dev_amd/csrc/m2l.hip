// M2L over the CSR interaction list: one 64-lane workgroup per target cell.
//
// Reference: kernels/LaplaceSphericalCPU.cxx:149-191 (non-mutual; MASS
// variant :157-172).  For target coefficient (j, k):
//   L_j^k += sum_{n < P-j} sum_{m=-n}^{n} s(j,k,m) Mhat_n^m S_{j+n}^{m-k}(dX)
// with Mhat_n^m = M_n^m (m >= 0) or conj(M_n^{-m}) (m < 0), S the singular
// harmonics of dX = X_i - X_j - X_periodic and
//   s = (-1)^j                     (m < 0)
//   s = (-1)^j (-1)^(k<m ? k : m)  (m >= 0).
// That is 8 P (P+1)^2 (P+2) / 12 flops per pair (9,680 at P = 10).
//
// Folding the signs into the tables, M'_n^m = (-1)^m Mhat_n^m (m >= 0),
// M'_n^m = Mhat_n^m (m < 0), S'_a^b = (-1)^b S_a^b (b > 0), S'_a^b = S_a^b
// (b <= 0), leaves a plain correlation in k:
//   L_j^k += (-1)^j sum_n sum_m M'_n^m S'_{j+n}^{m-k}.
//
// Work split: the (j, k, n) triples -- each a run of 2n+1 complex MACs --
// are packed onto the 64 lanes by longest-processing-time-first on the
// host, so every lane does ~1/64 of the pair's MACs (the per-output split
// would leave lane (0,0) with 100 MACs and lane (9,k) with 1 at P = 10).
// Per source the wavefront writes the P^2-entry tables M' and S' into LDS
// (one packed coefficient per lane, division-free recurrences), each lane
// runs its fixed item list (descriptors held in registers) and adds each
// finished triple into its own LDS slot; after the last source the slots of
// each output are summed in n order.  The next source's multipole and
// centre are loaded while the current pair is contracted.
#include <algorithm>
#include <vector>

#include "harmonics.h"
#include "internal.h"

namespace exafmm {

template <int P>
struct M2LShape {
  static constexpr int NF = P * P;  // full (n, m) tables
  static constexpr int NTRIPLE = [] {
    int s = 0;
    for (int j = 0; j < P; j++) s += (j + 1) * (P - j);
    return s;
  }();
  static constexpr int TOTAL = [] {
    int s = 0;
    for (int j = 0; j < P; j++) s += (j + 1) * (P - j) * (P - j);
    return s;
  }();
  static constexpr int MAXTRIPLE = 2 * P - 1;
  static constexpr int IPL = (TOTAL + 63) / 64 > MAXTRIPLE ? (TOTAL + 63) / 64 + 3 : MAXTRIPLE + 1;
};

// Lane schedule.  Each lane owns a list of (j, k, n) triples (LPT packing by
// 2n+1 MACs); its items run m = -n..n of each triple in order.  Descriptor:
//   bits [0,12)  byte offset of M'_n^m in the M' table (16 B entries)
//   bits [12,24) byte offset of S'_{j+n}^{m-k} in the S' table
//   bit 24       last item of a triple: add the partial into the lane's next
//                accumulator slot
//   bits [25,29) j of the triple (selects the triple's scale factor)
// Padding items point at the zero entry NF of both tables.  Slots are laid
// out lane-major (lane l owns slots [slot0[l], slot0[l] + #triples(l))), and
// `outslot` lists, for every output (j, k), its slots in n order.
struct M2LSchedule {
  std::vector<uint32_t> desc;     // IPL x 64, item-major
  std::vector<int> slot0;         // 64
  std::vector<int> outslot;       // NT x P (-1 padded)
  int nslots = 0;
};

template <typename R>
struct cplxT {
  R re, im;
};

static M2LSchedule build_schedule(int P, int ipl, int esz) {
  struct Triple {
    int j, k, n, work;
  };
  std::vector<Triple> tr;
  for (int j = 0; j < P; j++)
    for (int k = 0; k <= j; k++)
      for (int n = 0; n < P - j; n++) tr.push_back({j, k, n, 2 * n + 1});
  std::vector<int> order(tr.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tr[a].work > tr[b].work; });
  std::vector<std::vector<int>> lanes(64);
  std::vector<int> load(64, 0);
  for (int t : order) {
    int best = 0;
    for (int l = 1; l < 64; l++)
      if (load[l] < load[best]) best = l;
    lanes[best].push_back(t);
    load[best] += tr[t].work;
  }
  const int NF = P * P, NT = P * (P + 1) / 2;
  M2LSchedule s;
  const uint32_t pad = (uint32_t)(NF * esz) | ((uint32_t)(NF * esz) << 12);
  s.desc.assign(64 * (size_t)ipl, pad);
  s.slot0.assign(64, 0);
  s.outslot.assign((size_t)NT * P, -1);
  int slot = 0;
  for (int l = 0; l < 64; l++) {
    if (load[l] > ipl) throw Error(EXAFMM_HIP_EUNSUP, "M2L schedule exceeds the per-lane item budget");
    s.slot0[l] = slot;
    int it = 0;
    for (int t : lanes[l]) {
      const Triple &q = tr[t];
      for (int m = -q.n; m <= q.n; m++) {
        const int a = q.j + q.n, b = m - q.k;
        uint32_t d = (uint32_t)((q.n * q.n + q.n + m) * esz) | ((uint32_t)((a * a + a + b) * esz) << 12);
        if (m == q.n) d |= (1u << 24) | ((uint32_t)q.j << 25);
        s.desc[(size_t)it * 64 + l] = d;
        it++;
      }
      s.outslot[(size_t)(q.j * (q.j + 1) / 2 + q.k) * P + q.n] = slot++;
    }
  }
  s.nslots = slot;
  return s;
}

// R = double: the plain FP64 path.  R = float: the mixed path -- centres,
// the offset geometry and the accumulators stay FP64; the harmonic tables and
// the per-triple contraction are FP32.  To keep the FP32 tables in range at
// any cell size, they hold the expansions scaled by the pair distance r:
//   M~_n^m = M_n^m / r^n,  S~_a^b = S_a^b r^(a+1)  (unit-distance harmonics),
// so L_j^k = r^-(j+1) sum_n sum_m M~ S~, and each finished (j, k, n) partial
// is multiplied in FP64 by u^(j+1) = r^-(j+1).
template <int P, typename R>
__global__ void __launch_bounds__(64)
    k_m2l(const int *__restrict__ tgt_list, const int *__restrict__ off, const int *__restrict__ src,
          const unsigned char *__restrict__ img, const double4 *__restrict__ cX, const cplx *__restrict__ M,
          cplx *__restrict__ L, const uint32_t *__restrict__ sched, const int *__restrict__ slot0,
          const int *__restrict__ outslot, double cycle, int mass) {
  using S = M2LShape<P>;
  using C = cplxT<R>;
  constexpr bool F32 = sizeof(R) == 4;
  constexpr int NT = Expn<P>::NT;
  constexpr int NF = S::NF;
  constexpr int IPL = S::IPL;
  constexpr int NO = (NT + 63) / 64;
  // Single-buffered tables (two barriers per pair) keep the LDS footprint
  // small enough for five or more resident waves per SIMD.
  __shared__ C sM[NF + 1];
  __shared__ C sS[NF + 1];
  __shared__ cplx seg[S::NTRIPLE];
  __shared__ double sU[P];  // FP32 path: scale * u^(j+1)
  const int c = tgt_list[blockIdx.x];
  const int lane = threadIdx.x;
  uint32_t desc[IPL];
#pragma unroll
  for (int it = 0; it < IPL; it++) desc[it] = sched[it * 64 + lane];
  const int myslot0 = slot0[lane];
  for (int t = lane; t < S::NTRIPLE; t += 64) seg[t] = {0, 0};
  if (lane == 0) {
    sM[NF] = {0, 0};
    sS[NF] = {0, 0};
  }
  // Packed coefficients (n, m >= 0) this lane translates: t = lane + 64 o.
  int tn[NO], tm[NO];
#pragma unroll
  for (int o = 0; o < NO; o++) {
    const int t = lane + 64 * o < NT ? lane + 64 * o : 0;
    int n = 0;
    while (tri(n + 1) <= t) n++;
    tn[o] = n;
    tm[o] = t - tri(n);
  }
  const double4 Xi = cX[c];
  const double mi0 = mass ? M[(size_t)c * NT].re : 1.0;
  const int e0 = off[c], e1 = off[c + 1];
  // Prefetch of the first source.
  int s_next = e0 < e1 ? src[e0] : 0;
  int im_next = e0 < e1 ? img[e0] : 13;
  cplx m_next[NO];
#pragma unroll
  for (int o = 0; o < NO; o++) m_next[o] = (e0 < e1 && lane + 64 * o < NT) ? M[(size_t)s_next * NT + lane + 64 * o] : cplx{0, 0};
  double4 X_next = cX[s_next];
  double mj0_next = mass ? M[(size_t)s_next * NT].re : 1.0;
  for (int e = e0; e < e1; e++) {
    const double4 Xj = X_next;
    const int im = im_next;
    cplx mcur[NO];
#pragma unroll
    for (int o = 0; o < NO; o++) mcur[o] = m_next[o];
    const double mj0 = mj0_next;
    if (e + 1 < e1) {  // issue the next pair's loads now
      s_next = src[e + 1];
      im_next = img[e + 1];
#pragma unroll
      for (int o = 0; o < NO; o++) m_next[o] = lane + 64 * o < NT ? M[(size_t)s_next * NT + lane + 64 * o] : cplx{0, 0};
      X_next = cX[s_next];
      if (mass) mj0_next = M[(size_t)s_next * NT].re;
    }
    const double px = (im / 9 - 1) * cycle, py = ((im / 3) % 3 - 1) * cycle, pz = (im % 3 - 1) * cycle;
    const double dx = Xi.x - Xj.x - px, dy = Xi.y - Xj.y - py, dz = Xi.z - Xj.z - pz;
    // Geometry without divisions: u = 1/r, w = 1/rho_xy.
    const double rxy2 = dx * dx + dy * dy;
    const double u = rsqrt(rxy2 + dz * dz);
    const double w = rxy2 > 0 ? rsqrt(rxy2) : 0.0;
    const double ct = dz * u, st = rxy2 * w * u;  // cos(alpha), sin(alpha)
    const double cp = rxy2 > 0 ? dx * w : 1.0, sp = dy * w;
    const double us = F32 ? 1.0 : u;  // the FP32 tables use unit distance
    const double ru = us, rsu = st * us, rcp = cp, rsp = sp, cu = ct * us, u2 = us * us;
    const double scale = mass ? mi0 * mj0 : 1.0;
    __syncthreads();  // the previous pair's items are done with the tables
    if (F32 && lane < P) {
      double f = scale * u;
      for (int i = 0; i < lane; i++) f *= u;
      sU[lane] = f;
    }
#pragma unroll
    for (int o = 0; o < NO; o++) {
      if (lane + 64 * o >= NT) continue;
      const int a = tn[o], b = tm[o];
      // S_a^b = (-1)^a (a-b)! u^(a+1) P_a^b(ct) e^{ib phi} = G_a e^{ib phi}, with
      // G_b = (2b-1)!! (st u)^b u and the division-free upward recurrence
      // G_a = -(2a-1) ct u G_{a-1} - (a+b-1)(a-b-1) u^2 G_{a-2}.
      // Built in FP64 for both paths and rounded once when stored.
      double g = ru, gm1 = 0, ec = 1, es = 0;
      for (int i = 1; i <= b; i++) {
        g *= rsu * (double)(2 * i - 1);
        const double t = ec * rcp - es * rsp;
        es = ec * rsp + es * rcp;
        ec = t;
      }
      for (int t = b + 1; t <= a; t++) {
        const double gn = -(double)(2 * t - 1) * cu * g - (double)((t + b - 1) * (t - b - 1)) * u2 * gm1;
        gm1 = g;
        g = gn;
      }
      const R sr = (R)(g * ec), si = (R)(g * es);
      // S'_a^b = (-1)^b S_a^b for b > 0, S'_a^{-b} = conj(S_a^b).
      const R sg = (b & 1) ? (R)-1 : (R)1;
      sS[a * a + a + b] = {sg * sr, sg * si};
      if (b > 0) sS[a * a + a - b] = {sr, -si};
      // M'_n^m = (-1)^m M_n^m (m >= 0), M'_n^{-m} = conj(M_n^m).
      cplx mv = mcur[o];
      if (mass && a == 0) mv = {1.0, 0.0};
      if (F32) {  // M~_a^b = M_a^b u^a
        double f = 1.0;
        for (int i = 0; i < a; i++) f *= u;
        mv.re *= f;
        mv.im *= f;
      }
      sM[a * a + a + b] = {sg * (R)mv.re, sg * (R)mv.im};
      if (b > 0) sM[a * a + a - b] = {(R)mv.re, -(R)mv.im};
    }
    __syncthreads();
    const char *mb = reinterpret_cast<const char *>(sM);
    const char *sb = reinterpret_cast<const char *>(sS);
    R ar = 0, ai = 0;
    int slot = myslot0;
#pragma unroll
    for (int it = 0; it < IPL; it++) {
      const uint32_t d = desc[it];
      const C mv = *reinterpret_cast<const C *>(mb + (d & 0xFFF));
      const C sv = *reinterpret_cast<const C *>(sb + ((d >> 12) & 0xFFF));
      ar = fma(mv.re, sv.re, ar);
      ar = fma(-mv.im, sv.im, ar);
      ai = fma(mv.re, sv.im, ai);
      ai = fma(mv.im, sv.re, ai);
      if (d & (1u << 24)) {
        const double f = F32 ? sU[d >> 25] : scale;
        cplx t = seg[slot];
        t.re += f * (double)ar;
        t.im += f * (double)ai;
        seg[slot] = t;
        slot++;
        ar = ai = 0;
      }
    }
  }
  __syncthreads();
  for (int t = lane; t < NT; t += 64) {
    int j = 0;
    while (tri(j + 1) <= t) j++;
    double lr = 0, li = 0;
    for (int n = 0; n < P - j; n++) {
      const cplx v = seg[outslot[t * P + n]];
      lr += v.re;
      li += v.im;
    }
    const double sgn = (j & 1) ? -1.0 : 1.0;
    cplx o = L[(size_t)c * NT + t];
    o.re += sgn * lr;
    o.im += sgn * li;
    L[(size_t)c * NT + t] = o;
  }
}

template <int P>
static void m2l_launch(Handle &h, const M2LList &lst, const double4 *cX, const cplx *M, cplx *L, bool f32,
                       bool timed) {
  using S = M2LShape<P>;
  const int esz = f32 ? 8 : 16;  // bytes of one table entry
  const int key = P * 100 + esz;
  if (h.m2l_sched_P != key) {
    const M2LSchedule sc = build_schedule(P, S::IPL, esz);
    std::vector<uint32_t> blob(sc.desc);
    for (int v : sc.slot0) blob.push_back((uint32_t)v);
    for (int v : sc.outslot) blob.push_back((uint32_t)v);
    h.m2l_sched.resize(blob.size());
    EXA_HIP(hipMemcpyAsync(h.m2l_sched.p, blob.data(), blob.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    h.m2l_sched_P = key;
  }
  const uint32_t *desc = h.m2l_sched.p;
  const int *slot0 = reinterpret_cast<const int *>(desc + 64 * S::IPL);
  const int *outslot = slot0 + 64;
  if (lst.ntgt == 0) return;
  if (timed) {
    EXA_HIP(hipEventRecord(h.ev[14], h.stream));
    h.pair_timed = true;
  }
  if (esz == 8)
    hipLaunchKernelGGL((k_m2l<P, float>), dim3(lst.ntgt), dim3(64), 0, h.stream, lst.tgt, lst.off, lst.src, lst.img,
                       cX, M, L, desc, slot0, outslot, h.cfg.cycle, h.cfg.mass);
  else
    hipLaunchKernelGGL((k_m2l<P, double>), dim3(lst.ntgt), dim3(64), 0, h.stream, lst.tgt, lst.off, lst.src, lst.img,
                       cX, M, L, desc, slot0, outslot, h.cfg.cycle, h.cfg.mass);
  if (timed) EXA_HIP(hipEventRecord(h.ev[15], h.stream));
  EXA_HIP(hipGetLastError());
}

template <int P>
static void m2l_impl(Handle &h, const M2LList &lst) {
  m2l_launch<P>(h, lst, h.cX.p, h.M.p, h.L.p, h.cfg.precision == 1, true);
}

void launch_m2l_arrays(Handle &h, const M2LList &lst, const double4 *cX, const cplx *M, cplx *L) {
  if (h.cfg.basis == 1) return launch_m2l_cart_arrays(h, lst, cX, M, L);
  switch (h.cfg.P) {
    case 4: m2l_launch<4>(h, lst, cX, M, L, false, false); break;
    case 6: m2l_launch<6>(h, lst, cX, M, L, false, false); break;
    case 8: m2l_launch<8>(h, lst, cX, M, L, false, false); break;
    case 10: m2l_launch<10>(h, lst, cX, M, L, false, false); break;
    case 12: m2l_launch<12>(h, lst, cX, M, L, false, false); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
}

void run_m2l(Handle &h) {
  if (h.nm2l == 0) {
    // No pair reads the scaled multipoles the split upward pass may have
    // queued: a later traversal must rebuild them for its own tree.
    h.lat_ms_ready = false;
    h.lat_begun = false;
    return;
  }
  if (h.cfg.basis == 1) {
    join_multipoles(h);
    return run_m2l_cart(h);
  }
  // Lattice pairs on the matrix cores first (from the raw list; lst becomes
  // the irregular pairs grouped by target); otherwise every pair per pair.
  h.pair_timed = false;
  M2LList lst{};
  if (!run_m2l_lattice(h, lst)) {
    ensure_m2l_grouped(h);
    lst = M2LList{h.m2l_tgt.p, h.nm2l_tgt, h.m2l_off.p, h.m2l_src.p, h.m2l_img.p};
  }
  join_multipoles(h);  // (the lattice path joined them before its contraction)
  h.m2l_last = lst;
  switch (h.cfg.P) {
    case 4: m2l_impl<4>(h, lst); break;
    case 6: m2l_impl<6>(h, lst); break;
    case 8: m2l_impl<8>(h, lst); break;
    case 10: m2l_impl<10>(h, lst); break;
    case 12: m2l_impl<12>(h, lst); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
}

// The M2L of the last run_m2l again on the current multipoles: the lattice
// contraction from its set-up, then the per-pair kernel on the pairs the last
// run gave it (no host waits; spherical basis).
void run_m2l_replay(Handle &h) {
  if (h.nm2l == 0) return;
  lattice_replay(h);
  const bool f32 = h.cfg.precision == 1;
  switch (h.cfg.P) {
    case 4: m2l_launch<4>(h, h.m2l_last, h.cX.p, h.M.p, h.L.p, f32, false); break;
    case 6: m2l_launch<6>(h, h.m2l_last, h.cX.p, h.M.p, h.L.p, f32, false); break;
    case 8: m2l_launch<8>(h, h.m2l_last, h.cX.p, h.M.p, h.L.p, f32, false); break;
    case 10: m2l_launch<10>(h, h.m2l_last, h.cX.p, h.M.p, h.L.p, f32, false); break;
    case 12: m2l_launch<12>(h, h.m2l_last, h.cX.p, h.M.p, h.L.p, f32, false); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
}

}  // namespace exafmm
