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
// Work split: the (j, k, n) triples -- each a run of 2n+1 complex MACs --
// are packed onto the 64 lanes by longest-processing-time-first on the
// host, so every lane does ~1/64 of the pair's MACs (the per-output split
// would leave lane (0,0) with 100 MACs and lane (9,k) with 1 at P = 10).
// Per source the wavefront writes the P^2-entry tables Mhat and S into LDS,
// each lane runs its fixed item list (descriptors held in registers) and
// adds each finished triple into its own LDS slot; after the last source the
// slots of each output are summed in n order.
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

// Descriptor bits: [0,8) Mhat index, [8,16) S index, [16] negate,
// [17] end of triple, [18,27) triple id.  A padding item points both
// indices at the zero entry NF and never ends a triple.
static std::vector<uint32_t> build_schedule(int P, int ipl) {
  struct Triple {
    int j, k, n, id, work;
  };
  std::vector<Triple> tr;
  int id = 0;
  for (int j = 0; j < P; j++)
    for (int k = 0; k <= j; k++)
      for (int n = 0; n < P - j; n++) tr.push_back({j, k, n, id++, 2 * n + 1});
  std::vector<Triple> order = tr;
  std::stable_sort(order.begin(), order.end(), [](const Triple &a, const Triple &b) { return a.work > b.work; });
  std::vector<std::vector<int>> lanes(64);
  std::vector<int> load(64, 0);
  for (const Triple &t : order) {
    int best = 0;
    for (int l = 1; l < 64; l++)
      if (load[l] < load[best]) best = l;
    lanes[best].push_back(t.id);
    load[best] += t.work;
  }
  const int NF = P * P;
  std::vector<uint32_t> desc(64 * (size_t)ipl, (uint32_t)NF | ((uint32_t)NF << 8));
  for (int l = 0; l < 64; l++) {
    if (load[l] > ipl) throw Error(EXAFMM_HIP_EUNSUP, "M2L schedule exceeds the per-lane item budget");
    int it = 0;
    for (int tid : lanes[l]) {
      const Triple &t = tr[tid];
      for (int m = -t.n; m <= t.n; m++) {
        const int a = t.j + t.n, b = m - t.k;
        const uint32_t mi = t.n * t.n + t.n + m;
        const uint32_t yi = a * a + a + b;
        int sgn = (t.j & 1) ? -1 : 1;
        if (m >= 0) sgn *= (((t.k < m) ? t.k : m) & 1) ? -1 : 1;
        uint32_t d = mi | (yi << 8) | ((sgn < 0 ? 1u : 0u) << 16) | ((uint32_t)t.id << 18);
        if (m == t.n) d |= 1u << 17;
        desc[(size_t)it * 64 + l] = d;
        it++;
      }
    }
  }
  return desc;
}

template <int P>
__global__ void __launch_bounds__(64)
    k_m2l_f64(const int *__restrict__ tgt_list, const int *__restrict__ off, const int *__restrict__ src,
              const unsigned char *__restrict__ img, const double4 *__restrict__ cX, const cplx *__restrict__ M,
              cplx *__restrict__ L, const uint32_t *__restrict__ sched, double cycle, int mass) {
  using S = M2LShape<P>;
  constexpr int NT = Expn<P>::NT;
  constexpr int NF = S::NF;
  constexpr int IPL = S::IPL;
  __shared__ cplx sM[2][NF + 1];
  __shared__ cplx sS[2][NF + 1];
  __shared__ cplx seg[S::NTRIPLE];
  const int c = tgt_list[blockIdx.x];
  const int lane = threadIdx.x;
  uint32_t desc[IPL];
#pragma unroll
  for (int it = 0; it < IPL; it++) desc[it] = sched[it * 64 + lane];
  for (int t = lane; t < S::NTRIPLE; t += 64) seg[t] = {0, 0};
  if (lane == 0) {
    sM[0][NF] = sM[1][NF] = {0, 0};
    sS[0][NF] = sS[1][NF] = {0, 0};
  }
  const double4 Xi = cX[c];
  const double mi0 = mass ? M[(size_t)c * NT].re : 1.0;
  const int e0 = off[c], e1 = off[c + 1];
  for (int e = e0; e < e1; e++) {
    const int buf = (e - e0) & 1;
    const int s = src[e];
    const double4 Xj = cX[s];
    const int im = img[e];
    const double px = (im / 9 - 1) * cycle, py = ((im / 3) % 3 - 1) * cycle, pz = (im % 3 - 1) * cycle;
    const Geom g = make_geom(Xi.x - Xj.x - px, Xi.y - Xj.y - py, Xi.z - Xj.z - pz);
    for (int f = lane; f < NF; f += 64) {
      int a = 0;
      while ((a + 1) * (a + 1) <= f) a++;
      const int b = f - a * a - a;
      const int ab = b >= 0 ? b : -b;
      double re, imv;
      singular_nm(a, ab, g, re, imv);
      sS[buf][f] = {re, b >= 0 ? imv : -imv};
      cplx mv = M[(size_t)s * NT + tri(a) + ab];
      if (b < 0) mv.im = -mv.im;
      if (mass && f == 0) mv = {1.0, 0.0};
      sM[buf][f] = mv;
    }
    const double scale = mass ? mi0 * M[(size_t)s * NT].re : 1.0;
    __syncthreads();
    double ar = 0, ai = 0;
#pragma unroll
    for (int it = 0; it < IPL; it++) {
      const uint32_t d = desc[it];
      const cplx mv = sM[buf][d & 255];
      const cplx sv = sS[buf][(d >> 8) & 255];
      const double sg = (d >> 16) & 1 ? -1.0 : 1.0;
      const double mr = sg * mv.re, mim = sg * mv.im;
      ar += mr * sv.re - mim * sv.im;
      ai += mr * sv.im + mim * sv.re;
      if ((d >> 17) & 1) {
        const int tid = d >> 18;
        cplx t = seg[tid];
        t.re += scale * ar;
        t.im += scale * ai;
        seg[tid] = t;
        ar = ai = 0;
      }
    }
  }
  __syncthreads();
  for (int t = lane; t < NT; t += 64) {
    int j = 0;
    while (tri(j + 1) <= t) j++;
    const int k = t - tri(j);
    int base = 0;
    for (int jj = 0; jj < j; jj++) base += (jj + 1) * (P - jj);
    base += k * (P - j);
    double lr = 0, li = 0;
    for (int n = 0; n < P - j; n++) {
      lr += seg[base + n].re;
      li += seg[base + n].im;
    }
    cplx o = L[(size_t)c * NT + t];
    o.re += lr;
    o.im += li;
    L[(size_t)c * NT + t] = o;
  }
}

template <int P>
static void m2l_impl(Handle &h) {
  using S = M2LShape<P>;
  if (h.m2l_sched_P != P) {
    const std::vector<uint32_t> host = build_schedule(P, S::IPL);
    h.m2l_sched.resize(host.size());
    EXA_HIP(hipMemcpyAsync(h.m2l_sched.p, host.data(), host.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    h.m2l_sched_P = P;
  }
  hipLaunchKernelGGL(k_m2l_f64<P>, dim3(h.nm2l_tgt), dim3(64), 0, h.stream, h.m2l_tgt.p, h.m2l_off.p, h.m2l_src.p,
                     h.m2l_img.p, h.cX.p, h.M.p, h.L.p, h.m2l_sched.p, h.cfg.cycle, h.cfg.mass);
  EXA_HIP(hipGetLastError());
}

void run_m2l(Handle &h) {
  if (h.nm2l_tgt == 0) return;
  switch (h.cfg.P) {
    case 4: m2l_impl<4>(h); break;
    case 6: m2l_impl<6>(h); break;
    case 8: m2l_impl<8>(h); break;
    case 10: m2l_impl<10>(h); break;
    case 12: m2l_impl<12>(h); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
}

}  // namespace exafmm
