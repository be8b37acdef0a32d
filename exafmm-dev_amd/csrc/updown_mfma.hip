// M2M and L2L between a box and its children as batched products on the
// matrix cores.
//
// Reference: kernels/LaplaceSphericalCPU.cxx:120-147 (M2M) and :193-224
// (L2L), applied per cell by include/up_down_pass.h:45-148.  A child box's
// centre sits at X_parent + R_child u with u = (+-1, +-1, +-1), one of 8
// octant vectors.  The regular harmonics are homogeneous, R_n(s x) =
// s^n R_n(x), so with expansions scaled by the child radius,
//   M2M:  M_p(j) = R^j  sum_o A_o (M_c(o)(n) R^-n),
//   L2L:  L_c(j) = R^-j B_o (L_p(n) R^n),
// where A_o (110 x 110 real for P = 10) and B_o are the reference's operators
// evaluated at R = 1.  They are built once per handle on the device by
// applying the same per-coefficient code (m2m_lane, l2l_lane) to unit
// vectors, and stored in MFMA fragment order.  Each workgroup takes 16
// parent cells of one level (the MFMA columns).  M2M: 4 waves, wave w
// handling octants 2w and 2w + 1, partials summed in a fixed order
// (deterministic).  L2L: 8 waves, wave o writing the rows of the children in
// octant o.  Zero blocks (M2M: child degree > parent degree; L2L: parent
// degree < child degree) are skipped at compile time.
#include "harmonics.h"
#include "internal.h"

namespace exafmm {

typedef double v4d_ud __attribute__((ext_vector_type(4)));

template <int P>
struct UDShape {
  static constexpr int NT = P * (P + 1) / 2;
  static constexpr int NR = 2 * NT;
  static constexpr int RB = (NR + 15) / 16;  // row blocks of one octant
  static constexpr int KC = (NR + 3) / 4;    // column chunks of one octant
  static constexpr int deg_of_real(int q) {  // degree of real index q
    int t = q / 2, n = 0;
    while ((n + 1) * (n + 2) / 2 <= t) n++;
    return n;
  }
  // M2M: row block rb needs input columns of degree <= its largest degree.
  static constexpr int m2m_kend(int rb) {
    const int last = (rb * 16 + 15 < NR ? rb * 16 + 15 : NR - 1);
    const int j = deg_of_real(last);
    return (2 * ((j + 1) * (j + 2) / 2) + 3) / 4;
  }
  // L2L: row block rb needs input columns of degree >= its smallest degree.
  static constexpr int l2l_kbeg(int rb) {
    const int j = deg_of_real(rb * 16);
    return (2 * (j * (j + 1) / 2)) / 4;
  }
  static constexpr int OPS = 8 * RB * KC * 64;  // doubles per operator
};

// Degree of every real index, for runtime lookups.
template <int P>
struct DegTab {
  int d[UDShape<P>::KC * 4 + 16];
  constexpr DegTab() : d() {
    for (int q = 0; q < UDShape<P>::KC * 4 + 16; q++) d[q] = q < UDShape<P>::NR ? UDShape<P>::deg_of_real(q) : 0;
  }
};

__device__ inline void octant_unit(int o, double &ux, double &uy, double &uz) {
  ux = (o & 1) ? 1.0 : -1.0;
  uy = (o & 2) ? 1.0 : -1.0;
  uz = (o & 4) ? 1.0 : -1.0;
}

// acc += Op * b for one octant's operator in fragment order (row blocks
// rb, column chunks [lo(rb), hi(rb)) that are not identically zero).  The
// next row block's operator entries are loaded while the current block's
// MFMAs run, so each wave waits on memory once per operator, not per MFMA.
template <int P, bool M2M>
__device__ inline void op_apply(const double *__restrict__ Op, int lane, const double (&b)[UDShape<P>::KC],
                                v4d_ud (&acc)[UDShape<P>::RB]) {
  using U = UDShape<P>;
  constexpr int KC = U::KC, RB = U::RB;
  auto lo = [](int rb) { return M2M ? 0 : U::l2l_kbeg(rb); };
  auto hi = [](int rb) { return M2M ? U::m2m_kend(rb) : KC; };
  double cur[KC], nxt[KC];
#pragma unroll
  for (int kc = lo(0); kc < hi(0); kc++) cur[kc] = Op[kc * 64 + lane];
#pragma unroll
  for (int rb = 0; rb < RB; rb++) {
    if (rb + 1 < RB) {
#pragma unroll
      for (int kc = lo(rb + 1); kc < hi(rb + 1); kc++) nxt[kc] = Op[((rb + 1) * KC + kc) * 64 + lane];
    }
#pragma unroll
    for (int kc = lo(rb); kc < hi(rb); kc++)
      acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur[kc], b[kc], acc[rb], 0, 0, 0);
#pragma unroll
    for (int kc = 0; kc < KC; kc++) cur[kc] = nxt[kc];
  }
}

// Column q (real input index) of A_o (op 0) or B_o (op 1): grid (NR, 8, 2),
// one 64-lane workgroup each, lanes own output coefficients.  The operator
// buffers are zero-filled beforehand (padding rows and columns stay zero).
template <int P>
__global__ void __launch_bounds__(64) k_ud_ops(double *__restrict__ A, double *__restrict__ B) {
  using U = UDShape<P>;
  constexpr int NT = U::NT;
  __shared__ cplx sIn[NT];
  __shared__ cplx sY[NT];
  const int q = blockIdx.x, o = blockIdx.y, op = blockIdx.z, lane = threadIdx.x;
  double ux, uy, uz;
  octant_unit(o, ux, uy, uz);
  // M2M shifts by X_parent - X_child = -u, L2L by X_child - X_parent = u.
  const Geom g = op == 0 ? make_geom(-ux, -uy, -uz) : make_geom(ux, uy, uz);
  for (int t = lane; t < NT; t += 64) {
    int n, m;
    coef_nm(t, n, m);
    double re, im;
    regular_nm(n, m, g, re, im);
    sY[t] = {re, im};
    sIn[t] = {t == q / 2 && (q & 1) == 0 ? 1.0 : 0.0, t == q / 2 && (q & 1) == 1 ? 1.0 : 0.0};
  }
  __syncthreads();
  double *D = op == 0 ? A : B;
  for (int t = lane; t < NT; t += 64) {
    int j, k;
    coef_nm(t, j, k);
    double lr = 0, li = 0;
    if (op == 0)
      m2m_lane(j, k, sIn, sY, lr, li);
    else
      l2l_lane<P>(j, k, sIn, sY, lr, li);
    for (int part = 0; part < 2; part++) {
      const int row = 2 * t + part;
      // Fragment order: block (rb, kc), lane = row % 16 + 16 (q % 4).
      const size_t idx = (((size_t)o * U::RB + row / 16) * U::KC + q / 4) * 64 + (row % 16) + 16 * (q % 4);
      D[idx] = part == 0 ? lr : li;
    }
  }
}

// Per-column set-up shared by both products: the children of each of the
// block's 16 parent cells by octant, and powers of the child radius.
template <int P>
struct UDCols {
  int ch[16][8];
  int cell[16];
  int zero[16];           // child radius 0 (coincident bodies): identity shift
  double pw[16][P + 1];   // R^i
  double ipw[16][P + 1];  // R^-i
};

template <int P>
__device__ inline void ud_setup(UDCols<P> &s, const int *__restrict__ cells, int ncell, const double4 *__restrict__ cX,
                                const double *__restrict__ cRb, const int *__restrict__ cChild,
                                const int *__restrict__ cNChild) {
  const int tid = threadIdx.x;
  if (tid < 128) s.ch[tid >> 3][tid & 7] = -1;
  __syncthreads();
  if (tid < 16) {
    const int i = blockIdx.x * 16 + tid;
    const int p = i < ncell ? cells[i] : -1;
    double r = 0;
    if (p >= 0) {
      const double4 Xp = cX[p];
      const int ch0 = cChild[p], nch = cNChild[p];
      for (int a = 0; a < nch; a++) {
        const int c = ch0 + a;
        const double4 Xc = cX[c];
        const int o = (Xc.x > Xp.x ? 1 : 0) | (Xc.y > Xp.y ? 2 : 0) | (Xc.z > Xp.z ? 4 : 0);
        s.ch[tid][o] = c;
        r = cRb[c];
      }
    }
    s.cell[tid] = p >= 0 && cNChild[p] > 0 ? p : -1;  // -1: no column (empty slot or leaf)
    s.zero[tid] = !(r > 0);
    double f = 1, fi = 1;
    const double ir = r > 0 ? 1.0 / r : 0.0;  // zero-radius columns are summed directly
    for (int k = 0; k <= P; k++) {
      s.pw[tid][k] = f;
      s.ipw[tid][k] = fi;
      f *= r;
      fi *= ir;
    }
  }
  __syncthreads();
}

template <int P>
__global__ void __launch_bounds__(256)
    k_m2m_mfma(const int *__restrict__ cells, int ncell, const double4 *__restrict__ cX, const double *__restrict__ cRb,
               const int *__restrict__ cChild, const int *__restrict__ cNChild, const double *__restrict__ A,
               cplx *__restrict__ M) {
  using U = UDShape<P>;
  constexpr int NR = U::NR, RB = U::RB, KC = U::KC;
  constexpr DegTab<P> deg{};
  __shared__ UDCols<P> s;
  __shared__ double part[2][RB * 16][16];
  ud_setup<P>(s, cells, ncell, cX, cRb, cChild, cNChild);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, kq = lane >> 4;
  double *Mr = reinterpret_cast<double *>(M);
  v4d_ud acc[RB];
#pragma unroll
  for (int rb = 0; rb < RB; rb++) acc[rb] = v4d_ud{0, 0, 0, 0};
#pragma unroll 1
  for (int oo = 0; oo < 2; oo++) {
    const int o = 2 * wave + oo;
    const int c = s.ch[col][o];
    const double *src = Mr + (size_t)(c < 0 ? 0 : c) * NR;
    double b[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
      const int k = 4 * kc + kq;
      const double v = src[k < NR ? k : NR - 1] * s.ipw[col][deg.d[k]];
      b[kc] = (c >= 0 && k < NR) ? v : 0.0;
    }
    op_apply<P, true>(A + (size_t)o * RB * KC * 64, lane, b, acc);
  }
  // Octant partials summed as (w0 + w2) + (w1 + w3).
  auto stash = [&](int slot) {
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) part[slot][rb * 16 + (lane >> 4) + 4 * r][col] = acc[rb][r];
  };
  auto absorb = [&](int slot) {
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[rb][r] += part[slot][rb * 16 + (lane >> 4) + 4 * r][col];
  };
  if (wave >= 2) stash(wave - 2);
  __syncthreads();
  if (wave < 2) absorb(wave);
  __syncthreads();
  if (wave == 1) stash(0);
  __syncthreads();
  if (wave != 0) return;
  absorb(0);
  const int p = s.cell[col];
  if (p < 0) return;
#pragma unroll
  for (int rb = 0; rb < RB; rb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = rb * 16 + (lane >> 4) + 4 * r;
      if (row >= NR) continue;
      double v = acc[rb][r] * s.pw[col][deg.d[row]];
      if (s.zero[col]) {  // X_child = X_parent: the shift is the identity
        v = 0;
        for (int o = 0; o < 8; o++)
          if (s.ch[col][o] >= 0) v += Mr[(size_t)s.ch[col][o] * NR + row];
      }
      Mr[(size_t)p * NR + row] = v;
    }
}

template <int P>
__global__ void __launch_bounds__(512)
    k_l2l_mfma(const int *__restrict__ cells, int ncell, const double4 *__restrict__ cX, const double *__restrict__ cRb,
               const int *__restrict__ cChild, const int *__restrict__ cNChild, const double *__restrict__ B,
               cplx *__restrict__ L) {
  using U = UDShape<P>;
  constexpr int NR = U::NR, RB = U::RB, KC = U::KC;
  constexpr DegTab<P> deg{};
  __shared__ UDCols<P> s;
  ud_setup<P>(s, cells, ncell, cX, cRb, cChild, cNChild);
  const int tid = threadIdx.x, o = tid >> 6, lane = tid & 63;  // wave o: octant o
  const int col = lane & 15, kq = lane >> 4;
  double *Lr = reinterpret_cast<double *>(L);
  const int p = s.cell[col];
  const double *src = Lr + (size_t)(p < 0 ? 0 : p) * NR;
  double b[KC];
#pragma unroll
  for (int kc = 0; kc < KC; kc++) {
    const int k = 4 * kc + kq;
    const double v = src[k < NR ? k : NR - 1] * s.pw[col][deg.d[k]];
    b[kc] = (p >= 0 && k < NR) ? v : 0.0;
  }
  v4d_ud acc[RB];
#pragma unroll
  for (int rb = 0; rb < RB; rb++) acc[rb] = v4d_ud{0, 0, 0, 0};
  op_apply<P, false>(B + (size_t)o * RB * KC * 64, lane, b, acc);
  const int c = s.ch[col][o];
  if (c < 0 || p < 0) return;
#pragma unroll
  for (int rb = 0; rb < RB; rb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = rb * 16 + (lane >> 4) + 4 * r;
      if (row < NR) Lr[(size_t)c * NR + row] += s.zero[col] ? src[row] : acc[rb][r] * s.ipw[col][deg.d[row]];
    }
}

// L2L of a level with few parents: one workgroup per (16 parents, row
// block), so a level of a few cells still spreads over RB workgroups instead
// of running its whole operator product in one (the small levels took about
// 30 us each, bound by one workgroup's operator loads and MFMA chain).
template <int P, int R>
__device__ inline void l2l_row_block(const double *__restrict__ Op, int lane, const double (&b)[UDShape<P>::KC],
                                     v4d_ud &acc) {
  using U = UDShape<P>;
  constexpr int KC = U::KC, K0 = U::l2l_kbeg(R);
  double a[KC - K0];
#pragma unroll
  for (int kc = K0; kc < KC; kc++) a[kc - K0] = Op[(R * KC + kc) * 64 + lane];
#pragma unroll
  for (int kc = K0; kc < KC; kc++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kc - K0], b[kc], acc, 0, 0, 0);
}

template <int P, int R = 0>
__device__ inline void l2l_row_block_of(int rb, const double *__restrict__ Op, int lane,
                                        const double (&b)[UDShape<P>::KC], v4d_ud &acc) {
  if constexpr (R < UDShape<P>::RB) {
    if (rb == R) return l2l_row_block<P, R>(Op, lane, b, acc);
    l2l_row_block_of<P, R + 1>(rb, Op, lane, b, acc);
  }
}

template <int P>
__global__ void __launch_bounds__(512)
    k_l2l_mfma_rb(const int *__restrict__ cells, int ncell, const double4 *__restrict__ cX,
                  const double *__restrict__ cRb, const int *__restrict__ cChild, const int *__restrict__ cNChild,
                  const double *__restrict__ B, cplx *__restrict__ L) {
  using U = UDShape<P>;
  constexpr int NR = U::NR, KC = U::KC;
  constexpr DegTab<P> deg{};
  __shared__ UDCols<P> s;
  ud_setup<P>(s, cells, ncell, cX, cRb, cChild, cNChild);
  const int tid = threadIdx.x, o = tid >> 6, lane = tid & 63;  // wave o: octant o
  const int col = lane & 15, kq = lane >> 4, rb = blockIdx.y;
  double *Lr = reinterpret_cast<double *>(L);
  const int p = s.cell[col];
  const double *src = Lr + (size_t)(p < 0 ? 0 : p) * NR;
  double b[KC];
#pragma unroll
  for (int kc = 0; kc < KC; kc++) {
    const int k = 4 * kc + kq;
    const double v = src[k < NR ? k : NR - 1] * s.pw[col][deg.d[k]];
    b[kc] = (p >= 0 && k < NR) ? v : 0.0;
  }
  v4d_ud acc = v4d_ud{0, 0, 0, 0};
  l2l_row_block_of<P>(rb, B + (size_t)o * U::RB * KC * 64, lane, b, acc);
  const int c = s.ch[col][o];
  if (c < 0 || p < 0) return;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int row = rb * 16 + (lane >> 4) + 4 * r;
    if (row < NR) Lr[(size_t)c * NR + row] += s.zero[col] ? src[row] : acc[r] * s.ipw[col][deg.d[row]];
  }
}

__global__ void k_fill_f64(double *__restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Built once per handle, on the stream of the first product that needs
// them (later users on the other stream are ordered after it by the
// evaluation's stream joins).
template <int P>
static void ensure_ops(Handle &h, hipStream_t st) {
  using U = UDShape<P>;
  if (h.ud_ready) return;
  h.ud_ops.resize(2 * (size_t)U::OPS);
  hipLaunchKernelGGL(k_fill_f64, dim3(div_up(2 * (int64_t)U::OPS, 256)), dim3(256), 0, st, h.ud_ops.p,
                     2 * (int64_t)U::OPS);
  hipLaunchKernelGGL(k_ud_ops<P>, dim3(U::NR, 8, 2), dim3(64), 0, st, h.ud_ops.p, h.ud_ops.p + U::OPS);
  EXA_HIP(hipGetLastError());
  h.ud_ready = true;
}

template <int P>
static void m2m_level(Handle &h, const int *cells, int cnt, hipStream_t st) {
  ensure_ops<P>(h, st);
  hipLaunchKernelGGL(k_m2m_mfma<P>, dim3(div_up(cnt, 16)), dim3(256), 0, st, cells, cnt, h.cX.p, h.cRb.p,
                     h.cChild.p, h.cNChild.p, h.ud_ops.p, h.M.p);
}

template <int P>
static void l2l_level(Handle &h, const int *cells, int cnt) {
  using U = UDShape<P>;
  ensure_ops<P>(h, h.stream);
  // Levels of fewer than kL2LSplit groups of 16 parents: one workgroup per
  // (group, row block).
  constexpr int kL2LSplit = 64;
  if (div_up(cnt, 16) < kL2LSplit)
    hipLaunchKernelGGL(k_l2l_mfma_rb<P>, dim3(div_up(cnt, 16), U::RB), dim3(512), 0, h.stream, cells, cnt, h.cX.p,
                       h.cRb.p, h.cChild.p, h.cNChild.p, h.ud_ops.p + U::OPS, h.L.p);
  else
    hipLaunchKernelGGL(k_l2l_mfma<P>, dim3(div_up(cnt, 16)), dim3(512), 0, h.stream, cells, cnt, h.cX.p, h.cRb.p,
                       h.cChild.p, h.cNChild.p, h.ud_ops.p + U::OPS, h.L.p);
}

// Periodic runs keep the per-cell kernels: the image sums cancel strongly
// and amplify the operator form's last-bit differences in the root
// multipole (3e-12 relative at 4 image levels), above the 1e-12 held
// against the reference there.
bool updown_mfma_enabled(const Handle &h) { return h.cfg.basis == 0 && h.cfg.images == 0 && h.updown_mfma; }

void run_m2m_level_mfma(Handle &h, const int *cells, int cnt, hipStream_t st) {
  switch (h.cfg.P) {
    case 4: m2m_level<4>(h, cells, cnt, st); break;
    case 6: m2m_level<6>(h, cells, cnt, st); break;
    case 8: m2m_level<8>(h, cells, cnt, st); break;
    case 10: m2m_level<10>(h, cells, cnt, st); break;
    case 12: m2m_level<12>(h, cells, cnt, st); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
  EXA_HIP(hipGetLastError());
}

void run_l2l_level_mfma(Handle &h, const int *cells, int cnt) {
  switch (h.cfg.P) {
    case 4: l2l_level<4>(h, cells, cnt); break;
    case 6: l2l_level<6>(h, cells, cnt); break;
    case 8: l2l_level<8>(h, cells, cnt); break;
    case 10: l2l_level<10>(h, cells, cnt); break;
    case 12: l2l_level<12>(h, cells, cnt); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
  EXA_HIP(hipGetLastError());
}

}  // namespace exafmm
