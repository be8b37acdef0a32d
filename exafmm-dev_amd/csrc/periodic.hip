// Periodic far field and the dipole correction.
//
// traversePeriodic (include/traversal.h:127-185): the 27 nearest images of
// the tree are part of the interaction lists; beyond them, for each of
// images-1 sublevels, the root's local expansion receives M2L from the 26 x 27
// images at offsets (3 i + c) cycle, i != 0, of a periodic parent whose
// multipole starts as the root's; the parent's multipole then becomes the M2M
// of its 26 neighbour copies (the centre copy is not a child of the periodic
// parent in the reference, NCHILD = 26), and cycle triples.  MASS mode
// denormalises around the M2M and finally divides the root's L by M0.
//
// Work: 702 M2L per sublevel, into one cell.  One workgroup per ring cell
// (26) does its 27 translations; a one-workgroup kernel adds the 26 partial
// local expansions in ring order and forms the next sublevel's multipole.
//
// getDipole / dipoleCorrection (include/up_down_pass.h:151-168).
#include <cmath>

#include "harmonics.h"
#include "internal.h"

namespace exafmm {

// Partial local expansion of ring cell blockIdx.x (ix, iy, iz != 0) at this
// sublevel: sum over its 27 children of M2L(root <- parent image at Xp).
template <int P>
__global__ void __launch_bounds__(64)
    k_periodic_m2l(const double4 *__restrict__ cX, const cplx *__restrict__ M, const cplx *__restrict__ Mp,
                   double cycle, int mass, cplx *__restrict__ part) {
  constexpr int NT = Expn<P>::NT;
  constexpr int NO = (NT + 63) / 64;
  __shared__ cplx sS[P * P];
  __shared__ cplx sM[NT];
  const int lane = threadIdx.x;
  int r = blockIdx.x < 13 ? blockIdx.x : blockIdx.x + 1;  // skip the centre (13)
  const int ix = r / 9 - 1, iy = (r / 3) % 3 - 1, iz = r % 3 - 1;
  for (int t = lane; t < NT; t += 64) sM[t] = Mp[t];
  const double mi0 = M[0].re, mj0 = Mp[0].re;
  double ar[NO], ai[NO];
  int tj[NO], tk[NO];
#pragma unroll
  for (int o = 0; o < NO; o++) {
    ar[o] = ai[o] = 0;
    coef_nm(lane + 64 * o < NT ? lane + 64 * o : 0, tj[o], tk[o]);
  }
  const double4 X = cX[0];
  for (int cx = -1; cx <= 1; cx++)
    for (int cy = -1; cy <= 1; cy++)
      for (int cz = -1; cz <= 1; cz++) {
        const double px = (ix * 3 + cx) * cycle, py = (iy * 3 + cy) * cycle, pz = (iz * 3 + cz) * cycle;
        // dX = X_root - X_parent - Xp, with the parent a copy of the root.
        const Geom g = make_geom(X.x - X.x - px, X.y - X.y - py, X.z - X.z - pz);
        __syncthreads();
        for (int t = lane; t < NT; t += 64) {
          int a, b;
          coef_nm(t, a, b);
          double re, im;
          singular_nm(a, b, g, re, im);
          sS[a * a + a + b] = {re, im};
          if (b > 0) sS[a * a + a - b] = {re, -im};
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < NO; o++) {
          if (lane + 64 * o >= NT) continue;
          const int j = tj[o], k = tk[o];
          // LaplaceSphericalCPU.cxx:149-191 (MASS: :157-172) for one (j, k).
          const double Cnm = (mass ? mi0 * mj0 : 1.0) * ((j & 1) ? -1.0 : 1.0);
          double lr = 0, li = 0;
          int n0 = 0;
          if (mass) {
            const cplx y = sS[j * j + j - k];
            lr += Cnm * y.re;
            li += Cnm * y.im;
            n0 = 1;
          }
          for (int n = n0; n < P - j; n++) {
            const int a = j + n;
            for (int m = -n; m < 0; m++) {  // conj(M_n^{-m}) S_{j+n}^{m-k}
              const cplx mv = sM[tri(n) - m];
              const cplx y = sS[a * a + a + m - k];
              lr += Cnm * (mv.re * y.re + mv.im * y.im);
              li += Cnm * (mv.re * y.im - mv.im * y.re);
            }
            for (int m = 0; m <= n; m++) {
              const double c2 = Cnm * ((((k - m) * (k < m) + m) & 1) ? -1.0 : 1.0);
              const cplx mv = sM[tri(n) + m];
              const cplx y = sS[a * a + a + m - k];
              lr += c2 * (mv.re * y.re - mv.im * y.im);
              li += c2 * (mv.re * y.im + mv.im * y.re);
            }
          }
          ar[o] += lr;
          ai[o] += li;
        }
      }
#pragma unroll
  for (int o = 0; o < NO; o++)
    if (lane + 64 * o < NT) part[(size_t)blockIdx.x * NT + lane + 64 * o] = {ar[o], ai[o]};
}

// L_root += sum of the 26 partials (ring order); then, unless this was the
// last sublevel, Mp <- M2M of the 26 neighbour copies of the periodic parent.
template <int P>
__global__ void __launch_bounds__(64)
    k_periodic_update(const double4 *__restrict__ cX, cplx *__restrict__ L, const cplx *__restrict__ part,
                      cplx *__restrict__ Mp, double cycle, int mass, int last) {
  constexpr int NT = Expn<P>::NT;
  __shared__ cplx sM[NT];
  __shared__ cplx sY[NT];
  __shared__ double m0s;
  const int lane = threadIdx.x;
  for (int t = lane; t < NT; t += 64) {
    cplx l = L[t];
    for (int b = 0; b < 26; b++) {
      l.re += part[(size_t)b * NT + t].re;
      l.im += part[(size_t)b * NT + t].im;
    }
    L[t] = l;
  }
  if (last) return;
  if (lane == 0) m0s = Mp[0].re;
  __syncthreads();
  for (int t = lane; t < NT; t += 64) {
    cplx v = Mp[t];
    if (mass && t > 0) {  // traversal.h:155-156
      v.re *= m0s;
      v.im *= m0s;
    }
    sM[t] = v;
  }
  const double4 X = cX[0];
  double acc_re[2] = {0, 0}, acc_im[2] = {0, 0};
  for (int ix = -1; ix <= 1; ix++)
    for (int iy = -1; iy <= 1; iy++)
      for (int iz = -1; iz <= 1; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        const double xj = X.x + ix * cycle, yj = X.y + iy * cycle, zj = X.z + iz * cycle;
        const Geom g = make_geom(X.x - xj, X.y - yj, X.z - zj);
        __syncthreads();
        for (int t = lane; t < NT; t += 64) {
          int n, m;
          coef_nm(t, n, m);
          double re, im;
          regular_nm(n, m, g, re, im);
          sY[t] = {re, im};
        }
        __syncthreads();
        for (int o = 0; o < 2; o++) {
          const int t = lane + 64 * o;
          if (t >= NT) continue;
          int j, k;
          coef_nm(t, j, k);
          m2m_lane(j, k, sM, sY, acc_re[o], acc_im[o]);
        }
      }
  __syncthreads();
  for (int o = 0; o < 2; o++) {
    const int t = lane + 64 * o;
    if (t < NT) sY[t] = {acc_re[o], acc_im[o]};
  }
  __syncthreads();
  const double m0 = sY[0].re;
  for (int t = lane; t < NT; t += 64) {
    cplx v = sY[t];
    if (mass && t > 0) {  // traversal.h:176-178
      v.re /= m0;
      v.im /= m0;
    }
    Mp[t] = v;
  }
}

__global__ void k_periodic_mass_L(const cplx *__restrict__ M, cplx *__restrict__ L, int nt) {
  const double m0 = M[0].re;  // traversal.h:181-183
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    L[t].re /= m0;
    L[t].im /= m0;
  }
}

template <int P>
static void periodic_impl(Handle &h) {
  constexpr int NT = Expn<P>::NT;
  h.per_M.resize(NT);
  h.per_part.resize(26 * NT);
  copy_device(h.per_M.p, h.M.p, NT * sizeof(cplx), h.stream);
  double cycle = h.cfg.cycle;
  for (int lev = 0; lev < h.cfg.images - 1; lev++) {
    hipLaunchKernelGGL((k_periodic_m2l<P>), dim3(26), dim3(64), 0, h.stream, h.cX.p, h.M.p, h.per_M.p, cycle,
                       h.cfg.mass, h.per_part.p);
    hipLaunchKernelGGL((k_periodic_update<P>), dim3(1), dim3(64), 0, h.stream, h.cX.p, h.L.p, h.per_part.p,
                       h.per_M.p, cycle, h.cfg.mass, lev == h.cfg.images - 2 ? 1 : 0);
    cycle *= 3;
  }
  if (h.cfg.mass) hipLaunchKernelGGL(k_periodic_mass_L, dim3(1), dim3(64), 0, h.stream, h.M.p, h.L.p, NT);
  EXA_HIP(hipGetLastError());
}

void run_periodic(Handle &h) {
  if (h.cfg.images <= 0 || h.ncells == 0) return;
  if (h.cfg.basis == 1) {
    run_periodic_cart(h);
    if (h.cfg.mass) hipLaunchKernelGGL(k_periodic_mass_L, dim3(1), dim3(64), 0, h.stream, h.M.p, h.L.p, h.nterm);
    EXA_HIP(hipGetLastError());
    return;
  }
  switch (h.cfg.P) {
    case 4: periodic_impl<4>(h); break;
    case 6: periodic_impl<6>(h); break;
    case 8: periodic_impl<8>(h); break;
    case 10: periodic_impl<10>(h); break;
    case 12: periodic_impl<12>(h); break;
    default: throw Error(EXAFMM_HIP_EUNSUP, "P must be one of 4, 6, 8, 10, 12");
  }
}

// ---------------------------------------------------------------------------
// Dipole of the bodies about X0 (tree order, up_down_pass.h:151-157):
// per-block partial sums, then one block adds them in block order.
constexpr int kDipBlocks = 256;

__global__ void __launch_bounds__(256)
    k_dipole_partial(const double4 *__restrict__ body, int n, double x0, double y0, double z0,
                     double *__restrict__ part) {
  __shared__ double s[3][256];
  double d[3] = {0, 0, 0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double4 b = body[i];
    d[0] += (b.x - x0) * b.w;
    d[1] += (b.y - y0) * b.w;
    d[2] += (b.z - z0) * b.w;
  }
  for (int k = 0; k < 3; k++) s[k][threadIdx.x] = d[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_dipole_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
  if (threadIdx.x < 3) {
    double v = 0;
    for (int b = 0; b < nb; b++) v += part[b * 3 + threadIdx.x];
    out[threadIdx.x] = v;
  }
}

void run_get_dipole(Handle &h, const double *X0, double *out) {
  h.red.reserve(3 * kDipBlocks + 3);
  hipLaunchKernelGGL(k_dipole_partial, dim3(kDipBlocks), dim3(256), 0, h.stream, h.body4.p, h.n, X0[0], X0[1],
                     X0[2], h.red.p);
  hipLaunchKernelGGL(k_dipole_final, dim3(1), dim3(64), 0, h.stream, h.red.p, kDipBlocks, h.red.p + 3 * kDipBlocks);
  EXA_HIP(hipGetLastError());
  EXA_HIP(hipMemcpyAsync(out, h.red.p + 3 * kDipBlocks, 3 * sizeof(double), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
}

// dipoleCorrection (up_down_pass.h:160-168) on the tree-order results.
__global__ void k_dipole_correction(const double4 *__restrict__ body, double4 *__restrict__ trg, int n, double coef,
                                    double dx, double dy, double dz, int numBodies) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double nd = (dx * dx + dy * dy) + dz * dz;
  double4 t = trg[i];
  t.x -= coef * nd / numBodies / body[i].w;
  t.y -= coef * dx;
  t.z -= coef * dy;
  t.w -= coef * dz;
  trg[i] = t;
}

void run_dipole_correction(Handle &h, const double *dipole, int numBodies, double cycle) {
  const double coef = 4 * M_PI / (3 * cycle * cycle * cycle);
  if (h.n == 0) return;
  hipLaunchKernelGGL(k_dipole_correction, dim3(div_up(h.n, 256)), dim3(256), 0, h.stream, h.body4.p, h.trg4.p, h.n,
                     coef, dipole[0], dipole[1], dipole[2], numBodies);
  EXA_HIP(hipGetLastError());
}

}  // namespace exafmm
