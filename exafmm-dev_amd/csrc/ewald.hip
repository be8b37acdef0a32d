// Ewald summation of the periodic Coulomb potential and gradient (the check
// the periodic FMM is compared with).
//
// Reference: include/ewald.h:25-181 (driven by examples/ewald.cxx:91-112).
//   wave part:  W_k = sum_j q_j e^{i k.x_j s} (s = 2 pi / cycle), scaled by
//               coef exp(-|k|^2 coef2) / |k|^2, then per body
//               phi += Re(W e^{-i k.x s}),  grad -= s k Im(...)
//               over the half-space waves |k| <= ksize (initWaves order);
//   real part:  per target leaf, the source tree is walked; a source cell is
//               kept while the wrapped distance of the cell centres is below
//               3 cutoff, and each kept source leaf's bodies interact with
//               the image fixed by that wrap (erfc-screened, R^2 < cutoff^2);
//   self term:  phi_i -= 2 alpha q_i / sqrt(pi).
// The results are already normalised (no q_i factor), like the reference's.
//
// Kernels: one workgroup per wave for the forward DFT (body loop, fixed-order
// reduction), one thread per body for the inverse DFT (waves staged in LDS),
// one 64-lane workgroup per target leaf for the real part (source leaves
// selected by lanes, bodies staged in LDS).  Source leaves that cannot hold a
// pair inside the cut-off (centre distance minus both cells' extents) are
// skipped: they contribute nothing in the reference either.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace exafmm {

constexpr int kDftThreads = 256;

__global__ void __launch_bounds__(kDftThreads)
    k_ewald_dft(const double4 *__restrict__ body, int n, double *__restrict__ waves, double scale, double coef,
                double coef2) {
  __shared__ double sr[kDftThreads], si[kDftThreads];
  double *W = waves + 5 * (size_t)blockIdx.x;
  const double kx = W[0], ky = W[1], kz = W[2];
  double re = 0, im = 0;
  for (int b = threadIdx.x; b < n; b += kDftThreads) {
    const double4 x = body[b];
    const double th = kx * x.x * scale + ky * x.y * scale + kz * x.z * scale;
    double s, c;
    sincos(th, &s, &c);
    re += x.w * c;
    im += x.w * s;
  }
  sr[threadIdx.x] = re;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int w = kDftThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double K2 = kx * kx + ky * ky + kz * kz;
    const double factor = coef * exp(-K2 * coef2) / K2;
    W[3] = sr[0] * factor;
    W[4] = si[0] * factor;
  }
}

constexpr int kIdftStage = 512;

__global__ void __launch_bounds__(256)
    k_ewald_idft(const double4 *__restrict__ body, int n, const double *__restrict__ waves, int nw, double scale,
                 double4 *__restrict__ out) {
  __shared__ double sw[5 * kIdftStage];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const double4 x = b < n ? body[b] : make_double4(0, 0, 0, 0);
  double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
  for (int w0 = 0; w0 < nw; w0 += kIdftStage) {
    const int cnt = min(kIdftStage, nw - w0);
    __syncthreads();
    for (int i = threadIdx.x; i < 5 * cnt; i += blockDim.x) sw[i] = waves[5 * (size_t)w0 + i];
    __syncthreads();
    for (int w = 0; w < cnt; w++) {
      const double kx = sw[5 * w], ky = sw[5 * w + 1], kz = sw[5 * w + 2], wr = sw[5 * w + 3], wi = sw[5 * w + 4];
      const double th = kx * x.x * scale + ky * x.y * scale + kz * x.z * scale;
      double s, c;
      sincos(th, &s, &c);
      const double dtmp = wr * s - wi * c;
      t0 += wr * c + wi * s;
      t1 -= dtmp * kx;
      t2 -= dtmp * ky;
      t3 -= dtmp * kz;
    }
  }
  if (b < n) {
    double4 o = out[b];
    o.x += t0;
    o.y += t1 * scale;
    o.z += t2 * scale;
    o.w += t3 * scale;
    out[b] = o;
  }
}

__device__ inline double wrap1(double v, double w) {  // vec.h:193-205
  if (v < -w / 2) v += w;
  if (v > w / 2) v -= w;
  return v;
}

constexpr int kEwStage = 128;

__global__ void __launch_bounds__(64)
    k_ewald_real(int ncells, const double4 *__restrict__ cX, const double *__restrict__ cRb,
                 const int *__restrict__ cParent, const int *__restrict__ cNChild, const int *__restrict__ cBody,
                 const int *__restrict__ cNBody, const double4 *__restrict__ body, double4 *__restrict__ out,
                 double alpha, double cutoff, double cycle, int walk) {
  __shared__ double4 sb[kEwStage];
  __shared__ double4 sacc[64];
  const int c = blockIdx.x;
  if (cNChild[c] != 0 || cNBody[c] == 0) return;
  const int lane = threadIdx.x;
  const int tb0 = cBody[c], tn = cNBody[c];
  const double4 Xi = cX[c];
  const double Rbi = cRb[c];
  const int T = tn <= 16 ? 16 : (tn <= 32 ? 32 : 64);
  const int G = 64 / T;
  const int g = lane / T, ti = lane % T;
  const double cut2 = cutoff * cutoff, a3 = alpha * alpha * alpha, a2 = alpha * alpha;
  const double k2sqrtpi = 1.1283791670955126;  // 2 / sqrt(pi)
  for (int tbase = 0; tbase < tn; tbase += T) {
    const bool active = tbase + ti < tn;
    const double4 xi = active ? body[tb0 + tbase + ti] : make_double4(0, 0, 0, 0);
    double pot = 0, gx = 0, gy = 0, gz = 0;
    for (int base = 0; base < ncells; base += 64) {
      // Lanes select the source leaves of this chunk and their images.
      const int cj = base + lane;
      bool keep = false;
      double px = 0, py = 0, pz = 0;
      if (cj < ncells && cNChild[cj] == 0 && cNBody[cj] > 0) {
        const double4 Xj = cX[cj];
        const double dx = wrap1(Xi.x - Xj.x, cycle), dy = wrap1(Xi.y - Xj.y, cycle), dz = wrap1(Xi.z - Xj.z, cycle);
        px = Xi.x - Xj.x - dx;
        py = Xi.y - Xj.y - dy;
        pz = Xi.z - Xj.z - dz;
        const double R = sqrt(dx * dx + dy * dy + dz * dz);
        keep = R < 3 * cutoff;
        // Body extents: a body lies within sqrt(3) Rb of its cell centre
        // (1% margin for the rounding of the cell geometry).
        keep = keep && (R - 1.01 * 1.7320508075688772 * (Rbi + cRb[cj]) < cutoff);
        // Every ancestor must pass the 3 cutoff test too (depth-bounded walk).
        for (int a = cj, d = 0; keep && walk && a != 0 && d < 64; d++) {
          a = cParent[a];
          const double4 Xa = cX[a];
          const double ex = wrap1(Xi.x - Xa.x, cycle), ey = wrap1(Xi.y - Xa.y, cycle), ez = wrap1(Xi.z - Xa.z, cycle);
          keep = sqrt(ex * ex + ey * ey + ez * ez) < 3 * cutoff;
        }
      }
      uint64_t mask = __ballot(keep);
      int sbase = 0;
      while (mask) {
        // Fill: the kept leaves in cell order (or parts of one), each
        // shifted by its image, up to kEwStage bodies.
        int cnt = 0;
        __syncthreads();
        while (mask && cnt < kEwStage) {
          const int l = __builtin_ctzll(mask);
          const int s = base + l;
          const double qx = __shfl(px, l), qy = __shfl(py, l), qz = __shfl(pz, l);
          const int sb0 = cBody[s], sn = cNBody[s];
          const int take = min(sn - sbase, kEwStage - cnt);
          for (int k = lane; k < take; k += 64) {
            double4 v = body[sb0 + sbase + k];
            v.x += qx;
            v.y += qy;
            v.z += qz;
            sb[cnt + k] = v;
          }
          cnt += take;
          sbase += take;
          if (sbase == sn) {
            mask &= mask - 1;
            sbase = 0;
          }
        }
        __syncthreads();
        for (int k = g; k < cnt; k += G) {
          const double4 xj = sb[k];
          const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
          const double R2 = dx * dx + dy * dy + dz * dz;
          if (active && 0 < R2 && R2 < cut2) {
            const double R2s = R2 * a2, Rs = sqrt(R2s), invRs = 1 / Rs;
            const double invR2s = invRs * invRs, invR3s = invR2s * invRs, e = erfc(Rs);
            const double dtmp = xj.w * (k2sqrtpi * exp(-R2s) * invR2s + e * invR3s) * a3;
            pot += xj.w * e * invRs * alpha;
            gx -= dx * dtmp;
            gy -= dy * dtmp;
            gz -= dz * dtmp;
          }
        }
      }
    }
    __syncthreads();
    sacc[lane] = make_double4(pot, gx, gy, gz);
    __syncthreads();
    if (g == 0 && active) {
      double4 s = sacc[ti];
      for (int gg = 1; gg < G; gg++) {
        const double4 o = sacc[gg * T + ti];
        s.x += o.x;
        s.y += o.y;
        s.z += o.z;
        s.w += o.w;
      }
      double4 o = out[tb0 + tbase + ti];
      o.x += s.x;
      o.y += s.y;
      o.z += s.z;
      o.w += s.w;
      out[tb0 + tbase + ti] = o;
    }
  }
}

__global__ void k_ewald_self(const double4 *__restrict__ body, int n, double alpha, double4 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i].x -= 1.1283791670955126 * body[i].w * alpha;  // M_2_SQRTPI q alpha
}

__global__ void k_zero4(double4 *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_double4(0, 0, 0, 0);
}

__global__ void k_scatter_add(const double4 *__restrict__ src, const int *__restrict__ bidx, int n,
                              double4 *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 d = dst[bidx[i]];
  const double4 s = src[i];
  d.x += s.x;
  d.y += s.y;
  d.z += s.z;
  d.w += s.w;
  dst[bidx[i]] = d;
}

// initWaves, ewald.h:64-90: the half space l > 0, or l = 0 and m > 0, or
// l = m = 0 and n > 0, with l^2 + m^2 + n^2 <= ksize^2, in loop order.
static std::vector<double> init_waves(int ksize) {
  std::vector<double> w;
  const int kmax = ksize, kmaxsq = ksize * ksize;
  for (int l = 0; l <= kmax; l++)
    for (int m = l == 0 ? 0 : -kmax; m <= kmax; m++)
      for (int n = (l == 0 && m == 0) ? 1 : -kmax; n <= kmax; n++)
        if (l * l + m * m + n * n <= kmaxsq) {
          w.push_back(l);
          w.push_back(m);
          w.push_back(n);
          w.push_back(0);
          w.push_back(0);
        }
  return w;
}

void run_ewald(Handle &h, int ksize, double alpha, double sigma, double cutoff, double cycle, int parts,
               double *trg) {
  const int n = h.n;
  h.ew_trg.resize(n > 0 ? n : 1);
  hipLaunchKernelGGL(k_zero4, dim3(div_up(n, 256)), dim3(256), 0, h.stream, h.ew_trg.p, n);
  const double scale = 2 * M_PI / cycle;
  if (parts & 1) {  // wavePart, ewald.h:166-181
    const std::vector<double> w = init_waves(ksize);
    const int nw = (int)w.size() / 5;
    if (nw > 0) {
      h.ew_waves.resize(w.size());
      EXA_HIP(hipMemcpyAsync(h.ew_waves.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, h.stream));
      const double coef = .5 / M_PI / M_PI / sigma / cycle;
      const double coef2 = scale * scale / (4 * alpha * alpha);
      hipLaunchKernelGGL(k_ewald_dft, dim3(nw), dim3(kDftThreads), 0, h.stream, h.body4.p, n, h.ew_waves.p, scale,
                         coef, coef2);
      hipLaunchKernelGGL(k_ewald_idft, dim3(div_up(n, 256)), dim3(256), 0, h.stream, h.body4.p, n, h.ew_waves.p, nw,
                         scale, h.ew_trg.p);
      EXA_HIP(hipStreamSynchronize(h.stream));  // the host wave table is freed on return
    }
  }
  if (parts & 2) {  // realPart, ewald.h:145-156
    // The ancestor walk can be skipped when 3 cutoff exceeds every wrapped
    // centre distance: a wrapped coordinate difference is at most
    // max(cycle / 2, extent - cycle) for bodies spanning `extent`.
    sync_bounds(h);
    double ext = 0;
    for (int d = 0; d < 3; d++) ext = std::max(ext, h.xmax[d] - h.xmin[d]);
    const double bound = std::sqrt(3.0) * std::max(0.5 * cycle, ext - cycle);
    const int walk = 3 * cutoff > bound * 1.000001 ? 0 : 1;
    hipLaunchKernelGGL(k_ewald_real, dim3(h.ncells), dim3(64), 0, h.stream, h.ncells, h.cX.p, h.cRb.p, h.cParent.p,
                       h.cNChild.p, h.cBody.p, h.cNBody.p, h.body4.p, h.ew_trg.p, alpha, cutoff, cycle, walk);
  }
  if (parts & 4)  // selfTerm, ewald.h:159-163
    hipLaunchKernelGGL(k_ewald_self, dim3(div_up(n, 256)), dim3(256), 0, h.stream, h.body4.p, n, alpha, h.ew_trg.p);
  EXA_HIP(hipGetLastError());
  // Accumulate into the caller's array (caller order).
  DBuf<double4> o;
  o.resize(n);
  EXA_HIP(hipMemcpyAsync(o.p, trg, n * sizeof(double4), hipMemcpyHostToDevice, h.stream));
  hipLaunchKernelGGL(k_scatter_add, dim3(div_up(n, 256)), dim3(256), 0, h.stream, h.ew_trg.p, h.bidx.p, n, o.p);
  EXA_HIP(hipGetLastError());
  EXA_HIP(hipMemcpyAsync(trg, o.p, n * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  o.release();
}

}  // namespace exafmm
