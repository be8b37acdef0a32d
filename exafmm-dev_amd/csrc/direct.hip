// FP64 direct summation onto sampled targets (reference:
// include/traversal.h:266-337, Traversal::direct + normalize), the accuracy
// yardstick for full-size runs.  One 256-lane workgroup per target; lanes
// stride over the sources, images are looped outside the sources as in the
// reference, and the lane partials are reduced in a fixed tree order.
#include <cmath>

#include "internal.h"

namespace exafmm {

__global__ void __launch_bounds__(256)
    k_direct(const int *__restrict__ tgt, const double4 *__restrict__ tgt_xyzq, const double4 *__restrict__ src,
             int ns, int prange, double cycle, double4 *__restrict__ out) {
  __shared__ double4 red[256];
  const double4 xi = tgt ? src[tgt[blockIdx.x]] : tgt_xyzq[blockIdx.x];
  double pot = 0, ax = 0, ay = 0, az = 0;
  for (int ix = -prange; ix <= prange; ix++)
    for (int iy = -prange; iy <= prange; iy++)
      for (int iz = -prange; iz <= prange; iz++) {
        const double px = ix * cycle, py = iy * cycle, pz = iz * cycle;
        for (int j = threadIdx.x; j < ns; j += 256) {
          const double4 xj = src[j];
          const double dx = xi.x - xj.x - px, dy = xi.y - xj.y - py, dz = xi.z - xj.z - pz;
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 != 0) {
            const double invr = 1.0 / sqrt(r2);
            const double q3 = xj.w * invr * invr * invr;
            pot += xj.w * invr;
            ax += dx * q3;
            ay += dy * q3;
            az += dz * q3;
          }
        }
      }
  red[threadIdx.x] = make_double4(pot, ax, ay, az);
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      double4 a = red[threadIdx.x];
      const double4 b = red[threadIdx.x + w];
      a.x += b.x;
      a.y += b.y;
      a.z += b.z;
      a.w += b.w;
      red[threadIdx.x] = a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double4 s = red[0];
    out[blockIdx.x] = make_double4(s.x, -s.y, -s.z, -s.w);
  }
}

void run_direct(Handle &h, int nt, const int *target_ibody, double *trg) {
  if (nt <= 0) return;
  for (int i = 0; i < nt; i++)
    if (target_ibody[i] < 0 || target_ibody[i] >= h.n) throw Error(EXAFMM_HIP_EINVAL, "direct: target index out of range");
  int prange = 0;
  for (int i = 0; i < h.cfg.images; i++) prange += int(std::pow(3., i));
  DBuf<int> t;
  DBuf<double4> o;
  t.resize(nt);
  o.resize(nt);
  EXA_HIP(hipMemcpyAsync(t.p, target_ibody, nt * sizeof(int), hipMemcpyHostToDevice, h.stream));
  hipLaunchKernelGGL(k_direct, dim3(nt), dim3(256), 0, h.stream, t.p, nullptr, h.in4.p, h.n, prange, h.cfg.cycle,
                     o.p);
  EXA_HIP(hipGetLastError());
  EXA_HIP(hipMemcpyAsync(trg, o.p, nt * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  t.release();
  o.release();
}

void run_direct_arrays(Handle &h, int nt, const double *t_xyzq, int ns, const double *s_xyzq, double *trg) {
  if (nt <= 0) return;
  int prange = 0;
  for (int i = 0; i < h.cfg.images; i++) prange += int(std::pow(3., i));
  DBuf<double4> t, s, o;
  t.resize(nt);
  s.resize(ns > 0 ? ns : 1);
  o.resize(nt);
  EXA_HIP(hipMemcpyAsync(t.p, t_xyzq, nt * sizeof(double4), hipMemcpyHostToDevice, h.stream));
  if (ns > 0) EXA_HIP(hipMemcpyAsync(s.p, s_xyzq, ns * sizeof(double4), hipMemcpyHostToDevice, h.stream));
  hipLaunchKernelGGL(k_direct, dim3(nt), dim3(256), 0, h.stream, nullptr, t.p, s.p, ns, prange, h.cfg.cycle, o.p);
  EXA_HIP(hipGetLastError());
  EXA_HIP(hipMemcpyAsync(trg, o.p, nt * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  t.release();
  s.release();
  o.release();
}

}  // namespace exafmm
