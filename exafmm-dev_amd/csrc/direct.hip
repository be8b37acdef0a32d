// FP64 direct summation onto sampled targets (reference:
// include/traversal.h:266-337, Traversal::direct + normalize), the accuracy
// yardstick for full-size runs.  One 256-lane workgroup per target; lanes
// stride over the sources, images are looped outside the sources as in the
// reference, and the lane partials are reduced in a fixed tree order.
#include <cmath>
#include <mutex>
#include <string>

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

static int image_range(const Handle &h) {
  int prange = 0;
  for (int i = 0; i < h.cfg.images; i++) prange += int(std::pow(3., i));
  return prange;
}

// Direct summation of nt staged targets t against ns staged sources s into
// o (device arrays), results copied to trg on the host.
static void direct_staged(Handle &h, int nt, const double *t_xyzq, double4 *t, int ns, const double *s_xyzq,
                          double4 *s, double4 *o, double *trg) {
  EXA_HIP(hipMemcpyAsync(t, t_xyzq, nt * sizeof(double4), hipMemcpyHostToDevice, h.stream));
  if (ns > 0) EXA_HIP(hipMemcpyAsync(s, s_xyzq, ns * sizeof(double4), hipMemcpyHostToDevice, h.stream));
  hipLaunchKernelGGL(k_direct, dim3(nt), dim3(256), 0, h.stream, nullptr, t, s, ns, image_range(h), h.cfg.cycle, o);
  EXA_HIP(hipGetLastError());
  EXA_HIP(hipMemcpyAsync(trg, o, nt * sizeof(double4), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
}

void run_direct_arrays(Handle &h, int nt, const double *t_xyzq, int ns, const double *s_xyzq, double *trg) {
  if (nt <= 0) return;
  DBuf<double4> t, s, o;
  t.resize(nt);
  s.resize(ns > 0 ? ns : 1);
  o.resize(nt);
  direct_staged(h, nt, t_xyzq, t.p, ns, s_xyzq, s.p, o.p, trg);
  t.release();
  s.release();
  o.release();
}

// ---------------------------------------------------------------------------
// The allocation-free variant: the same staging in a static device array,
// `scratch`, instead of device buffers allocated per call.  Each device
// holds its own copy of the array (for the life of the process); its device
// address comes from hipGetSymbolAddress, taken once per device with that
// device current and kept in scratch_ptr.  Host code uses only those
// addresses, never the array's name.
constexpr size_t kScratchElems = EXAFMM_HIP_STATIC_SCRATCH_BODIES;  // double4 entries: 16 MiB
static __device__ __align__(256) double4 scratch[kScratchElems];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static double4 *scratch_ptr[kMaxDevices] = {};

static double4 *scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(EXAFMM_HIP_EINVAL, "direct_arrays_static: device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the handle's device is current
    if (e != hipSuccess || !p)
      throw Error(EXAFMM_HIP_EDEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    scratch_ptr[device] = static_cast<double4 *>(p);
  }
  return scratch_ptr[device];
}

void run_direct_arrays_static(Handle &h, int nt, const double *t_xyzq, int ns, const double *s_xyzq, double *trg) {
  const size_t need = 2 * (size_t)nt + (size_t)ns;  // targets, results, sources
  if (need > kScratchElems)
    throw Error(EXAFMM_HIP_EINVAL, "direct_arrays_static: 2 * nt + ns = " + std::to_string(need) +
                                       " exceeds the static scratch of " + std::to_string(kScratchElems) + " bodies");
  double4 *base = scratch_on(h.cfg.device);
  if (nt <= 0) return;
  direct_staged(h, nt, t_xyzq, base, ns, s_xyzq, base + 2 * (size_t)nt, base + nt, trg);
}

}  // namespace exafmm

extern "C" int exafmm_hip_direct_arrays_static(exafmm_hip_handle *H, int nt, const double *t_xyzq, int ns,
                                               const double *s_xyzq, double *trg) {
  return guarded(H, [&](exafmm::Handle &h) {
    if (!(nt >= 0 && ns >= 0 && (nt == 0 || (t_xyzq && trg)) && (ns == 0 || s_xyzq)))
      throw exafmm::Error(EXAFMM_HIP_EINVAL, "direct_arrays_static: bad arguments");
    exafmm::run_direct_arrays_static(h, nt, t_xyzq, ns, s_xyzq, trg);
  });
}
