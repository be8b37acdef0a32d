// P2P over the CSR interaction list: one 64-lane workgroup per target leaf.
//
// Reference: kernels/LaplaceP2PCPU.cxx:141-169 (pair) and :172-331 (self).
// Per pair: dX = x_i - x_j - X_periodic, R2 = |dX|^2, skipped when R2 == 0;
// pot += q_j / r, grad -= q_j dX / r^3.  The reference multiplies by q_i
// here and divides it out in L2P/normalize (LaplaceSphericalCPU.cxx:235,
// traversal.h:333-337); the GPU accumulates the normalised quantities
// directly.
//
// Lane layout: a leaf's targets are replicated over G = 64 / T lane groups
// (T = 16, 32 or 64 targets per group) and group g takes every G-th source
// body, so short leaves still fill the wavefront.  Source bodies are staged
// through LDS up to 128 at a time (consecutive list entries, each shifted by
// its periodic image); the G partial sums are combined in a fixed order at
// the end (deterministic, no atomics).
#include "internal.h"

namespace exafmm {

__device__ inline void image_shift(int img, double cycle, double &px, double &py, double &pz) {
  px = (img / 9 - 1) * cycle;
  py = ((img / 3) % 3 - 1) * cycle;
  pz = (img % 3 - 1) * cycle;
}

constexpr int kP2PStage = 256;  // source bodies staged in LDS per fill

// 1/sqrt(x) for finite x > 0: the hardware estimate refined by one
// higher-order Newton step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2, which
// takes the estimate's ~2^-22 relative error below FP64 rounding.  (No
// special-case fix-ups: the caller selects 0 for x == 0.)
__device__ inline double rsqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Walker over a target's source list (CSR entries [e, e1)).  The
// descriptors of up to 64 consecutive entries -- first body, body count,
// image -- and the exclusive prefix of their counts are loaded in one
// parallel step; a stage fill then maps each stage slot to (entry, body) by
// binary search over that prefix, so all body loads of a fill are
// independent of each other (no per-leaf dependent load chain).  The
// workgroup is one wave; every call is uniform across it.
struct SrcWalker {
  int e, sbase, dbase, dcount;
};

struct SrcDesc {
  int b0[64];
  int pre[65];
  unsigned char img[64];
};

template <int STAGE, typename Put>
__device__ inline int fill_stage(SrcWalker &w, int e1, int lane, const int *__restrict__ src,
                                 const unsigned char *__restrict__ img, const int *__restrict__ cBody,
                                 const int *__restrict__ cNBody, SrcDesc &sd, Put put) {
  __syncthreads();  // the previous stage (and descriptor batch) are consumed
  if (w.e >= w.dbase + w.dcount) {
    w.dbase = w.e;
    w.dcount = min(64, e1 - w.e);
    int n = 0;
    if (lane < w.dcount) {
      const int s = src[w.e + lane];
      sd.b0[lane] = cBody[s];
      n = cNBody[s];
      sd.img[lane] = img[w.e + lane];
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
  for (int k = lane; k < cnt; k += 64) {
    const int q = q0 + k;
    int lo = first, hi = w.dcount;  // largest i in [first, dcount) with pre[i] <= q
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sd.pre[mid] <= q)
        lo = mid;
      else
        hi = mid;
    }
    put(k, sd.b0[lo] + (q - sd.pre[lo]), sd.img[lo]);
  }
  // Advance to the first unconsumed body.
  const int qn = q0 + cnt;
  if (qn >= qend) {
    w.e = w.dbase + w.dcount;
    w.sbase = 0;
  } else {
    int lo = first, hi = w.dcount;
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

__global__ void __launch_bounds__(64)
    k_p2p_f64(const int *__restrict__ tgt_list, const int *__restrict__ off, const int *__restrict__ src,
              const unsigned char *__restrict__ img, const int *__restrict__ cBody, const int *__restrict__ cNBody,
              const double4 *__restrict__ body, double4 *__restrict__ trg, double cycle) {
  __shared__ double4 sb[kP2PStage];
  __shared__ double4 sacc[64];
  __shared__ SrcDesc sd;
  const int c = tgt_list[blockIdx.x];
  const int lane = threadIdx.x;
  const int tb0 = cBody[c], tn = cNBody[c];
  const int T = tn <= 16 ? 16 : (tn <= 32 ? 32 : 64);
  const int G = 64 / T;
  const int g = lane / T, ti = lane % T;
  const int e0 = off[c], e1 = off[c + 1];
  for (int tbase = 0; tbase < tn; tbase += T) {
    const bool active = tbase + ti < tn;
    const double4 xi = active ? body[tb0 + tbase + ti] : make_double4(0, 0, 0, 0);
    double pot = 0, ax = 0, ay = 0, az = 0;
    SrcWalker w{e0, 0, e0, 0};
    while (w.e < e1) {
      // Fill: consecutive source leaves (or parts of one) up to kP2PStage
      // bodies, each shifted by its periodic image.
      const int cnt = fill_stage<kP2PStage>(w, e1, lane, src, img, cBody, cNBody, sd, [&](int k, int b, int im) {
        double px, py, pz;
        image_shift(im, cycle, px, py, pz);
        double4 v = body[b];
        v.x += px;
        v.y += py;
        v.z += pz;
        sb[k] = v;
      });
#pragma unroll 4
      for (int k = g; k < cnt; k += G) {
        const double4 xj = sb[k];
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rs = rsqrt_pos(r2);  // unconditional: a select, not a branch, skips R2 == 0
        const double invr = r2 > 0 ? rs : 0.0;
        const double qinvr = xj.w * invr;
        const double qinvr3 = qinvr * invr * invr;
        pot += qinvr;
        ax += dx * qinvr3;
        ay += dy * qinvr3;
        az += dz * qinvr3;
      }
    }
    // Combine the G group partials in group order.
    __syncthreads();
    sacc[lane] = make_double4(pot, ax, ay, az);
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
      double4 tv = trg[tb0 + tbase + ti];
      tv.x += s.x;
      tv.y -= s.y;
      tv.z -= s.z;
      tv.w -= s.w;
      trg[tb0 + tbase + ti] = tv;
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
constexpr int kP2PStageF = 256;
constexpr int kSeg = 8;  // FP32 pair terms summed before each compensated add

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
    k_p2p_mixed(const int *__restrict__ tgt_list, const int *__restrict__ off, const int *__restrict__ src,
                const unsigned char *__restrict__ img, const int *__restrict__ cBody, const int *__restrict__ cNBody,
                const double4 *__restrict__ cX, const double4 *__restrict__ body, double4 *__restrict__ trg,
                double cycle) {
  __shared__ float4 sb[kP2PStageF];   // hi parts and charge
  __shared__ float4 sbl[kP2PStageF];  // lo parts
  __shared__ double4 sacc[64];
  __shared__ SrcDesc sd;
  const int c = tgt_list[blockIdx.x];
  const int lane = threadIdx.x;
  const int tb0 = cBody[c], tn = cNBody[c];
  const double4 Xc = cX[c];
  const int T = tn <= 16 ? 16 : (tn <= 32 ? 32 : 64);
  const int G = 64 / T;
  const int g = lane / T, ti = lane % T;
  const int e0 = off[c], e1 = off[c + 1];
  for (int tbase = 0; tbase < tn; tbase += T) {
    const bool active = tbase + ti < tn;
    const double4 xd = active ? body[tb0 + tbase + ti] : Xc;
    const double rx = xd.x - Xc.x, ry = xd.y - Xc.y, rz = xd.z - Xc.z;
    const float xi = (float)rx, yi = (float)ry, zi = (float)rz;
    const float xil = (float)(rx - xi), yil = (float)(ry - yi), zil = (float)(rz - zi);
    Kahan32 kp, kx, ky, kz;
    SrcWalker w{e0, 0, e0, 0};
    while (w.e < e1) {
      const int cnt = fill_stage<kP2PStageF>(w, e1, lane, src, img, cBody, cNBody, sd, [&](int k, int b, int im) {
        double px, py, pz;
        image_shift(im, cycle, px, py, pz);
        const double4 v = body[b];
        const double rx = (v.x + px) - Xc.x, ry = (v.y + py) - Xc.y, rz = (v.z + pz) - Xc.z;
        const float hx = (float)rx, hy = (float)ry, hz = (float)rz;
        sb[k] = make_float4(hx, hy, hz, (float)v.w);
        sbl[k] = make_float4((float)(rx - hx), (float)(ry - hy), (float)(rz - hz), 0.0f);
      });
      // Pad the tile to whole segments with zero charges (they add exactly 0).
      const int cntp = (cnt + kSeg * G - 1) / (kSeg * G) * (kSeg * G);
      for (int k = cnt + lane; k < cntp; k += 64) {
        sb[k] = make_float4(0, 0, 0, 0);
        sbl[k] = make_float4(0, 0, 0, 0);
      }
      __syncthreads();
      for (int k0 = g; k0 < cntp; k0 += kSeg * G) {
        float pot = 0, ax = 0, ay = 0, az = 0;
#pragma unroll
        for (int u = 0; u < kSeg; u++) {
          const int k = k0 + u * G;
          const float4 xj = sb[k];
          const float4 xl = sbl[k];
          const float dx = (xi - xj.x) + (xil - xl.x), dy = (yi - xj.y) + (yil - xl.y),
                      dz = (zi - xj.z) + (zil - xl.z);
          const float r2 = dx * dx + dy * dy + dz * dz;
          const float rs = __builtin_amdgcn_rsqf(r2);
          const float invr = r2 > 0.0f ? rs : 0.0f;
          const float qinvr = xj.w * invr;
          const float qinvr3 = qinvr * invr * invr;
          pot += qinvr;
          ax += dx * qinvr3;
          ay += dy * qinvr3;
          az += dz * qinvr3;
        }
        kp.add(pot);
        kx.add(ax);
        ky.add(ay);
        kz.add(az);
      }
    }
    __syncthreads();
    sacc[lane] = make_double4(kp.value(), kx.value(), ky.value(), kz.value());
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
      double4 tv = trg[tb0 + tbase + ti];
      tv.x += s.x;
      tv.y -= s.y;
      tv.z -= s.z;
      tv.w -= s.w;
      trg[tb0 + tbase + ti] = tv;
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
  EXA_HIP(hipMemsetAsync(h.trg4.p, 0, (size_t)h.n * sizeof(double4), st));
  if (h.np2p_tgt == 0) return;
  if (h.cfg.precision == 1)
    hipLaunchKernelGGL(k_p2p_mixed, dim3(h.np2p_tgt), dim3(64), 0, st, h.p2p_tgt.p, h.p2p_off.p, h.p2p_src.p,
                       h.p2p_img.p, h.cBody.p, h.cNBody.p, h.cX.p, h.body4.p, h.trg4.p, h.cfg.cycle);
  else
    hipLaunchKernelGGL(k_p2p_f64, dim3(h.np2p_tgt), dim3(64), 0, st, h.p2p_tgt.p, h.p2p_off.p, h.p2p_src.p,
                       h.p2p_img.p, h.cBody.p, h.cNBody.p, h.body4.p, h.trg4.p, h.cfg.cycle);
  EXA_HIP(hipGetLastError());
}

int64_t count_p2p_body_pairs(Handle &h) {
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
