// Scan and sort helpers over hipCUB (rocPRIM) device primitives, plus the
// handle destructor.  Sorting and scanning are bookkeeping around the hot
// path (tree build and list grouping); the hot kernels are hand-written.
#include <algorithm>

#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "internal.h"

namespace exafmm {

Handle::~Handle() {
  drop_graph(*this);
  for (hipEvent_t e : {ev_fork, ev_gmult, ev_gp2p})
    if (e) (void)hipEventDestroy(e);
  qin.release();
  in4.release();
  body4.release();
  bidx.release();
  trg4.release();
  trgf.release();
  cX.release();
  cRb.release();
  cParent.release();
  cChild.release();
  cNChild.release();
  cBody.release();
  cNBody.release();
  cLevel.release();
  cKey.release();
  M.release();
  L.release();
  lvl_cells.release();
  p2p_off.release();
  p2p_src.release();
  m2l_off.release();
  m2l_src.release();
  p2p_img.release();
  m2l_img.release();
  p2p_tgt.release();
  m2l_tgt.release();
  m2l_sched.release();
  for (auto *b : {&at_X, &at_aX, &at_body2}) b->release();
  for (auto *b : {&at_start, &at_count, &at_parent, &at_leaf, &at_level, &at_first, &at_nchild, &at_astart,
                  &at_acount, &at_agid, &at_idx2, &at_nnode, &at_cellid, &at_cursor, &at_nl})
    b->release();
  raw_mi.release();
  raw_mj.release();
  raw_pi.release();
  raw_pj.release();
  raw_mimg.release();
  raw_pimg.release();
  tmp.release();
  k64a.release();
  k64b.release();
  ia.release();
  ib.release();
  ic.release();
  id.release();
  ie.release();
  ifl.release();
  fa.release();
  fb.release();
  cnt.release();
  red.release();
  box.release();
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (ev_mult) (void)hipEventDestroy(ev_mult);
  if (ev_cnt) (void)hipEventDestroy(ev_cnt);
  if (ev_lists) (void)hipEventDestroy(ev_lists);
  if (ev_p2pcnt) (void)hipEventDestroy(ev_p2pcnt);
  if (ev_bounds) (void)hipEventDestroy(ev_bounds);
  if (hpin) (void)hipHostFree(hpin);
  if (stream) (void)hipStreamDestroy(stream);
  if (stream2) (void)hipStreamDestroy(stream2);
  if (stream3) (void)hipStreamDestroy(stream3);
}

// Zero / copy device memory by kernel (graph-capturable; bytes a multiple of 8).
__global__ void k_zero_u64(unsigned long long *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_copy_u64(unsigned long long *__restrict__ d, const unsigned long long *__restrict__ s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

void zero_device(void *p, size_t bytes, hipStream_t st) {
  const int64_t n = (int64_t)(bytes / 8);
  if (n == 0) return;
  hipLaunchKernelGGL(k_zero_u64, dim3((int)std::min<int64_t>(div_up(n, 256), 8192)), dim3(256), 0, st,
                     (unsigned long long *)p, n);
  EXA_HIP(hipGetLastError());
}

void copy_device(void *d, const void *s, size_t bytes, hipStream_t st) {
  const int64_t n = (int64_t)(bytes / 8);
  if (n == 0) return;
  hipLaunchKernelGGL(k_copy_u64, dim3((int)std::min<int64_t>(div_up(n, 256), 8192)), dim3(256), 0, st,
                     (unsigned long long *)d, (const unsigned long long *)s, n);
  EXA_HIP(hipGetLastError());
}

void exclusive_scan_int(Handle &h, const int *in, int *out, int n) { exclusive_scan_int(in, out, n, h.stream, h.tmp); }

void exclusive_scan_int(const int *in, int *out, int n, hipStream_t st, DBuf<unsigned char> &tmp) {
  size_t bytes = 0;
  EXA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
  tmp.reserve(bytes);
  EXA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, st));
}

struct Int4Sum {
  __host__ __device__ int4 operator()(const int4 &a, const int4 &b) const {
    return make_int4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
  }
};

void exclusive_scan_int4(Handle &h, const int4 *in, int4 *out, int n) {
  size_t bytes = 0;
  const int4 zero = make_int4(0, 0, 0, 0);
  EXA_HIP(hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, in, out, Int4Sum{}, zero, n, h.stream));
  h.tmp.reserve(bytes);
  EXA_HIP(hipcub::DeviceScan::ExclusiveScan(h.tmp.p, bytes, in, out, Int4Sum{}, zero, n, h.stream));
}

// Stable radix sorts of (key, int) pairs over key bits [0, bits).  rocprim's
// default configuration sorts up to 2^20 items by a block sort and merge
// passes (about 20 launches for the 2^20 body keys of a 1M-body tree); from
// kOnesweepMin items on, the configuration here takes the onesweep radix sort
// (one pass per 8 key bits).  Smaller sorts keep the default: they may run
// beside P2P, where onesweep's chained blocks wait for CUs (a 37K-item sort
// took 2 ms there instead of 0.17).
using OnesweepConfig =
    rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, rocprim::default_config, 0>;
constexpr int kOnesweepMin = 1 << 18;

template <typename Config, typename K>
static void radix_sort_pairs_cfg(K *keys, K *keys_alt, int *vals, int *vals_alt, int n, int bits,
                                 bool &result_in_alt, hipStream_t st, DBuf<unsigned char> &tmp) {
  rocprim::double_buffer<K> k(keys, keys_alt);
  rocprim::double_buffer<int> v(vals, vals_alt);
  const unsigned int end_bit = bits > 0 ? bits : 1;
  size_t bytes = 0;
  EXA_HIP(rocprim::radix_sort_pairs<Config>(nullptr, bytes, k, v, n, 0u, end_bit, st));
  tmp.reserve(bytes);
  EXA_HIP(rocprim::radix_sort_pairs<Config>(tmp.p, bytes, k, v, n, 0u, end_bit, st));
  result_in_alt = k.current() != keys;
}

template <typename K>
static void radix_sort_pairs(K *keys, K *keys_alt, int *vals, int *vals_alt, int n, int bits, bool &result_in_alt,
                             hipStream_t st, DBuf<unsigned char> &tmp) {
  if (n >= kOnesweepMin)
    radix_sort_pairs_cfg<OnesweepConfig>(keys, keys_alt, vals, vals_alt, n, bits, result_in_alt, st, tmp);
  else
    radix_sort_pairs_cfg<rocprim::default_config>(keys, keys_alt, vals, vals_alt, n, bits, result_in_alt, st, tmp);
}

void sort_pairs_u64_int(Handle &h, uint64_t *keys, uint64_t *keys_alt, int *vals, int *vals_alt, int n, int bits,
                        bool &result_in_alt) {
  radix_sort_pairs(keys, keys_alt, vals, vals_alt, n, bits, result_in_alt, h.stream, h.tmp);
}

void sort_pairs_int_int(Handle &h, int *keys, int *keys_alt, int *vals, int *vals_alt, int n, int bits,
                        bool &result_in_alt) {
  radix_sort_pairs(keys, keys_alt, vals, vals_alt, n, bits, result_in_alt, h.stream, h.tmp);
}

void sort_pairs_int_int(int *keys, int *keys_alt, int *vals, int *vals_alt, int n, int bits, bool &result_in_alt,
                        hipStream_t st, DBuf<unsigned char> &tmp) {
  radix_sort_pairs(keys, keys_alt, vals, vals_alt, n, bits, result_in_alt, st, tmp);
}

}  // namespace exafmm
