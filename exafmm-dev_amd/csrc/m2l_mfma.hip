// M2L of lattice pairs as dense contractions on the matrix cores.
//
// Reference: kernels/LaplaceSphericalCPU.cxx:149-191 (non-MASS).  Octree cell
// centres of one level lie on that level's lattice, so the offset between a
// target of width w and a source of the same level or one level up or down
// is X_i - X_j = d u with d an integer vector and u = w / 4.  The singular
// harmonics are homogeneous, S_a(u d) = u^-(a+1) S_a(d), so with
// M~_n = M_n u^-n the translation of every such pair is one fixed real matrix
// per offset d:  L_j = u^-(j+1) T(d) M~.  In real form (rows (j, k, re|im),
// columns (n, m, re|im), m >= 0) the reference's sum over m = -n..n folds
// into 2 x 2 blocks, with M'_n^m = (-1)^m M_n^m, M'_n^-m = conj M_n^m and the
// signed harmonics S' of the spherical kernel:
//   T[(j,k,re),(n,m,re)] = (-1)^j ((-1)^m S1r + [m>0] S2r)
//   T[(j,k,re),(n,m,im)] = (-1)^j (-(-1)^m S1i + [m>0] S2i)
//   T[(j,k,im),(n,m,re)] = (-1)^j ((-1)^m S1i + [m>0] S2i)
//   T[(j,k,im),(n,m,im)] = (-1)^j ((-1)^m S1r - [m>0] S2r)
// S1 = S'_{j+n}^{m-k}(d), S2 = S'_{j+n}^{-m-k}(d), zero unless j + n < P.
//
// Work split: targets are grouped into tiles of 64 cells of one level and
// one octant, ordered by their distances to the domain's faces (cells of one
// octant share most of their interaction offsets, and those of one face
// class also the missing ones, so a tile's columns are mostly busy at every
// offset);
// a tile's local expansions (112 rows x 64 columns) stay in MFMA
// accumulators while the workgroup walks the offsets: per offset the matrix
// T(d) is staged in LDS in MFMA fragment order (only the 16 x 4 blocks that
// are not identically zero: j + n < P), each lane gathers its column's source
// multipole scaled by u^-n, and the 4 waves (16 columns each) run
// v_mfma_f64_16x16x4 (v_mfma_f32_16x16x4 in the mixed form).  Pairs that are
// not lattice pairs (levels two or more apart, periodic images, MASS mode,
// offsets beyond kLatD) go to the per-pair kernel.
#include <algorithm>
#include <climits>
#include <type_traits>
#include <vector>

#include "harmonics.h"
#include "internal.h"

namespace exafmm {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr double kLatUnit = 4;                   // u = w / kLatUnit
constexpr int kLatD = 60;                         // |d_c| <= kLatD (in units of u)
constexpr int kLatSide = 2 * kLatD + 1;
constexpr int kLatKeys = kLatSide * kLatSide * kLatSide;

template <int P>
struct MfmaShape {
  static constexpr int NT = P * (P + 1) / 2;
  static constexpr int NR = 2 * NT;              // real rows / columns
  static constexpr int RB = (NR + 15) / 16;      // row blocks
  static constexpr int KC = (NR + 3) / 4;        // column chunks
  static constexpr int jmin(int rb) {
    int r = rb * 16 / 2, j = 0;
    while ((j + 1) * (j + 2) / 2 <= r) j++;
    return j;
  }
  static constexpr int kc_end(int rb) {
    const int j = jmin(rb);
    const int cols = 2 * ((P - j) * (P - j + 1) / 2);  // columns with n <= P-1-j
    return (cols + 3) / 4;
  }
  static constexpr int base(int rb) {
    int s = 0;
    for (int i = 0; i < rb; i++) s += kc_end(i);
    return s;
  }
  static constexpr int NBLK = base(RB);
};

// n of packed coefficient t = n(n+1)/2 + m.
template <int P>
struct DegreeTable {
  int n[P * (P + 1) / 2 + 1];
  constexpr DegreeTable() : n() {
    int t = 0;
    for (int a = 0; a < P; a++)
      for (int b = 0; b <= a; b++) n[t++] = a;
    n[t] = P;
  }
};

// Component r of lane l of the 16x16x4 MFMA result holds D[row][l%16] with
// row = (l/16) + 4 r (FP64) or 4 (l/16) + r (FP32); A[l%16][l/16] and
// B[l/16][l%16] are lane l's operands for both.
template <typename R>
__device__ inline int d_row(int lane, int r) {
  return sizeof(R) == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
}

template <typename R>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
  typedef v4d type;
  static __device__ inline v4d mac(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};
template <>
struct MfmaAcc<float> {
  typedef v4f type;
  static __device__ inline v4f mac(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

// ---------------------------------------------------------------------------
// Pair classification: key of the lattice offset, or -1 (ptgt: the target
// cell of every CSR entry).
__global__ void k_lat_classify(const int *__restrict__ ptgt, const int *__restrict__ src,
                               const unsigned char *__restrict__ img, int64_t np, const double4 *__restrict__ cX,
                               const double *__restrict__ cRb, int *__restrict__ pkey, int *__restrict__ hist) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np) return;
  const int t = ptgt[e], s = src[e];
  int key = -1;
  // Same level or one level apart: then the offset's residue mod 4 fixes
  // the level relation (0: same, 2: source one up, odd: source one down).
  const double rt = cRb[t], rs = cRb[s];
  if (img[e] == 13 && (rs == rt || rs == 2 * rt || 2 * rs == rt)) {
    // Offsets in units of u = 2 R_t / kLatUnit; the lattice values are whole
    // numbers to far better than the 1e-6 test below.
    const double iu = kLatUnit / (2 * rt);
    const double4 a = cX[t], b = cX[s];
    const double dx = (a.x - b.x) * iu, dy = (a.y - b.y) * iu, dz = (a.z - b.z) * iu;
    const double rx = rint(dx), ry = rint(dy), rz = rint(dz);
    if (fabs(dx - rx) < 1e-6 && fabs(dy - ry) < 1e-6 && fabs(dz - rz) < 1e-6 && fabs(rx) <= kLatD &&
        fabs(ry) <= kLatD && fabs(rz) <= kLatD)
      key = (((int)rx + kLatD) * kLatSide + ((int)ry + kLatD)) * kLatSide + ((int)rz + kLatD);
  }
  pkey[e] = key;
  // Presence flag (no counting): a read first, so the many pairs of a
  // common offset do not all store to one address.
  if (key >= 0 && hist[key] == 0) hist[key] = 1;
}

// T(d) rows in memory and LDS: TE entries padded to whole 1 KiB chunks (one
// chunk = one wave-wide 16-byte-per-lane LDS DMA).
template <int P, typename R>
struct TPad {
  static constexpr int TE = MfmaShape<P>::NBLK * 64;
  static constexpr int CH = 1024 / (int)sizeof(R);
  static constexpr int NCH = (TE + CH - 1) / CH;
  static constexpr int TEP = NCH * CH;
};

// Source multipole rows for the contraction: lane (column n, kq) of a wave
// needs entries 4 kc + kq (kc = 0 .. KC-1) of its column's row.  A row keeps
// them in groups of VL = 16 bytes / sizeof(R) consecutive kc per kq, the four
// kq groups of one run of kc side by side: every 16-byte vector load of a
// lane is one group, and the 4 lanes of a column read 64 contiguous bytes.
// KC is padded to KCP (a multiple of VL) with zeros.
template <int P, typename R>
struct MsRow {
  static constexpr int VL = 16 / (int)sizeof(R);
  static constexpr int KCP = (MfmaShape<P>::KC + VL - 1) / VL * VL;
  static constexpr int NKP = KCP * 4;  // entries per row
};

// T(d) of offset key okey[o] into row o of T, in fragment order: block
// (rb, kc) holds T[rb*16 + l%16][kc*4 + l/16] at lane l.  T depends on the
// key alone (offset in units of u and, through its residue, the level
// relation), so rows are kept across evaluations.
template <int P, typename R>
__global__ void __launch_bounds__(64) k_lat_tmat(const int *__restrict__ okey, R *__restrict__ T) {
  using S = MfmaShape<P>;
  constexpr int NT = S::NT;
  __shared__ cplx sS[P * P];
  const int o = blockIdx.x, lane = threadIdx.x;
  const int key = okey[o];
  const double dx = key / (kLatSide * kLatSide) - kLatD, dy = (key / kLatSide) % kLatSide - kLatD,
               dz = key % kLatSide - kLatD;
  const Geom g = make_geom(dx, dy, dz);
  // Source multipoles arrive scaled by their own cell's u^-n; u_s = 2^D u_t
  // with D = 0, +1, -1 for offsets = 0, 2, odd mod 4, so column (n, m) of T
  // takes the exact factor 2^(D n).
  const int rx = ((key / (kLatSide * kLatSide)) - kLatD) & 3;
  const int dlev = rx == 0 ? 0 : (rx == 2 ? 1 : -1);
  for (int t = lane; t < NT; t += 64) {
    int a, b;
    coef_nm(t, a, b);
    double re, im;
    singular_nm(a, b, g, re, im);
    const double sg = (b & 1) ? -1.0 : 1.0;  // S'_a^b = (-1)^b S_a^b (b > 0)
    sS[a * a + a + b] = {sg * re, sg * im};
    if (b > 0) sS[a * a + a - b] = {re, -im};  // S'_a^-b = S_a^-b = conj S_a^b
  }
  __syncthreads();
  for (int rb = 0; rb < S::RB; rb++)
    for (int kc = 0; kc < S::kc_end(rb); kc++) {
      const int row = rb * 16 + (lane & 15), col = kc * 4 + (lane >> 4);
      double v = 0;
      if (row < S::NR && col < S::NR) {
        int j, k, n, m;
        coef_nm(row >> 1, j, k);
        coef_nm(col >> 1, n, m);
        if (j + n < P) {
          const int a = j + n;
          const cplx S1 = sS[a * a + a + m - k];
          const cplx S2 = m > 0 ? sS[a * a + a - m - k] : cplx{0, 0};
          const double sm = (m & 1) ? -1.0 : 1.0, sj = (j & 1) ? -1.0 : 1.0;
          const int pr = row & 1, pc = col & 1;
          if (pr == 0 && pc == 0) v = sm * S1.re + S2.re;
          if (pr == 0 && pc == 1) v = -sm * S1.im + S2.im;
          if (pr == 1 && pc == 0) v = sm * S1.im + S2.im;
          if (pr == 1 && pc == 1) v = sm * S1.re - S2.re;
          v = ldexp(v * sj, dlev * n);
        }
      }
      T[(size_t)o * TPad<P, R>::TEP + (S::base(rb) + kc) * 64 + lane] = (R)v;
    }
  for (int i = S::NBLK * 64 + lane; i < TPad<P, R>::TEP; i += 64) T[(size_t)o * TPad<P, R>::TEP + i] = 0;
}

// Target tiles: cells with M2L pairs grouped by (level, octant), and inside
// a group ordered by face class, then cell order (a radix sort keeps cell
// order among equal keys).  The face class is the six distances of the cell
// to the faces of its level's grid, each clipped at kTileK: cells near the
// domain boundary lack the offsets that point outside it, and cells of one
// class lack the same ones, so tiles cut from class-ordered cells have
// fewer steps with idle columns.  The class only orders work: any order
// gives the same pairs.
constexpr int kTileK = 4;
constexpr int kTileClassBits = 14;  // (kTileK + 1)^6 = 15625 classes
static_assert((kTileK + 1) * (kTileK + 1) * (kTileK + 1) * (kTileK + 1) * (kTileK + 1) * (kTileK + 1) <=
                  (1 << kTileClassBits),
              "face classes fit their key bits");
constexpr int kTileNoGroup = 255 << kTileClassBits;  // cells without pairs sort last

// Face class of the cell whose Morton index on its level's grid is kk
// (coordinate bits interleaved x, y, z from bit 0).
__device__ inline int face_class(uint64_t kk, int l) {
  int g[3] = {0, 0, 0};
  for (int i = 0; i < l && i < 21; i++)
    for (int a = 0; a < 3; a++) g[a] |= (int)((kk >> (3 * i + a)) & 1) << i;
  const int64_t side = (int64_t)1 << (l < 21 ? l : 21);
  int cls = 0;
  for (int a = 0; a < 3; a++) {
    const int lo = (int)min<int64_t>(g[a], kTileK), hi = (int)min<int64_t>(side - 1 - g[a], kTileK);
    cls = (cls * (kTileK + 1) + lo) * (kTileK + 1) + hi;
  }
  return cls;
}

__global__ void k_tile_key(int nc, const int *__restrict__ has, const int *__restrict__ lvl,
                           const uint64_t *__restrict__ ckey, bool key_lvloff, int *__restrict__ key,
                           int *__restrict__ val) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int k = kTileNoGroup;
  const int l = lvl[c];
  if (has[c] && l > 0 && l < 31) {
    const uint64_t lvloff = key_lvloff ? ((1ull << (3 * l)) - 1) / 7 : 0;
    const uint64_t kk = ckey[c] - lvloff;
    k = ((l * 8 + (int)(kk & 7)) << kTileClassBits) | face_class(kk, l);
  }
  key[c] = k;
  val[c] = c;
}

// has[t] = 1 for every target t of the M2L list (plain stores of 1).
__global__ void k_mark_has(const int *__restrict__ tgt, int64_t np, int *__restrict__ has) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < np) has[tgt[e]] = 1;
}

// Group sizes from the sorted keys: hist[g] = end - start of group g.
__global__ void k_tile_hist(int nc, const int *__restrict__ key, int *__restrict__ hist) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nc) return;
  const int k = key[p] >> kTileClassBits;
  if (k >= 255) return;
  if (p == 0 || key[p - 1] >> kTileClassBits != k) atomicSub(&hist[k], p);  // one start per group
  if (p == nc - 1 || key[p + 1] >> kTileClassBits != k) atomicAdd(&hist[k], p + 1);  // one end per group
}

// gb[0..255] = group start in sorted order, gb[256..511] = first chunk of
// the group; tmap[chunk] = tile number (chunks of the 8 octant groups of a
// level interleaved, so consecutive tiles cover one region of space).
__global__ void k_tile_assign(int nc, const int *__restrict__ skey, const int *__restrict__ sval,
                              const int *__restrict__ gb, const int *__restrict__ tmap, int *__restrict__ tile_tgt,
                              int *__restrict__ tile_of, int *__restrict__ col_of) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nc) return;
  const int g = skey[p] >> kTileClassBits, c = sval[p];
  if (g >= 255) {
    tile_of[c] = -1;
    return;
  }
  const int r = p - gb[g];
  const int t = tmap[gb[256 + g] + r / 64], col = r % 64;
  tile_tgt[t * 64 + col] = c;
  tile_of[c] = t;
  col_of[c] = col;
}

// Mark the (tile, offset) steps that have at least one pair; pstep[e] = the
// pair's step tile * noff + o, or -1 (not a lattice pair).
__global__ void k_lat_mark(const int *__restrict__ ptgt, int64_t np, const int *__restrict__ pkey,
                           const int *__restrict__ o_of_key, const int *__restrict__ tile_of, int noff,
                           int *__restrict__ act, int *__restrict__ pstep) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np) return;
  int v = -1;
  const int key = pkey[e];
  const int o = key >= 0 ? o_of_key[key] : -1;
  if (o >= 0) {
    const int tile = tile_of[ptgt[e]];
    if (tile >= 0) {
      v = tile * noff + o;
      act[v] = 1;
    }
  }
  pstep[e] = v;
}

// Compact step list: alist[k] = offset of step k (steps of a tile in offset order).
// Active step i = (tile, offset o): its T row, row_of_key[okey[o]].
__global__ void k_lat_alist(const int *__restrict__ act, const int *__restrict__ aidx, int64_t n, int noff,
                            const int *__restrict__ okey, const int *__restrict__ row_of_key, int *__restrict__ alist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && act[i]) alist[aidx[i]] = row_of_key[okey[i % noff]];
}

// Lattice pairs and their (step, column) slots, in two passes of plain
// stores: (1) each lattice pair's slot aidx[step] * 64 + column (kept in
// pslot[e]) takes the pair's index; (2) a pair whose slot holds another
// index (a repeated source: same target and offset) and every non-lattice
// pair is irregular, counted per wave into *nirr, and the pair a slot holds
// writes its source cell s there as -2 - s (never equal to a pair index, so
// the comparisons of other pairs stay right in any order).  Slots left at -1
// are empty columns.
__global__ void k_lat_slots(const int *__restrict__ ptgt, int64_t np, const int *__restrict__ col_of,
                            const int *__restrict__ aidx, int *__restrict__ pslot, int *__restrict__ slot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np) return;
  const int v = pslot[e];  // the step, from k_lat_mark
  if (v < 0) return;
  const int k = aidx[v] * 64 + col_of[ptgt[e]];
  pslot[e] = k;
  slot[k] = (int)e;
}

__global__ void k_lat_check(int64_t np, const int *__restrict__ pslot, const int *__restrict__ src,
                            int *__restrict__ slot, int *__restrict__ irr, int *__restrict__ nirr) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ir = false;
  if (e < np) {
    const int k = pslot[e];
    ir = k < 0 || slot[k] != (int)e;
    if (!ir) slot[k] = -2 - src[e];
    irr[e] = ir ? 1 : 0;
  }
  const unsigned long long m = __ballot(ir);
  if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(nirr, (int)__popcll(m));
}

// v into p[0 .. 64 * (*nsteps)) (the step count is on the device).
__global__ void k_fill_slots(int *__restrict__ p, const int *__restrict__ nsteps, int v) {
  const int64_t n = (int64_t)*nsteps * 64;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Offsets present in the list (numbered in key order), and those of them
// with no row in the T table yet.
__global__ void k_lat_flag(const int *__restrict__ hist, const int *__restrict__ row_of_key, int n,
                           int *__restrict__ flag, int *__restrict__ nflag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  const bool used = k < n && hist[k] != 0;
  flag[k] = used ? 1 : 0;
  nflag[k] = used && row_of_key[k] < 0 ? 1 : 0;
}

// New rows base + scan[k] of the T table for the flagged keys.
__global__ void k_tc_assign(const int *__restrict__ nflag, const int *__restrict__ scan, int n, int base,
                            int *__restrict__ row_of_key, int *__restrict__ keys) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && nflag[k]) {
    row_of_key[k] = base + scan[k];
    keys[scan[k]] = k;
  }
}

__global__ void k_lat_number(const int *__restrict__ flag, const int *__restrict__ scan, int n,
                             int *__restrict__ o_of_key, int *__restrict__ okey) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  o_of_key[k] = flag[k] ? scan[k] : -1;
  if (flag[k]) okey[scan[k]] = k;
}

__global__ void k_fill_int(int *__restrict__ p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// CSR of the irregular pairs: offsets from the exclusive scan of the flags.
// The irregular pairs (irr[e] = 1) in emission order.
__global__ void k_irr_compact(const int *__restrict__ irr, const int *__restrict__ scan, int64_t np,
                              const int *__restrict__ tgt, const int *__restrict__ src,
                              const unsigned char *__restrict__ img, int *__restrict__ tgt_irr,
                              int *__restrict__ src_irr, unsigned char *__restrict__ img_irr) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np || !irr[e]) return;
  tgt_irr[scan[e]] = tgt[e];
  src_irr[scan[e]] = src[e];
  img_irr[scan[e]] = img[e];
}

// ---------------------------------------------------------------------------
// Source multipoles scaled by their own cell's u^-n (u = w / kLatUnit), in
// the precision of the contraction, as MsRow rows (zero padded); row nc is
// all zero and stands for an empty column.
template <int P, typename R>
__global__ void __launch_bounds__(256)
    k_lat_mscale(int nc, const double *__restrict__ cRb, const cplx *__restrict__ M, R *__restrict__ Ms) {
  using S = MfmaShape<P>;
  using MR = MsRow<P, R>;
  constexpr int NR = S::NR, NKP = MR::NKP, VL = MR::VL;
  constexpr DegreeTable<P> deg{};
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)(nc + 1) * NKP) return;
  const int c = (int)(i / NKP), p = (int)(i % NKP);
  const int r = p % (4 * VL), kq = r / VL, kc = (p / (4 * VL)) * VL + r % VL;
  const int k = 4 * kc + kq;  // real row of the packed expansion (re, im interleaved)
  double v = 0;
  if (c < nc && k < NR) {
    const double iu = kLatUnit / (2 * cRb[c]);
    double f = 1;
    for (int q = 0; q < deg.n[k >> 1]; q++) f *= iu;
    v = reinterpret_cast<const double *>(M)[(size_t)c * NR + k] * f;
  }
  Ms[i] = (R)v;
}

// Steps of a tile are split into kLatSeg segments; segment s runs on XCD s
// (work item 8 tile + s is taken by workgroups w = item mod gridDim.x, the
// grid is a multiple of 8, and workgroups are dealt round-robin over the
// 8 XCDs), so an XCD only ever needs the T(d) of one eighth of the offsets
// and keeps them in its L2.
constexpr int kLatSeg = 8;
// Words of Handle::hpin the lattice set-up reads back.
constexpr int kPinGhist = 64, kPinNoff = 64 + 256, kPinNnew = 64 + 257, kPinSteps = 64 + 258;
static_assert(kPinSteps < Handle::kPinWords, "pinned words");

typedef __attribute__((address_space(3))) void lds_void;
// s_waitcnt immediate for vmcnt(0) with expcnt and lgkmcnt left alone.
constexpr int kVmcnt0 = 0x0F70;
typedef __attribute__((address_space(1))) void glb_void;

// One workgroup (4 waves) per (tile of 64 target cells, segment of the
// tile's steps).  R = double: FP64 MFMA.  R = float (mixed form): FP32
// tables, multipoles and MFMA, the FP32 accumulators flushed into FP64 after
// every step (a step's K-sum is the only FP32 accumulation).  Software
// pipeline: while step i runs on the matrix cores, T of step i + 1 streams
// into the other LDS buffer by LDS DMA (global_load_lds_dwordx4), the
// multipole columns of step i + 1 stream into the other register set and the
// slot indices of step i + 2 are read; one barrier per step.  The two LDS
// buffers are separate arrays, the two column sets separate register arrays,
// and the step loop is unrolled by two, so every access names its buffer
// statically (no register copy between steps).  The segment's partial
// expansions (unscaled) go to part[] for k_lat_reduce.
template <int P, typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(P <= 10 ? 2 : 1, 2)))
    k_m2l_mfma(int nc, int nitems, const int *__restrict__ step0, const int *__restrict__ alist,
               const int *__restrict__ slot, const R *__restrict__ T, const R *__restrict__ Ms,
               double *__restrict__ part) {
  using S = MfmaShape<P>;
  using A = MfmaAcc<R>;
  using TP = TPad<P, R>;
  constexpr bool F32 = sizeof(R) == 4;
  using MR = MsRow<P, R>;
  typedef R vec __attribute__((ext_vector_type(16 / sizeof(R))));
  constexpr int RB = S::RB, KCP = MR::KCP, NKP = MR::NKP;
  constexpr int TEP = TP::TEP, CH = TP::CH, NCH = TP::NCH, VL = MR::VL;
  __shared__ __attribute__((aligned(16))) R sT0[TEP];
  __shared__ __attribute__((aligned(16))) R sT1[TEP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // Items (tile, segment) = (item / 8, item % 8); a grid that is a multiple
  // of 8 keeps every item of a segment on one XCD.
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int tile = item / kLatSeg, seg = item % kLatSeg;
    const int col = wave * 16 + (lane & 15);
    const int kq = lane >> 4;
    const int a0 = step0[tile], a1 = step0[tile + 1];
    const int i0 = a0 + (a1 - a0) * seg / kLatSeg, i1 = a0 + (a1 - a0) * (seg + 1) / kLatSeg;
    typename A::type acc[RB];
    v4d acc64[F32 ? RB : 1];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) acc[rb] = typename A::type{0, 0, 0, 0};
#pragma unroll
    for (int rb = 0; rb < (F32 ? RB : 1); rb++) acc64[rb] = v4d{0, 0, 0, 0};
    // T(d) of step i into an LDS buffer: chunk ch by wave ch % 4.
    auto fetch_T = [&](R *dst, int i) {
      const R *Tn = T + (size_t)alist[i] * TEP + lane * VL;
      for (int ch = wave; ch < NCH; ch += 4)
        __builtin_amdgcn_global_load_lds((glb_void *)(Tn + ch * CH), (lds_void *)(dst + ch * CH), 16, 0, 0);
    };
    R b0[KCP], b1[KCP];
    int sn = -1;
    // Column of slot value sl (-2 - source cell, or -1: row nc of zeros)
    // into v[0 .. KCP): one 16-byte load per VL entries (its kq group of the
    // MsRow row).
    auto load_col = [&](R *v, int sl) {
      const int cell = sl <= -2 ? -2 - sl : nc;
      const vec *src = reinterpret_cast<const vec *>(Ms + (size_t)cell * NKP) + kq;
#pragma unroll
      for (int j = 0; j < KCP / VL; j++) {
        const vec x = src[j * 4];
#pragma unroll
        for (int u = 0; u < VL; u++) v[j * VL + u] = x[u];
      }
    };
    // Step i from buffers CUR = (i - i0) & 1 while step i + 1 is fetched
    // into the other ones.
    auto step = [&](auto curc, int i) {
      constexpr int CUR = decltype(curc)::value;
      const R *sTc = CUR ? sT1 : sT0;
      R *sTn = CUR ? sT0 : sT1;
      const R *b = CUR ? b1 : b0;
      R *bn = CUR ? b0 : b1;
      const bool more = i + 1 < i1;
      int snn = -1;
      if (more) {
        load_col(bn, sn);
        fetch_T(sTn, i + 1);
        if (i + 2 < i1) snn = slot[(size_t)(i + 2) * 64 + col];
      }
#pragma unroll
      for (int rb = 0; rb < RB; rb++) {
#pragma unroll
        for (int kc = 0; kc < S::kc_end(rb); kc++)
          acc[rb] = A::mac(sTc[(S::base(rb) + kc) * 64 + lane], b[kc], acc[rb]);
      }
      if (F32) {
#pragma unroll
        for (int rb = 0; rb < (F32 ? RB : 1); rb++)
          for (int r = 0; r < 4; r++) {
            acc64[rb][r] += (double)acc[rb][r];
            acc[rb][r] = 0;
          }
      }
      if (more) sn = snn;
      __builtin_amdgcn_s_waitcnt(kVmcnt0);  // this wave's DMA and loads have landed
      __syncthreads();
    };
    // Prologue: step i0's T and columns.
    if (i0 < i1) {
      fetch_T(sT0, i0);
      sn = slot[(size_t)i0 * 64 + col];
      load_col(b0, sn);
      sn = i0 + 1 < i1 ? slot[(size_t)(i0 + 1) * 64 + col] : -1;
    }
    __builtin_amdgcn_s_waitcnt(kVmcnt0);
    __syncthreads();
    for (int i = i0; i < i1; i += 2) {
      step(std::integral_constant<int, 0>{}, i);
      if (i + 1 < i1) step(std::integral_constant<int, 1>{}, i + 1);
    }
    double *pp = part + (size_t)item * (RB * 16) * 64;
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = rb * 16 + d_row<R>(lane, r);
        const double v = F32 ? acc64[F32 ? rb : 0][r] : (double)acc[rb][r];
        pp[(size_t)row * 64 + col] = v;
      }
  }
}

// L of each tile column += u^-(j+1) * (sum of the segments' partials, in
// segment order), u = w / kLatUnit of the column's own target cell (the cells
// of a tile share a level, not necessarily a size: batched systems).
template <int P>
__global__ void __launch_bounds__(256)
    k_lat_reduce(const int *__restrict__ tile_tgt, const double *__restrict__ cRb, const double *__restrict__ part,
                 cplx *__restrict__ L) {
  using S = MfmaShape<P>;
  constexpr int NR = S::NR, NRP = S::RB * 16;
  constexpr DegreeTable<P> deg{};
  const int tile = blockIdx.x;
  double *Lr = reinterpret_cast<double *>(L);
  // Rows r0 .. r0 + 15 of the tile: the segment sums read column-fastest
  // (as the partials lie), then written row-fastest through LDS (the 16 rows
  // of one target are 128 contiguous bytes of L).
  __shared__ double sv[16][65];
  const int r0 = 16 * blockIdx.y, nrow = min(NR, r0 + 16) - r0;
  const int col = threadIdx.x & 63, rq = threadIdx.x >> 6;
  double v[4] = {0, 0, 0, 0};
  const double *pp = part + ((size_t)tile * kLatSeg * NRP + r0) * 64 + col;
  for (int sg = 0; sg < kLatSeg; sg++)
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (rq + 4 * k < nrow) v[k] += pp[((size_t)sg * NRP + rq + 4 * k) * 64];
#pragma unroll
  for (int k = 0; k < 4; k++) sv[rq + 4 * k][col] = v[k];
  __syncthreads();
  const int row = threadIdx.x & 15;
  if (row >= nrow) return;
  for (int c = threadIdx.x >> 4; c < 64; c += 16) {
    const int tgt = tile_tgt[tile * 64 + c];
    if (tgt < 0) continue;
    const double iw = 1.0 / (2 * cRb[tgt] / kLatUnit);
    double f = iw;
    for (int q = 0; q < deg.n[(r0 + row) >> 1]; q++) f *= iw;
    Lr[(size_t)tgt * NR + r0 + row] += sv[row][c] * f;
  }
}

// ---------------------------------------------------------------------------
template <int P>
static void mscale_impl(Handle &h, hipStream_t st) {
  using S = MfmaShape<P>;
  const int nc = h.ncells;
  const int64_t nms = (int64_t)(nc + 1) * MsRow<P, float>::NKP;
  const int64_t nmd = (int64_t)(nc + 1) * MsRow<P, double>::NKP;
  if (h.cfg.precision == 1) {
    h.lat_Msf.resize(nms);
    hipLaunchKernelGGL((k_lat_mscale<P, float>), dim3(div_up(nms, 256)), dim3(256), 0, st, nc, h.cRb.p, h.M.p,
                       h.lat_Msf.p);
  } else {
    h.lat_Ms.resize(nmd);
    hipLaunchKernelGGL((k_lat_mscale<P, double>), dim3(div_up(nmd, 256)), dim3(256), 0, st, nc, h.cRb.p, h.M.p,
                       h.lat_Ms.p);
  }
}

// Largest slot array sized by the all-steps-active bound (512 MB of int).
constexpr int64_t kSlotBound = (int64_t)1 << 27;

// An empty T table for expansion order P in FP32 or FP64.
static void tc_reset(Handle &h, int P, bool f32) {
  h.lat_tc_map.resize(kLatKeys + 1);
  h.lat_tc_flag.resize(kLatKeys + 1);
  h.lat_tc_scan.resize(kLatKeys + 1);
  hipLaunchKernelGGL(k_fill_int, dim3(div_up(kLatKeys + 1, 256)), dim3(256), 0, h.stream, h.lat_tc_map.p,
                     (int64_t)kLatKeys + 1, -1);
  h.lat_tc_rows = 0;
  h.lat_tc_P = P;
  h.lat_tc_f32 = f32;
}

// Grow a buffer to `want` elements, keeping its first `keep`.
template <typename R>
static void grow_keep_rows(Handle &h, DBuf<R> &b, size_t want, size_t keep) {
  if (want <= b.cap) return;
  DBuf<R> nb;
  nb.reserve(std::max(want, 2 * b.cap));
  if (keep) EXA_HIP(hipMemcpyAsync(nb.p, b.p, keep * sizeof(R), hipMemcpyDeviceToDevice, h.stream));
  EXA_HIP(hipStreamSynchronize(h.stream));
  b.release();
  b.p = nb.p;
  b.cap = nb.cap;
  b.n = want;
  nb.p = nullptr;
  nb.cap = 0;
}

// The contraction and the reduction of the last lattice set-up (its step
// lists, slots and tiles), on the main stream.
template <int P>
static void lattice_contract(Handle &h) {
  const int nitems = h.lat_run_ntiles * kLatSeg;
  if (h.cfg.precision == 1)
    hipLaunchKernelGGL((k_m2l_mfma<P, float>), dim3(h.lat_run_ngrid), dim3(256), 0, h.stream, h.ncells, nitems,
                       h.lat_step0.p, h.lat_alist.p, h.lat_slot.p, h.lat_Tf.p, h.lat_Msf.p, h.lat_part.p);
  else
    hipLaunchKernelGGL((k_m2l_mfma<P, double>), dim3(h.lat_run_ngrid), dim3(256), 0, h.stream, h.ncells, nitems,
                       h.lat_step0.p, h.lat_alist.p, h.lat_slot.p, h.lat_T.p, h.lat_Ms.p, h.lat_part.p);
  EXA_HIP(hipGetLastError());
}

template <int P>
static void lattice_reduce(Handle &h) {
  hipLaunchKernelGGL(k_lat_reduce<P>, dim3(h.lat_run_ntiles, MfmaShape<P>::RB), dim3(256), 0, h.stream,
                     h.lat_tile_tgt.p, h.cRb.p, h.lat_part.p, h.L.p);
  EXA_HIP(hipGetLastError());
}

template <int P>
static void lattice_begin(Handle &h) {
  const int nc = h.ncells;
  const int64_t np = h.nm2l;
  const int TPB = 256;
  const int gp = div_up(np, TPB);
  const bool f32 = h.cfg.precision == 1;
  // 0. Source multipoles scaled per cell (independent of the pair lists;
  // already queued behind the multipoles when the upward pass was split).
  if (!h.lat_ms_ready) mscale_impl<P>(h, h.stream);
  h.lat_ms_ready = false;
  // 1. Classify pairs (offset key, target cell) and flag the offsets used.
  // The raw (emission-order) M2L list: target, source, image per pair.
  const int *ptgt = h.raw_mi.p, *psrc = h.raw_mj.p;
  const unsigned char *pimg = h.raw_mimg.p;
  h.lat_pkey.resize(np > 0 ? np : 1);
  h.lat_hist.resize(kLatKeys);
  hipLaunchKernelGGL(k_fill_int, dim3(div_up(kLatKeys, TPB)), dim3(TPB), 0, h.stream, h.lat_hist.p, (int64_t)kLatKeys,
                     0);
  hipLaunchKernelGGL(k_lat_classify, dim3(gp), dim3(TPB), 0, h.stream, ptgt, psrc, pimg, np,
                     h.cX.p, h.cRb.p, h.lat_pkey.p, h.lat_hist.p);
  // 2. Offsets present, numbered in key order; offsets without a T row.
  if (h.lat_tc_P != P || h.lat_tc_f32 != f32) tc_reset(h, P, f32);
  h.lat_flag.resize(kLatKeys + 1);
  h.lat_kscan.resize(kLatKeys + 1);
  h.lat_ookey.resize(kLatKeys);
  h.lat_okey.resize(kLatKeys);
  hipLaunchKernelGGL(k_lat_flag, dim3(div_up(kLatKeys + 1, TPB)), dim3(TPB), 0, h.stream, h.lat_hist.p,
                     h.lat_tc_map.p, kLatKeys, h.lat_flag.p, h.lat_tc_flag.p);
  exclusive_scan_int(h, h.lat_flag.p, h.lat_kscan.p, kLatKeys + 1);
  exclusive_scan_int(h, h.lat_tc_flag.p, h.lat_tc_scan.p, kLatKeys + 1);
  hipLaunchKernelGGL(k_lat_number, dim3(div_up(kLatKeys, TPB)), dim3(TPB), 0, h.stream, h.lat_flag.p, h.lat_kscan.p,
                     kLatKeys, h.lat_ookey.p, h.lat_okey.p);
  // 3. Target tiles of one level and one octant (face class, then cell order).
  h.lat_tkey.resize(2 * (size_t)nc);
  h.lat_tval.resize(2 * (size_t)nc);
  h.lat_thist.resize(256);
  int *tk = h.lat_tkey.p, *tv = h.lat_tval.p;
  h.lat_has.resize(nc);
  hipLaunchKernelGGL(k_fill_int, dim3(div_up(nc, TPB)), dim3(TPB), 0, h.stream, h.lat_has.p, (int64_t)nc, 0);
  hipLaunchKernelGGL(k_mark_has, dim3(gp), dim3(TPB), 0, h.stream, ptgt, np, h.lat_has.p);
  hipLaunchKernelGGL(k_tile_key, dim3(div_up(nc, TPB)), dim3(TPB), 0, h.stream, nc, h.lat_has.p, h.cLevel.p, h.cKey.p,
                     h.key_lvloff, tk, tv);
  bool alt = false;
  sort_pairs_int_int(h, tk, tk + nc, tv, tv + nc, nc, 8 + kTileClassBits, alt);
  h.lat_tile_alt = alt;
  const int *skey = alt ? tk + nc : tk;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, h.stream, h.lat_thist.p, (int64_t)256, 0);
  hipLaunchKernelGGL(k_tile_hist, dim3(div_up(nc, TPB)), dim3(TPB), 0, h.stream, nc, skey, h.lat_thist.p);
  // Tile histogram, offset count and new-row count for the host (read by
  // lattice_run after the main stream has passed these copies).
  EXA_HIP(hipMemcpyAsync(h.hpin + kPinGhist, h.lat_thist.p, 256 * sizeof(int), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipMemcpyAsync(h.hpin + kPinNoff, h.lat_kscan.p + kLatKeys, sizeof(int), hipMemcpyDeviceToHost, h.stream));
  EXA_HIP(hipMemcpyAsync(h.hpin + kPinNnew, h.lat_tc_scan.p + kLatKeys, sizeof(int), hipMemcpyDeviceToHost,
                         h.stream));
  h.lat_begun = true;
}

// The rest of the lattice M2L, after lattice_begin (which the evaluation
// queues before P2P, so the set-up's first kernels start while the P2P
// list is grouped, before P2P fills the GPU).
template <int P>
static void lattice_run(Handle &h, M2LList &irr) {
  using S = MfmaShape<P>;
  const int nc = h.ncells;
  const int64_t np = h.nm2l;
  const int TPB = 256;
  const int gp = div_up(np, TPB);
  const bool f32 = h.cfg.precision == 1;
  if (!h.lat_begun) lattice_begin<P>(h);
  h.lat_begun = false;
  const int *ptgt = h.raw_mi.p, *psrc = h.raw_mj.p;
  const unsigned char *pimg = h.raw_mimg.p;
  const int *tk = h.lat_tkey.p, *tv = h.lat_tval.p;
  const int *skey = h.lat_tile_alt ? tk + nc : tk, *sval = h.lat_tile_alt ? tv + nc : tv;
  EXA_HIP(hipStreamSynchronize(h.stream));
  int ghist[256];
  std::copy(h.hpin + kPinGhist, h.hpin + kPinGhist + 256, ghist);
  const int noff = h.hpin[kPinNoff];
  int nnew = h.hpin[kPinNnew];
  if (h.lat_tc_rows + nnew > h.lat_tc_max) {  // start the table afresh
    tc_reset(h, P, f32);
    hipLaunchKernelGGL(k_lat_flag, dim3(div_up(kLatKeys + 1, TPB)), dim3(TPB), 0, h.stream, h.lat_hist.p,
                       h.lat_tc_map.p, kLatKeys, h.lat_flag.p, h.lat_tc_flag.p);
    exclusive_scan_int(h, h.lat_tc_flag.p, h.lat_tc_scan.p, kLatKeys + 1);
    nnew = noff;
  }
  // T(d) of the offsets not in the table yet: runs while the host lays out
  // the tiles and steps.
  if (nnew > 0) {
    h.lat_tc_keys.resize(nnew);
    hipLaunchKernelGGL(k_tc_assign, dim3(div_up(kLatKeys, TPB)), dim3(TPB), 0, h.stream, h.lat_tc_flag.p,
                       h.lat_tc_scan.p, kLatKeys, h.lat_tc_rows, h.lat_tc_map.p, h.lat_tc_keys.p);
    if (f32) {
      constexpr size_t TEP = TPad<P, float>::TEP;
      grow_keep_rows(h, h.lat_Tf, (size_t)(h.lat_tc_rows + nnew) * TEP, (size_t)h.lat_tc_rows * TEP);
      hipLaunchKernelGGL((k_lat_tmat<P, float>), dim3(nnew), dim3(64), 0, h.stream, h.lat_tc_keys.p,
                         h.lat_Tf.p + (size_t)h.lat_tc_rows * TEP);
    } else {
      constexpr size_t TEP = TPad<P, double>::TEP;
      grow_keep_rows(h, h.lat_T, (size_t)(h.lat_tc_rows + nnew) * TEP, (size_t)h.lat_tc_rows * TEP);
      hipLaunchKernelGGL((k_lat_tmat<P, double>), dim3(nnew), dim3(64), 0, h.stream, h.lat_tc_keys.p,
                         h.lat_T.p + (size_t)h.lat_tc_rows * TEP);
    }
    h.lat_tc_rows += nnew;
  }
  int gb[512], ntiles = 0, start = 0;
  for (int g = 0; g < 256; g++) {
    gb[g] = start;
    gb[256 + g] = ntiles;
    start += g < 255 ? ghist[g] : 0;
    ntiles += g < 255 ? div_up(ghist[g], 64) : 0;
  }
  // Tile numbers: per level, chunk q of octant 0..7, then chunk q + 1, ...
  std::vector<int> tmap(ntiles > 0 ? ntiles : 1);
  for (int l = 0, id = 0; l < 32; l++) {
    int qmax = 0;
    for (int o = 0; o < 8 && l * 8 + o < 255; o++) qmax = std::max(qmax, div_up(ghist[l * 8 + o], 64));
    for (int q = 0; q < qmax; q++)
      for (int o = 0; o < 8 && l * 8 + o < 255; o++) {
        const int g = l * 8 + o;
        if (q < div_up(ghist[g], 64)) tmap[gb[256 + g] + q] = id++;
      }
  }
  h.lat_gb.resize(512 + tmap.size());
  EXA_HIP(hipMemcpyAsync(h.lat_gb.p, gb, sizeof(gb), hipMemcpyHostToDevice, h.stream));
  EXA_HIP(hipMemcpyAsync(h.lat_gb.p + 512, tmap.data(), tmap.size() * sizeof(int), hipMemcpyHostToDevice, h.stream));
  h.lat_tile_tgt.resize((size_t)(ntiles > 0 ? ntiles : 1) * 64);
  h.lat_tile_of.resize(nc);
  h.lat_col_of.resize(nc);
  if (ntiles)
    hipLaunchKernelGGL(k_fill_int, dim3(div_up((int64_t)ntiles * 64, TPB)), dim3(TPB), 0, h.stream, h.lat_tile_tgt.p,
                       (int64_t)ntiles * 64, -1);
  hipLaunchKernelGGL(k_tile_assign, dim3(div_up(nc, TPB)), dim3(TPB), 0, h.stream, nc, skey, sval, h.lat_gb.p,
                     h.lat_gb.p + 512, h.lat_tile_tgt.p, h.lat_tile_of.p, h.lat_col_of.p);
  // 4. Steps: (tile, offset) with at least one pair, compacted per tile.
  const int64_t nsteps_all = (int64_t)ntiles * noff;
  h.lat_act.resize(nsteps_all + 1);
  h.lat_aidx.resize(nsteps_all + 1);
  hipLaunchKernelGGL(k_fill_int, dim3(div_up(nsteps_all + 1, TPB)), dim3(TPB), 0, h.stream, h.lat_act.p,
                     nsteps_all + 1, 0);
  h.lat_pslot.resize(np);
  hipLaunchKernelGGL(k_lat_mark, dim3(gp), dim3(TPB), 0, h.stream, ptgt, np, h.lat_pkey.p, h.lat_ookey.p,
                     h.lat_tile_of.p, noff, h.lat_act.p, h.lat_pslot.p);
  exclusive_scan_int(h, h.lat_act.p, h.lat_aidx.p, (int)(nsteps_all + 1));
  // Step lists sized for every (tile, offset) being active, so the host need
  // not wait for the active count (the device has it in aidx[nsteps_all]),
  // unless that bound is large; then the count is read back.
  int64_t nact = nsteps_all;
  if (nsteps_all * 64 > kSlotBound) {
    int na = 0;
    EXA_HIP(hipMemcpyAsync(&na, h.lat_aidx.p + nsteps_all, sizeof(int), hipMemcpyDeviceToHost, h.stream));
    EXA_HIP(hipStreamSynchronize(h.stream));
    nact = na;
  }
  if (nact * 64 > INT_MAX) throw Error(EXAFMM_HIP_EUNSUP, "lattice M2L: too many (tile, offset) steps");
  h.lat_alist.resize(nact > 0 ? nact : 1);
  h.lat_step0.resize(ntiles + 1);
  if (nsteps_all) {
    hipLaunchKernelGGL(k_lat_alist, dim3(div_up(nsteps_all, TPB)), dim3(TPB), 0, h.stream, h.lat_act.p, h.lat_aidx.p,
                       nsteps_all, noff, h.lat_okey.p, h.lat_tc_map.p, h.lat_alist.p);
    // step0[t] = aidx[t * noff]: a strided copy.
    EXA_HIP(hipMemcpy2DAsync(h.lat_step0.p, sizeof(int), h.lat_aidx.p, (size_t)noff * sizeof(int), sizeof(int),
                             ntiles + 1, hipMemcpyDeviceToDevice, h.stream));
  }
  // 5. Slots of the lattice pairs; the rest (and repeats) are irregular.
  const int64_t nslot = nact * 64;
  h.lat_slot.resize(nslot > 0 ? nslot : 1);
  if (nslot)
    hipLaunchKernelGGL(k_fill_slots, dim3(std::min<int64_t>(div_up(nslot, TPB), 4096)), dim3(TPB), 0, h.stream,
                       h.lat_slot.p, h.lat_aidx.p + nsteps_all, -1);
  h.lat_irr.resize(np + 1);
  h.lat_cnt.resize(1);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, h.stream, h.lat_cnt.p, (int64_t)1, 0);
  hipLaunchKernelGGL(k_lat_slots, dim3(gp), dim3(TPB), 0, h.stream, ptgt, np, h.lat_col_of.p, h.lat_aidx.p,
                     h.lat_pslot.p, h.lat_slot.p);
  hipLaunchKernelGGL(k_lat_check, dim3(gp), dim3(TPB), 0, h.stream, np, h.lat_pslot.p, psrc, h.lat_slot.p,
                     h.lat_irr.p, h.lat_cnt.p);
  // The irregular count (and the step count, for the statistics) is read
  // back behind the contraction's launch (pinned host words; the copies
  // complete before the contraction starts).
  EXA_HIP(hipMemcpyAsync(h.hpin, h.lat_cnt.p, sizeof(int), hipMemcpyDeviceToHost, h.stream));
  if (nsteps_all)
    EXA_HIP(hipMemcpyAsync(h.hpin + kPinSteps, h.lat_aidx.p + nsteps_all, sizeof(int), hipMemcpyDeviceToHost,
                           h.stream));
  EXA_HIP(hipEventRecord(h.ev_cnt, h.stream));
  h.lat_stats_noff = noff;
  h.lat_stats_tiles = ntiles;
  // 6. The contraction (timed by ev[12..13]) and the reduction.
  const bool run_mfma = noff > 0 && ntiles > 0 && nsteps_all > 0;
  join_multipoles(h);  // the set-up above did not need them
  h.lat_run_ntiles = 0;
  if (run_mfma) {
    h.lat_part.resize((size_t)ntiles * kLatSeg * S::RB * 16 * 64);
    // One workgroup per item, or lat_wg_cu workgroups per CU looping over
    // the items (room left on every CU for the P2P waves of the other stream).
    const int nitems = ntiles * kLatSeg;
    int ngrid = nitems;
    if (h.lat_wg_cu > 0) ngrid = std::min(ngrid, std::max(kLatSeg, (h.lat_wg_cu * h.ncu) / kLatSeg * kLatSeg));
    h.lat_run_ntiles = ntiles;
    h.lat_run_ngrid = ngrid;
    if (h.serial_heavy) EXA_HIP(hipStreamWaitEvent(h.stream, h.ev[10], 0));  // P2P's end (second stream)
    EXA_HIP(hipEventRecord(h.ev[12], h.stream));
    h.lat_timed = true;
    lattice_contract<P>(h);
    EXA_HIP(hipEventRecord(h.ev[13], h.stream));
    lattice_reduce<P>(h);
  }
  EXA_HIP(hipEventSynchronize(h.ev_cnt));
  const int nirr = *h.hpin;
  h.lat_stats_steps = nsteps_all ? h.hpin[kPinSteps] : 0;
  // 7. The irregular pairs (none for a uniform free-space tree), in
  // emission order, grouped by target (a stable sort: within a target the
  // order of the full list grouped by target), queued behind the contraction.
  int nitl = 0;
  if (nirr) {
    h.lat_scan.resize(np + 1);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, h.stream, h.lat_irr.p + np, (int64_t)1, 0);
    exclusive_scan_int(h, h.lat_irr.p, h.lat_scan.p, (int)(np + 1));
    h.irr_rt.resize(nirr);
    h.irr_rs.resize(nirr);
    h.irr_ri.resize(nirr);
    hipLaunchKernelGGL(k_irr_compact, dim3(gp), dim3(TPB), 0, h.stream, h.lat_irr.p, h.lat_scan.p, np, ptgt, psrc,
                       pimg, h.irr_rt.p, h.irr_rs.p, h.irr_ri.p);
    group_pairs(h, nirr, h.irr_rt, h.irr_rs, h.irr_ri, h.irr_off, h.irr_src, h.irr_img, h.irr_tgt, nitl);
  }
  irr = M2LList{h.irr_tgt.p, nitl, h.irr_off.p, h.irr_src.p, h.irr_img.p};
  h.lat_stats_pairs = np - nirr;
  EXA_HIP(hipGetLastError());
}

bool lattice_applies(const Handle &h) {
  return h.cfg.basis == 0 && !h.cfg.mass && h.lattice_m2l && h.cfg.P >= 4 && h.cfg.P <= 12 && (h.cfg.P % 2) == 0;
}

// Scaled source multipoles for the lattice contraction, queued on stream st
// right behind the multipoles (the split upward pass).
void run_lattice_mscale(Handle &h, hipStream_t st) {
  if (!lattice_applies(h)) return;
  switch (h.cfg.P) {
    case 4: mscale_impl<4>(h, st); break;
    case 6: mscale_impl<6>(h, st); break;
    case 8: mscale_impl<8>(h, st); break;
    case 10: mscale_impl<10>(h, st); break;
    case 12: mscale_impl<12>(h, st); break;
  }
  EXA_HIP(hipGetLastError());
  h.lat_ms_ready = true;
}

void m2l_begin(Handle &h) {
  h.lat_begun = false;
  if (!lattice_applies(h) || h.nm2l == 0) {
    h.lat_ms_ready = false;  // queued scaled multipoles (if any) go unused
    return;
  }
  switch (h.cfg.P) {
    case 4: lattice_begin<4>(h); break;
    case 6: lattice_begin<6>(h); break;
    case 8: lattice_begin<8>(h); break;
    case 10: lattice_begin<10>(h); break;
    case 12: lattice_begin<12>(h); break;
  }
  EXA_HIP(hipGetLastError());
}

// The lattice contraction and reduction again on the current multipoles,
// with the set-up of the last run_m2l_lattice (same tree and lists).
void lattice_replay(Handle &h) {
  if (!lattice_applies(h) || h.nm2l == 0 || h.lat_run_ntiles == 0) return;
  switch (h.cfg.P) {
    case 4: lattice_contract<4>(h); lattice_reduce<4>(h); break;
    case 6: lattice_contract<6>(h); lattice_reduce<6>(h); break;
    case 8: lattice_contract<8>(h); lattice_reduce<8>(h); break;
    case 10: lattice_contract<10>(h); lattice_reduce<10>(h); break;
    case 12: lattice_contract<12>(h); lattice_reduce<12>(h); break;
  }
}

bool run_m2l_lattice(Handle &h, M2LList &irr) {
  h.lat_timed = false;
  h.lat_run_ntiles = 0;
  h.lat_stats_pairs = 0;
  h.lat_stats_noff = h.lat_stats_tiles = 0;
  h.lat_stats_steps = 0;
  if (!lattice_applies(h) || h.nm2l == 0) {
    h.lat_ms_ready = false;
    return false;
  }
  switch (h.cfg.P) {
    case 4: lattice_run<4>(h, irr); return true;
    case 6: lattice_run<6>(h, irr); return true;
    case 8: lattice_run<8>(h, irr); return true;
    case 10: lattice_run<10>(h, irr); return true;
    case 12: lattice_run<12>(h, irr); return true;
    default: return false;
  }
}

}  // namespace exafmm
