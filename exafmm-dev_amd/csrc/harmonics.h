// Device-side solid harmonics for the Laplace spherical basis.
//
// Conventions (those of kernels/LaplaceSphericalCPU.cxx:29-101):
//   regular  R_n^m = (-1)^n rho^n / (n+m)!  P_n^m(cos a) e^{i m b}
//   singular S_n^m = (-1)^n (n-m)! rho^{-n-1} P_n^m(cos a) e^{i m b}
// with P_n^m the associated Legendre function including the Condon-Shortley
// phase (P_m^m = (-1)^m (2m-1)!! sin^m a) and X_n^{-m} = conj(X_n^m).
// Coefficient index of (n, m >= 0) in a packed expansion: n(n+1)/2 + m.
//
// The GPU takes cos a = z/r and e^{ib} = (x + i y)/sqrt(x^2 + y^2) straight
// from the Cartesian offset instead of going through acos/atan2.
#pragma once
#include <hip/hip_runtime.h>

namespace exafmm {

template <int P>
struct Expn {
  static constexpr int NT = P * (P + 1) / 2;  // packed terms, m >= 0
};

__host__ __device__ constexpr int tri(int n) { return n * (n + 1) / 2; }

// |a - b| with the reference's rounding: norm() sums x^2, y^2, z^2 left to
// right without contraction (vec.h:163-167), then std::sqrt.
__device__ inline double exact_dist(double ax, double ay, double az, double bx, double by, double bz) {
#pragma clang fp contract(off)
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  double t = 0;
  t += dx * dx;
  t += dy * dy;
  t += dz * dz;
  return sqrt(t);
}

struct Geom {
  double r;     // |dX|
  double ct;    // cos(alpha)
  double st;    // sin(alpha) >= 0
  double cp;    // cos(beta)
  double sp;    // sin(beta)
};

__device__ inline Geom make_geom(double dx, double dy, double dz) {
  Geom g;
  const double rxy2 = dx * dx + dy * dy;
  g.r = sqrt(rxy2 + dz * dz);
  const double rxy = sqrt(rxy2);
  if (g.r == 0) {
    g.ct = 1;
    g.st = 0;
  } else {
    g.ct = dz / g.r;
    g.st = rxy / g.r;
  }
  if (rxy == 0) {
    g.cp = 1;
    g.sp = 0;
  } else {
    g.cp = dx / rxy;
    g.sp = dy / rxy;
  }
  return g;
}

namespace detail {
template <int N>
struct Reciprocals {
  double v[N];
  constexpr Reciprocals() : v() {
    for (int i = 1; i < N; i++) v[i] = 1.0 / i;
  }
};
template <int N>
struct InvFactorials {
  double v[N];
  constexpr InvFactorials() : v() {
    v[0] = 1;
    for (int i = 1; i < N; i++) v[i] = v[i - 1] / i;
  }
};
}  // namespace detail

// 1/k (k < 256) and 1/k! (k < 32), so the recurrences below never divide.
__constant__ constexpr detail::Reciprocals<256> kInv{};
__constant__ constexpr detail::InvFactorials<32> kInvFact{};

__device__ inline double factorial(int k) {
  double f = 1;
  for (int i = 2; i <= k; i++) f *= i;
  return f;
}

// Associated Legendre P_n^m(ct) for one (n, m >= 0), upward recurrence in n.
__device__ inline double legendre(int n, int m, double ct, double st) {
  double pmm = 1;
  for (int i = 1; i <= m; i++) pmm *= -(2 * i - 1) * st;
  double p = pmm, p1 = 0;
  for (int k = m; k < n; k++) {
    const double pn = (ct * (2 * k + 1) * p - (k + m) * p1) * kInv.v[k - m + 1];
    p1 = p;
    p = pn;
  }
  return p;
}

// e^{i m b} by repeated multiplication.
__device__ inline void eim(int m, double cp, double sp, double &c, double &s) {
  c = 1;
  s = 0;
  for (int i = 0; i < m; i++) {
    const double t = c * cp - s * sp;
    s = c * sp + s * cp;
    c = t;
  }
}

// One regular harmonic R_n^m, m >= 0, without divisions:
//   H_m = (r sin(a) / 2)^m / m!,  H_{m+1} = -r cos(a) H_m,
//   H_k = -((2k-1) r cos(a) H_{k-1} + r^2 H_{k-2}) / (k^2 - m^2),
// R_n^m = H_n e^{i m b}.
__device__ inline void regular_nm(int n, int m, const Geom &g, double &re, double &im) {
  const double rs = 0.5 * g.r * g.st;
  double h = kInvFact.v[m];
  for (int i = 0; i < m; i++) h *= rs;
  double hm1 = 0;
  const double cr = g.ct * g.r, r2 = g.r * g.r;
  for (int k = m + 1; k <= n; k++) {
    const double hn = -((2 * k - 1) * cr * h + r2 * hm1) * kInv.v[k * k - m * m];
    hm1 = h;
    h = hn;
  }
  double c, s;
  eim(m, g.cp, g.sp, c, s);
  re = h * c;
  im = h * s;
}

// One singular harmonic S_n^m, m >= 0.
__device__ inline void singular_nm(int n, int m, const Geom &g, double &re, double &im) {
  const double inv_r = 1.0 / g.r;
  double sc = factorial(n - m) * inv_r;
  for (int i = 0; i < n; i++) sc *= inv_r;
  if (n & 1) sc = -sc;
  const double p = legendre(n, m, g.ct, g.st) * sc;
  double c, s;
  eim(m, g.cp, g.sp, c, s);
  re = p * c;
  im = p * s;
}

__device__ inline void coef_nm(int t, int &n, int &m) {
  n = 0;
  while (tri(n + 1) <= t) n++;
  m = t - tri(n);
}

// One output coefficient (j, k) of M2M (LaplaceSphericalCPU.cxx:120-147):
// the child multipole sM (packed, m >= 0) shifted by the regular harmonics
// sY (packed, m >= 0) of X_parent - X_child; adds into (lr, li).
template <typename C>
__device__ inline void m2m_lane(int j, int k, const C *sM, const C *sY, double &lr, double &li) {
  for (int n = 0; n <= j; n++) {
    // m from max(-n, -j+k+n) to min(k-1, n): source (j-n, k-m), Y_n^{-m}
    const int mlo = max(-n, -j + k + n), mhi = min(k - 1, n);
    for (int m = mlo; m <= mhi; m++) {
      const C s = sM[tri(j - n) + k - m];
      C y = sY[tri(n) + (m >= 0 ? m : -m)];
      if (m >= 0) y.im = -y.im;  // Y_n^{-m} = conj(Y_n^m)
      const double f = ((m < 0 && (m & 1)) ? -1.0 : 1.0) * ((n & 1) ? -1.0 : 1.0);
      lr += f * (s.re * y.re - s.im * y.im);
      li += f * (s.re * y.im + s.im * y.re);
    }
    // m from k to min(n, j+k-n): conj(source (j-n, m-k)), Y_n^{-m}
    const int mhi2 = min(n, j + k - n);
    for (int m = k; m <= mhi2; m++) {
      const C s = sM[tri(j - n) - k + m];
      const C y = sY[tri(n) + m];
      // conj(s) * conj(y) = conj(s * y)
      const double f = ((k + n + m) & 1) ? -1.0 : 1.0;
      lr += f * (s.re * y.re - s.im * y.im);
      li -= f * (s.re * y.im + s.im * y.re);
    }
  }
}

// One output coefficient (j, k) of L2L (LaplaceSphericalCPU.cxx:193-224):
// the parent local expansion sL (packed, m >= 0) shifted by the regular
// harmonics sY (packed, m >= 0) of X_child - X_parent; adds into (lr, li).
template <int P, typename C>
__device__ inline void l2l_lane(int j, int k, const C *sL, const C *sY, double &lr, double &li) {
  for (int n = j; n < P; n++) {
    // m from j+k-n to -1: conj(Lp_n^{-m}) Y_{n-j}^{m-k} (-1)^k
    for (int m = j + k - n; m < 0; m++) {
      const C s = sL[tri(n) - m];
      const int a = n - j, b = m - k;  // b < 0
      C y = sY[tri(a) - b];
      y.im = -y.im;
      const double f = (k & 1) ? -1.0 : 1.0;
      lr += f * (s.re * y.re + s.im * y.im);
      li += f * (s.re * y.im - s.im * y.re);
    }
    for (int m = 0; m <= n; m++) {
      const int a = n - j, b = m - k;
      if (a < (b >= 0 ? b : -b)) continue;
      const C s = sL[tri(n) + m];
      C y = sY[tri(a) + (b >= 0 ? b : -b)];
      if (b < 0) y.im = -y.im;
      const double f = (m < k && ((m - k) & 1)) ? -1.0 : 1.0;
      lr += f * (s.re * y.re - s.im * y.im);
      li += f * (s.re * y.im + s.im * y.re);
    }
  }
}

}  // namespace exafmm
