// Small fixed-size vector used by the host API (the role of the reference's
// include/vec.h generic class, vec.h:8-212): element access, arithmetic with
// vectors and scalars, norm (sum of squares), element-wise min/max.
#ifndef vec_h
#define vec_h
#include <ostream>

template <int N, typename T>
class vec {
  T d_[N];

 public:
  vec() {}
  vec(const T &v) {
    for (int i = 0; i < N; i++) d_[i] = v;
  }
  template <typename U>
  explicit vec(const vec<N, U> &o) {
    for (int i = 0; i < N; i++) d_[i] = T(o[i]);
  }
  T &operator[](int i) { return d_[i]; }
  const T &operator[](int i) const { return d_[i]; }
  vec &operator=(const T &v) {
    for (int i = 0; i < N; i++) d_[i] = v;
    return *this;
  }
#define EXAFMM_VEC_OP(op)                           \
  vec &operator op##=(const vec & o) {              \
    for (int i = 0; i < N; i++) d_[i] op## = o.d_[i]; \
    return *this;                                   \
  }                                                 \
  vec &operator op##=(const T & v) {                \
    for (int i = 0; i < N; i++) d_[i] op## = v;     \
    return *this;                                   \
  }                                                 \
  vec operator op(const vec &o) const {             \
    vec r(*this);                                   \
    r op## = o;                                     \
    return r;                                       \
  }                                                 \
  vec operator op(const T &v) const {               \
    vec r(*this);                                   \
    r op## = v;                                     \
    return r;                                       \
  }
  EXAFMM_VEC_OP(+)
  EXAFMM_VEC_OP(-)
  EXAFMM_VEC_OP(*)
  EXAFMM_VEC_OP(/)
#undef EXAFMM_VEC_OP
  vec operator-() const {
    vec r;
    for (int i = 0; i < N; i++) r.d_[i] = -d_[i];
    return r;
  }
  friend T sum(const vec &v) {
    T t = 0;
    for (int i = 0; i < N; i++) t += v[i];
    return t;
  }
  friend T norm(const vec &v) {
    T t = 0;
    for (int i = 0; i < N; i++) t += v[i] * v[i];
    return t;
  }
  friend vec min(const vec &a, const vec &b) {
    vec r;
    for (int i = 0; i < N; i++) r[i] = a[i] < b[i] ? a[i] : b[i];
    return r;
  }
  friend vec max(const vec &a, const vec &b) {
    vec r;
    for (int i = 0; i < N; i++) r[i] = a[i] > b[i] ? a[i] : b[i];
    return r;
  }
  friend std::ostream &operator<<(std::ostream &s, const vec &v) {
    for (int i = 0; i < N; i++) s << v[i] << ' ';
    return s;
  }
};
#endif
