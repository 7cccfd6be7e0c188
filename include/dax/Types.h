// Base value types of the Dax API slice served by the HIP adapter.
// Mirrors dax/Types.h:244-268 (Id, Scalar) and the Tuple/Vector3 family.
#ifndef DAX_HIP_SLICE_TYPES_H
#define DAX_HIP_SLICE_TYPES_H

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define DAX_EXEC_EXPORT __device__ __host__ inline
#define DAX_EXEC_CONT_EXPORT __device__ __host__ inline
#else
#define DAX_EXEC_EXPORT inline
#define DAX_EXEC_CONT_EXPORT inline
#endif
#define DAX_CONT_EXPORT inline

namespace dax {

// dax::Id is 32-bit by default (dax/internal/Configure.h.in:67-75); define
// DAX_USE_64BIT_IDS for 64-bit ids.
#ifdef DAX_USE_64BIT_IDS
typedef int64_t Id;
#else
typedef int32_t Id;
#endif
typedef float Scalar;

template <typename T, int N>
struct Tuple {
  enum { NUM_COMPONENTS = N };
  typedef T ComponentType;
  T c[N];
  DAX_EXEC_CONT_EXPORT Tuple() {}
  DAX_EXEC_CONT_EXPORT explicit Tuple(const T& v) {
    for (int i = 0; i < N; ++i) c[i] = v;
  }
  // Three components (dax/Types.h:444-449: Tuple<T,3>(x, y, z)).
  DAX_EXEC_CONT_EXPORT Tuple(T x, T y, T z) {
    static_assert(N == 3, "three components");
    c[0] = x;
    c[1] = y;
    c[2] = z;
  }
  DAX_EXEC_CONT_EXPORT T& operator[](int i) { return c[i]; }
  DAX_EXEC_CONT_EXPORT const T& operator[](int i) const { return c[i]; }
  DAX_EXEC_CONT_EXPORT bool operator==(const Tuple& o) const {
    for (int i = 0; i < N; ++i)
      if (!(c[i] == o.c[i])) return false;
    return true;
  }
  DAX_EXEC_CONT_EXPORT bool operator!=(const Tuple& o) const { return !(*this == o); }
  DAX_EXEC_CONT_EXPORT bool operator<(const Tuple& o) const {
    for (int i = 0; i < N; ++i) {
      if (c[i] < o.c[i]) return true;
      if (o.c[i] < c[i]) return false;
    }
    return false;
  }
};

typedef Tuple<Id, 3> Id3;
typedef Tuple<Scalar, 3> Vector3;

DAX_EXEC_CONT_EXPORT Id3 make_Id3(Id x, Id y, Id z) {
  Id3 r;
  r[0] = x;
  r[1] = y;
  r[2] = z;
  return r;
}
DAX_EXEC_CONT_EXPORT Vector3 make_Vector3(Scalar x, Scalar y, Scalar z) {
  Vector3 r;
  r[0] = x;
  r[1] = y;
  r[2] = z;
  return r;
}

template <typename T, int N>
DAX_EXEC_CONT_EXPORT Tuple<T, N> operator+(const Tuple<T, N>& a, const Tuple<T, N>& b) {
  Tuple<T, N> r;
  for (int i = 0; i < N; ++i) r[i] = a[i] + b[i];
  return r;
}
template <typename T, int N>
DAX_EXEC_CONT_EXPORT Tuple<T, N> operator-(const Tuple<T, N>& a, const Tuple<T, N>& b) {
  Tuple<T, N> r;
  for (int i = 0; i < N; ++i) r[i] = a[i] - b[i];
  return r;
}
template <typename T, int N>
DAX_EXEC_CONT_EXPORT Tuple<T, N> operator*(T s, const Tuple<T, N>& a) {
  Tuple<T, N> r;
  for (int i = 0; i < N; ++i) r[i] = s * a[i];
  return r;
}

DAX_EXEC_CONT_EXPORT Scalar dot(const Vector3& a, const Vector3& b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

}  // namespace dax

#endif
