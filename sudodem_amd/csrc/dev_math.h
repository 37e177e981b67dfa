// dev_math.h -- small-vector and quaternion arithmetic for the DEM kernels.
//
// The reference evaluates its 3-vectors with Eigen 3.3.5 built without vectorisation
// (SudoDEM3D/CMakeLists.txt:104), which sums three terms as x0+(x1+x2) and four as
// (x0+x1)+(x2+x3) (lib/Eigen/src/Core/Redux.h:91-105).  These helpers keep that order so
// that, with contraction disabled, additions and products round exactly as the reference's.
#pragma once
#include <hip/hip_runtime.h>

#define SDM_HD __host__ __device__ __forceinline__

#include "crmath.h"

template <typename T>
struct V3 {
  T x, y, z;
};

template <typename T>
SDM_HD V3<T> mk3(T x, T y, T z) {
  V3<T> r;
  r.x = x; r.y = y; r.z = z;
  return r;
}
template <typename T>
SDM_HD V3<T> operator+(V3<T> a, V3<T> b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T>
SDM_HD V3<T> operator-(V3<T> a, V3<T> b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T>
SDM_HD V3<T> operator-(V3<T> a) { return mk3(-a.x, -a.y, -a.z); }
template <typename T>
SDM_HD V3<T> operator*(V3<T> a, T s) { return mk3(a.x * s, a.y * s, a.z * s); }
template <typename T>
SDM_HD V3<T> operator*(T s, V3<T> a) { return mk3(s * a.x, s * a.y, s * a.z); }
template <typename T>
SDM_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
template <typename T>
SDM_HD T sqn(V3<T> a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
template <typename T>
SDM_HD V3<T> cross(V3<T> a, V3<T> b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SDM_HD double dsqrt(double x) { return sqrt(x); }
SDM_HD float dsqrt(float x) { return sqrtf(x); }
template <typename T>
SDM_HD T norm(V3<T> a) { return dsqrt(sqn(a)); }
// MatrixBase::normalize: divide by the norm, leave a zero vector unchanged (Dot.h:145-151).
template <typename T>
SDM_HD V3<T> normalized(V3<T> a) {
  T z = sqn(a);
  if (z > T(0)) {
    T s = dsqrt(z);
    return mk3(a.x / s, a.y / s, a.z / s);
  }
  return a;
}
// Mathr::Sign (lib/base/Math.hpp:186).
template <typename T>
SDM_HD T sgn(T f) { return f < T(0) ? T(-1) : (f > T(0) ? T(1) : T(0)); }

// Row-major 3x3 matrix.
template <typename T>
struct M3 {
  T m[9];
};
template <typename T>
SDM_HD V3<T> mul(const M3<T>& M, V3<T> v) {
  return mk3(M.m[0] * v.x + (M.m[1] * v.y + M.m[2] * v.z), M.m[3] * v.x + (M.m[4] * v.y + M.m[5] * v.z),
             M.m[6] * v.x + (M.m[7] * v.y + M.m[8] * v.z));
}

// Quaternion (w,x,y,z).
struct Q4 {
  double w, x, y, z;
};
// QuaternionBase::toRotationMatrix (lib/Eigen/src/Geometry/Quaternion.h:536-568).
SDM_HD M3<double> quat_rot(Q4 q) {
  const double tx = 2.0 * q.x, ty = 2.0 * q.y, tz = 2.0 * q.z;
  const double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
  const double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
  const double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
  M3<double> R;
  R.m[0] = 1.0 - (tyy + tzz); R.m[1] = txy - twz;         R.m[2] = txz + twy;
  R.m[3] = txy + twz;         R.m[4] = 1.0 - (txx + tzz); R.m[5] = tyz - twx;
  R.m[6] = txz - twy;         R.m[7] = tyz + twx;         R.m[8] = 1.0 - (txx + tyy);
  return R;
}
SDM_HD Q4 quat_conj(Q4 q) {
  Q4 r;
  r.w = q.w; r.x = -q.x; r.y = -q.y; r.z = -q.z;
  return r;
}
// Quaternion::normalize over the coefficient order (x,y,z,w).
SDM_HD Q4 quat_normalized(Q4 q) {
  double z = (q.x * q.x + q.y * q.y) + (q.z * q.z + q.w * q.w);
  if (z > 0.0) {
    double s = sqrt(z);
    q.w /= s; q.x /= s; q.y /= s; q.z /= s;
  }
  return q;
}
// QuaternionBase::_transformVector (Quaternion.h:476-485).
SDM_HD V3<double> quat_apply(Q4 q, V3<double> v) {
  V3<double> vec = mk3(q.x, q.y, q.z);
  V3<double> uv = cross(vec, v);
  uv = uv + uv;
  V3<double> c = cross(vec, uv);
  return mk3((v.x + q.w * uv.x) + c.x, (v.y + q.w * uv.y) + c.y, (v.z + q.w * uv.z) + c.z);
}
SDM_HD Q4 quat_mul(Q4 a, Q4 b) {
  Q4 o;
  o.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  o.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  o.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
  o.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
  return o;
}

// Transcendentals by precision.  Double pow, log, acos, sin, cos, atan2 and atan are the
// correctly rounded versions of crmath.h (the values glibc returns to the reference for all
// but ~0.1 % of arguments), log10 glibc's own formula on that log; float maps to the device's
// float functions.
SDM_HD double xpow(double a, double b) { return crm::cr_pow(a, b); }
SDM_HD float xpow(float a, float b) { return powf(a, b); }
SDM_HD double xsin(double a) { return crm::cr_sin(a); }
SDM_HD float xsin(float a) { return sinf(a); }
SDM_HD double xcos(double a) { return crm::cr_cos(a); }
SDM_HD float xcos(float a) { return cosf(a); }
SDM_HD double xatan2(double a, double b) { return crm::cr_atan2(a, b); }
SDM_HD float xatan2(float a, float b) { return atan2f(a, b); }
SDM_HD double xatan(double a) { return crm::cr_atan(a); }
SDM_HD double xfabs(double a) { return fabs(a); }
SDM_HD float xfabs(float a) { return fabsf(a); }
SDM_HD double xsqrt(double a) { return sqrt(a); }
SDM_HD float xsqrt(float a) { return sqrtf(a); }
SDM_HD double xlog(double a) { return crm::cr_log(a); }
SDM_HD float xlog(float a) { return logf(a); }
SDM_HD double xlog10(double a) { return crm::cr_log10(a); }
SDM_HD float xlog10(float a) { return log10f(a); }
SDM_HD double xacos(double a) { return crm::cr_acos(a); }
SDM_HD float xacos(float a) { return acosf(a); }

template <typename T, typename S>
SDM_HD V3<T> cast3(V3<S> a) { return mk3(T(a.x), T(a.y), T(a.z)); }
template <typename T, typename S>
SDM_HD M3<T> cast9(const M3<S>& a) {
  M3<T> r;
  for (int i = 0; i < 9; i++) r.m[i] = T(a.m[i]);
  return r;
}
