// sdm_oracle.cpp -- CPU oracle of the SudoDEM3D superellipsoid step loop.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain scalar C++ and FP64, the
// reference path the MI355X library replaces.  It is the checker that tests/ and the
// CPU-baseline leg of bench.py call; nothing in the product links or loads it.
//
// Every function cites the reference source it follows (paths relative to
// SudoDEM3D/).  The arithmetic is written in the evaluation order the reference binary
// uses: Eigen 3.3.5 is built with EIGEN_DONT_VECTORIZE (CMakeLists.txt:104), so a sum
// over 3 coefficients is x0+(x1+x2) and over 4 is (x0+x1)+(x2+x3)
// (lib/Eigen/src/Core/Redux.h:91-105), and small products are coefficient-wise dot
// products (lib/Eigen/src/Core/ProductEvaluators.h:550-553).  Reference quirks are kept
// on purpose; each one is marked QUIRK with the line that has it.
//
// Deviations (all documented, none changes a result the tests compare):
//  * neldermead_simplex leaves `nfunc` uninitialised (Superquadrics_Ig2.cpp:1713); it is
//    0 here, so the NMAX=5000 guard only fires after 5000 evaluations.
//  * The initial std::sort of the collider (InsertionSortCollider.cpp:339) can be replaced
//    by std::stable_sort (sdmcpu_set_init_sort) to pin the order of equal bounds.
#include "sdm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

// Math of the hot path.  Built as-is, glibc's functions, as the reference binary calls them.
// Built with SDM_ORACLE_CRMATH (libsdm_oracle_cr.so), the correctly rounded functions of
// crmath.h, so that a comparison with the device's FP64 form isolates everything but the
// last bits glibc gets wrong (~0.1 % of arguments).  The body-construction arithmetic
// (gammln) stays on glibc in both builds, as in the device library's host code.
#ifdef SDM_ORACLE_CRMATH
#define CRM_FN inline
#include "crmath.h"
namespace om {
inline double pow(double a, double b) { return crm::cr_pow(a, b); }
inline double atan2(double a, double b) { return crm::cr_atan2(a, b); }
inline double atan(double a) { return crm::cr_atan(a); }
inline double sin(double a) { return crm::cr_sin(a); }
inline double cos(double a) { return crm::cr_cos(a); }
inline double log(double a) { return crm::cr_log(a); }
inline double log10(double a) { return crm::cr_log10(a); }
inline double acos(double a) { return crm::cr_acos(a); }
}  // namespace om
#else
namespace om {
using std::acos;
using std::atan;
using std::atan2;
using std::cos;
using std::log;
using std::log10;
using std::pow;
using std::sin;
}  // namespace om
#endif

namespace {

const double NaN = std::numeric_limits<double>::quiet_NaN();
const double Inf = std::numeric_limits<double>::infinity();

// ---------------------------------------------------------------------------------------
// Vector helpers in Eigen's operation order.
// ---------------------------------------------------------------------------------------
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }
inline double sqn3(const double* a) { return a[0] * a[0] + (a[1] * a[1] + a[2] * a[2]); }
inline double norm3(const double* a) { return std::sqrt(sqn3(a)); }
// MatrixBase::normalize (lib/Eigen/src/Core/Dot.h:145-151): unchanged when the norm is 0.
inline void normalize3(double* a) {
  double z = sqn3(a);
  if (z > 0.0) {
    double s = std::sqrt(z);
    a[0] /= s; a[1] /= s; a[2] /= s;
  }
}
inline void cross3(const double* a, const double* b, double* o) {
  double x = a[1] * b[2] - a[2] * b[1];
  double y = a[2] * b[0] - a[0] * b[2];
  double z = a[0] * b[1] - a[1] * b[0];
  o[0] = x; o[1] = y; o[2] = z;
}
inline void matvec(const double* M, const double* v, double* o) {
  double x = M[0] * v[0] + (M[1] * v[1] + M[2] * v[2]);
  double y = M[3] * v[0] + (M[4] * v[1] + M[5] * v[2]);
  double z = M[6] * v[0] + (M[7] * v[1] + M[8] * v[2]);
  o[0] = x; o[1] = y; o[2] = z;
}
inline void set3(double* o, double x, double y, double z) { o[0] = x; o[1] = y; o[2] = z; }
inline void copy3(double* o, const double* a) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
// Mathr::Sign (lib/base/Math.hpp:186): Sign(0) = 0.
inline double Sign(double f) { if (f < 0) return -1; if (f > 0) return 1; return 0; }

// Quaternions are stored (w,x,y,z); Eigen stores coefficients (x,y,z,w).
// QuaternionBase::toRotationMatrix (lib/Eigen/src/Geometry/Quaternion.h:536-568).
void quat_to_rot(const double* q, double* R) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double tx = 2.0 * x, ty = 2.0 * y, tz = 2.0 * z;
  const double twx = tx * w, twy = ty * w, twz = tz * w;
  const double txx = tx * x, txy = ty * x, txz = tz * x;
  const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
  R[0] = 1.0 - (tyy + tzz); R[1] = txy - twz;         R[2] = txz + twy;
  R[3] = txy + twz;         R[4] = 1.0 - (txx + tzz); R[5] = tyz - twx;
  R[6] = txz - twy;         R[7] = tyz + twx;         R[8] = 1.0 - (txx + tyy);
}
void quat_conj(const double* q, double* o) { o[0] = q[0]; o[1] = -q[1]; o[2] = -q[2]; o[3] = -q[3]; }
// Quaternion::normalize = coeffs().normalize(); 4-term redux over (x,y,z,w).
void quat_normalize(double* q) {
  double z = (q[1] * q[1] + q[2] * q[2]) + (q[3] * q[3] + q[0] * q[0]);
  if (z > 0.0) {
    double s = std::sqrt(z);
    q[0] /= s; q[1] /= s; q[2] /= s; q[3] /= s;
  }
}
// QuaternionBase::_transformVector (lib/Eigen/src/Geometry/Quaternion.h:476-485).
void quat_rotate(const double* q, const double* v, double* o) {
  double vec[3] = {q[1], q[2], q[3]};
  double uv[3];
  cross3(vec, v, uv);
  uv[0] += uv[0]; uv[1] += uv[1]; uv[2] += uv[2];
  double c[3];
  cross3(vec, uv, c);
  for (int i = 0; i < 3; i++) o[i] = (v[i] + q[0] * uv[i]) + c[i];
}
// Quaternion product (lib/Eigen/src/Geometry/Quaternion.h, quat_product).
void quat_mul(const double* a, const double* b, double* o) {
  double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  double y = a[0] * b[2] + a[2] * b[0] + a[3] * b[1] - a[1] * b[3];
  double z = a[0] * b[3] + a[3] * b[0] + a[1] * b[2] - a[2] * b[1];
  o[0] = w; o[1] = x; o[2] = y; o[3] = z;
}

// Mathr::gammln / Mathr::beta (lib/base/Math.hpp:193-212), Numerical Recipes 3 Lanczos form.
double gammln(const double xx) {
  static const double cof[14] = {57.1562356658629235, -59.5979603554754912, 14.1360979747417471,
                                 -0.491913816097620199, .339946499848118887e-4, .465236289270485756e-4,
                                 -.983744753048795646e-4, .158088703224912494e-3, -.210264441724104883e-3,
                                 .217439618115212643e-3, -.164318106536763890e-3, .844182239838527433e-4,
                                 -.261908384015814087e-4, .368991826595316234e-5};
  double x, tmp, y, ser;
  y = x = xx;
  tmp = x + 5.24218750000000000;
  tmp = (x + 0.5) * std::log(tmp) - tmp;
  ser = 0.999999999999997092;
  for (int j = 0; j < 14; j++) ser += cof[j] / ++y;
  return tmp + std::log(2.5066282746310005 * ser / x);
}
double beta_fn(const double z, const double w) { return std::exp(gammln(z) + gammln(w) - gammln(z + w)); }

// ---------------------------------------------------------------------------------------
// Superquadrics shape (pkg/dem/Superquadrics.cpp, Superquadrics.hpp:48-166)
// ---------------------------------------------------------------------------------------
struct SQ {
  double rx, ry, rz, eps1, eps2, r_max;
  int isSphere;
  const double* Rl;  // rot_mat2local (cached, Superquadrics.hpp:101)
  const double* Rg;  // rot_mat2global
};

// Superquadrics::Initial (Superquadrics.cpp:47-93).
void sq_initial(double rx, double ry, double rz, double eps1, double eps2, double* r_max, double* volume,
                double* inertia) {
  double r = (rx > ry) ? rx : ry;
  r = r > rz ? r : rz;
  if ((eps1 < 1.0) || (eps2 < 1.0)) r *= std::pow(2.0, 0.5);
  *r_max = r;
  double a = rx * ry * rz * eps1 * eps2;
  double beta1 = beta_fn(1.5 * eps1, 0.5 * eps1) * beta_fn(0.5 * eps2, 2. * eps2 + 1);
  double beta2 = beta_fn(0.5 * eps1, 0.5 * eps1 + 1) * beta_fn(1.5 * eps2, eps2 + 1);
  *volume = 2. * a * beta_fn(eps1 * 0.5, eps1 * 0.5) * beta_fn(eps2 * 0.5 + 1, eps2);
  inertia[0] = 0.5 * a * ((ry * ry) * beta1 + 4. * (rz * rz) * beta2);
  inertia[1] = 0.5 * a * ((rx * rx) * beta1 + 4. * (rz * rz) * beta2);
  inertia[2] = 0.5 * a * ((rx * rx) + (ry * ry)) * beta1;
}

// Superquadrics::Normal2Phi (Superquadrics.cpp:269-307; the active branch is :284 and :288).
void normal2phi(const SQ& s, const double* n, double* phi) {
  phi[0] = om::atan2(Sign(n[1]) * om::pow((std::fabs(s.ry * n[1])), 1 / (2. - s.eps1)),
                      Sign(n[0]) * om::pow((std::fabs(s.rx * n[0])), 1 / (2. - s.eps1)));
  phi[1] = om::atan2(Sign(n[2]) * om::pow(std::fabs(s.rz * n[2] * om::pow(std::fabs(om::cos(phi[0])), 2. - s.eps1)),
                                            1 / (2. - s.eps2)),
                      om::pow(std::fabs(s.rx * n[0]), 1 / (2. - s.eps2)));
}

// Superquadrics::getSurface (Superquadrics.cpp:110-120), local coordinates.
void get_surface(const SQ& s, const double* phi, double* p) {
  p[0] = Sign(om::cos(phi[0])) * s.rx * om::pow(std::fabs(om::cos(phi[0])), s.eps1) *
         om::pow(std::fabs(om::cos(phi[1])), s.eps2);
  p[1] = Sign(om::sin(phi[0])) * s.ry * om::pow(std::fabs(om::sin(phi[0])), s.eps1) *
         om::pow(std::fabs(om::cos(phi[1])), s.eps2);
  p[2] = Sign(om::sin(phi[1])) * s.rz * om::pow(std::fabs(om::sin(phi[1])), s.eps2);
}

// Superquadrics::isInside (Superquadrics.cpp:103-108): p relative to the centre, global frame.
bool is_inside(const SQ& s, const double* pg) {
  double p[3];
  matvec(s.Rl, pg, p);
  return 1 > om::pow(om::pow(std::fabs(p[0] / s.rx), 2.0 / s.eps1) + om::pow(std::fabs(p[1] / s.ry), 2.0 / s.eps1),
                      s.eps1 / s.eps2) +
                 om::pow(std::fabs(p[2] / s.rz), 2.0 / s.eps2);
}

// Superquadrics::Curvatures (Superquadrics.cpp:131-255).  QUIRK: K ignores the M term
// (:183-188); R11, R22 use the coefficient choices of :172-178; R12 is never read.
void curvatures(const SQ& s, double* phi_in, double* k, double* dir_max, double* dir_min, int* max_flag) {
  *max_flag = 1;
  double phi[2] = {phi_in[0], phi_in[1]};
  if (std::fabs(phi[0]) < 1E-20) phi[0] = 1E-20;
  if (std::fabs(phi[1]) < 1E-20) phi[1] = 1E-20;
  const double rx = s.rx, ry = s.ry, rz = s.rz, eps1 = s.eps1, eps2 = s.eps2;
  double s1 = om::sin(phi[0]), s2 = om::sin(phi[1]);
  double c1 = om::cos(phi[0]), c2 = om::cos(phi[1]);
  double s_s1 = Sign(s1), s_s2 = Sign(s2), s_c1 = Sign(c1), s_c2 = Sign(c2);
  double abs_s1 = std::fabs(s1), abs_s2 = std::fabs(s2), abs_c1 = std::fabs(c1), abs_c2 = std::fabs(c2);
  double s10 = s_s1 * om::pow(abs_s1, eps1);
  double s20 = s_s2 * om::pow(abs_s2, eps2);
  double c10 = s_c1 * om::pow(abs_c1, eps1);
  double c20 = s_c2 * om::pow(abs_c2, eps2);
  double s12 = s_s1 * om::pow(abs_s1, eps1 - 2.0);
  double s22 = s_s2 * om::pow(abs_s2, eps2 - 2.0);
  double c12 = s_c1 * om::pow(abs_c1, eps1 - 2.0);
  double c22 = s_c2 * om::pow(abs_c2, eps2 - 2.0);
  double R1[3] = {-rx * eps1 * c12 * c1 * c20 * s1, ry * eps1 * c1 * c20 * s12 * s1, 0};
  double R2[3] = {-rx * eps2 * c10 * c22 * c2 * s2, -ry * eps2 * c22 * c2 * s10 * s2, rz * eps2 * c2 * s22 * s2};
  double R11[3] = {rx * eps1 * c12 * s20 * s1 * s1 * (eps1 - 1.0) - rx * eps1 * c10 * c20,
                   ry * eps1 * c1 * c1 * c20 * s12 * (eps1 - 1.0) - ry * eps1 * c20 * s10, 0};
  double R22[3] = {rx * eps2 * c10 * c22 * s2 * s2 * (eps2 - 1.0) - rx * eps2 * c10 * c20,
                   ry * eps2 * s10 * s1 * s1 * (eps2 - 1.0) - ry * eps2 * c20 * s10,
                   rz * eps2 * c2 * c2 * s22 * (eps2 - 1.0) - rz * eps2 * s20};
  double RR[3];
  cross3(R1, R2, RR);
  double nrm = norm3(RR);
  double n[3] = {RR[0] / nrm, RR[1] / nrm, RR[2] / nrm};
  double L = dot3(R11, n);
  double N = dot3(R22, n);
  double E = dot3(R1, R1);
  double F = dot3(R1, R2);
  double G = dot3(R2, R2);
  double K = L * N / (E * G - F * F);
  double H = 0.5 * (E * N + G * L) / (E * G - F * F);
  double tmp0 = std::sqrt(std::fabs(H * H - K));
  double k1 = H + tmp0;
  double k2 = H - tmp0;
  double dir1[3], dir2[3];
  for (int i = 0; i < 3; i++) dir1[i] = R1[i] * (N - k1 * G) + R2[i] * k1 * F;
  if (norm3(dir1) == 0) {
    *max_flag = 0;
  } else {
    normalize3(dir1);
    matvec(s.Rg, dir1, dir_max);
  }
  for (int i = 0; i < 3; i++) dir2[i] = R1[i] * (N - k2 * G) + R2[i] * k2 * F;
  normalize3(dir2);
  matvec(s.Rg, dir2, dir_min);
  k[0] = k1;
  k[1] = k2;
}

// ---------------------------------------------------------------------------------------
// Bodies, interactions, collider bounds.
// ---------------------------------------------------------------------------------------
const unsigned DOF_ALL = 63;
inline unsigned axisDOF(int axis, bool rot) { return 1u << (axis + (rot ? 3 : 0)); }

struct Body {
  // state (core/State.hpp)
  double pos[3], ori[4], vel[3], angVel[3], angMom[3], mass, inertia[3];
  int32_t specimen;  // batching attribute of the device library; the oracle runs one specimen
  int32_t blockedDOFs, isDamped;
  // shape
  int32_t kind;  // SDM_BODY_SUPERQUADRIC or SDM_BODY_WALL
  double rxyz[3], eps[2], r_max;
  int32_t isSphere, wallAxis, wallSense, material;
  double rot_local[9], rot_global[9];
  // bound (core/Bound.hpp:24-31)
  int32_t hasBound;
  double bmin[3], bmax[3], refPos[3], sweepLength;
  int32_t lastUpdateIter;
  // forces (ForceContainer, read back for checks)
  double force[3], torque[3];
};

struct Pair {
  int32_t id1, id2;
  int32_t hasGeom, hasPhys, functorSet, geomExists;
  // SuperquadricsGeom (Superquadrics.hpp:186-201)
  double depth, contactAngle[2], contactPoint[3], shearInc[3], relVn[3], relVs[3], normal[3], twist[3], point1[3],
      point2[3], orth[3];
  int32_t isShearNew;
  // SuperquadricsGeom2 (Superquadrics.hpp:216-224)
  double curv1[2], curv2[2], curvAngle, alpha, K, E, curvSum;
  int32_t geomIsSphere;
  // SuperquadricsPhys / HertzMindlinPhys (Superquadrics.hpp:303-337, FrictPhys, NormShearPhys)
  double kn, ks, tanPhi, betan, betas, normalForce[3], shearForce[3], nuab, Gab, mu1, mu2;
  // per-contact scatter of the last step: force on id1, torques on id1 and id2
  double F[3], T1[3], T2[3];
  int32_t contributed;
  // narrowphase diagnostics: evaluations and solver stop code of the last step
  int32_t nEval, stopCode;
  int64_t iterMadeReal, iterBorn;
  int32_t cellDist[3];  // Interaction::cellDist (core/Interaction.hpp): period of id2 relative to id1
};

void geom_reset(Pair& p) {
  p.depth = 0;
  p.contactAngle[0] = p.contactAngle[1] = 0;
  for (int i = 0; i < 3; i++)
    p.contactPoint[i] = p.shearInc[i] = p.relVn[i] = p.relVs[i] = p.normal[i] = p.twist[i] = p.point1[i] =
        p.point2[i] = p.orth[i] = 0;
  p.isShearNew = 1;
  p.curv1[0] = p.curv1[1] = p.curv2[0] = p.curv2[1] = 0;
  p.curvAngle = p.alpha = p.K = p.E = p.curvSum = 0;
  p.geomIsSphere = 0;
}
void phys_reset(Pair& p) {
  p.kn = p.ks = 0;
  p.tanPhi = NaN;
  p.betan = p.betas = 0;
  for (int i = 0; i < 3; i++) p.normalForce[i] = p.shearForce[i] = 0;
  p.nuab = p.Gab = p.mu1 = p.mu2 = 0;
}
Pair new_pair(int id1, int id2) {
  Pair p;
  std::memset(&p, 0, sizeof(p));
  p.id1 = id1;
  p.id2 = id2;
  p.geomExists = 1;
  geom_reset(p);
  phys_reset(p);
  p.iterMadeReal = -1;
  return p;
}

struct Bounds {  // InsertionSortCollider::Bounds (InsertionSortCollider.hpp:90-104)
  double coord;
  int32_t id;
  int32_t isMin, hasBB;
  int32_t period;  // periodic cell coordinate of the bound
  bool operator<(const Bounds& b) const {
    if (id == b.id && coord == b.coord) return isMin;
    return coord < b.coord;
  }
  bool operator>(const Bounds& b) const {
    if (id == b.id && coord == b.coord) return !isMin;
    return coord > b.coord;
  }
};

inline uint64_t pair_key(int a, int b) {
  if (a > b) std::swap(a, b);
  return (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
}

}  // namespace

struct sdm_ctx {
  sdm_params prm;
  std::vector<Body> bodies;
  std::vector<sdm_material> mats;
  std::vector<Pair> intrs;                  // InteractionContainer::linIntrs
  std::unordered_set<uint64_t> found_set;  // Body::intrs maps, as a set of unordered pairs
  bool found_dirty = false;
  // InsertionSortCollider state
  std::vector<Bounds> BB[3];
  long loIdx[3] = {0, 0, 0};  // VecBounds::loIdx: the element before which each ring wraps
  bool periodicState = false;  // InsertionSortCollider::periodic
  bool periErr = false;        // spatialOverlapPeri met a body larger than half the cell
  std::vector<double> minima, maxima;
  int32_t strideActive = 0;
  double fastestBodyMaxDist = -1;
  int32_t numReinit = 0;
  int32_t initSortStable = 0;
  // NewtonIntegrator state
  double maxVelocitySq = NaN;
  // CompressionEngine (cubic specimen, SOLE_GAIN + NEW_SCHEME)
  struct Servo {
    bool on = false, firstRun = true, quiet = false;
    sdm_servo p;
    int32_t iterate_num = 0, solve_num = 0, rampNum = 0;
    double width0 = 0, depth0 = 0, height0 = 0, initVol = 0;
    double width = 0, depth = 0, height = 0, x_area = 0, y_area = 0, z_area = 0;
    double ws[6] = {0, 0, 0, 0, 0, 0};  // wsxx_left, wsxx_right, wsyy_front, wsyy_back, wszz_bottom, wszz_top
    double wsxx = 0, wsyy = 0, wszz = 0;
    double stiff[6] = {0, 0, 0, 0, 0, 0}, gain[6] = {0, 0, 0, 0, 0, 0};
    double mean = 0, dev = 0, unbf = 1;
    int64_t completedIter = -1;
    bool completedNow = false;
  } servo;
  bool servoEverSet = false;  // a servo state exists (set once; kept while switched off)
  // Scene
  int64_t iter = 0;
  double time = 0;
  int nthreads = 1;
  std::vector<double> thrForce;  // [nthreads][N][6]
  sdm_counters cnt;
  std::string err;
};

namespace {

std::string g_err;

void build_found(sdm_ctx* c) {
  c->found_set.clear();
  c->found_set.reserve(c->intrs.size() * 2 + 16);
  for (auto& p : c->intrs) c->found_set.insert(pair_key(p.id1, p.id2));
  c->found_dirty = false;
}

SQ make_sq(const Body& b) {
  SQ s;
  s.rx = b.rxyz[0]; s.ry = b.rxyz[1]; s.rz = b.rxyz[2];
  s.eps1 = b.eps[0]; s.eps2 = b.eps[1];
  s.r_max = b.r_max;
  s.isSphere = b.isSphere;
  s.Rl = b.rot_local;
  s.Rg = b.rot_global;
  return s;
}

// ---------------------------------------------------------------------------------------
// Levenberg-Marquardt narrowphase (Superquadrics_Ig2.cpp:621-687, 1032-1354, 1880-2317)
// ---------------------------------------------------------------------------------------
struct LMOut {
  double p1[3], p2[3], center[3];
};

// Disfunction, active overload (Superquadrics_Ig2.cpp:621-687).  QUIRK: stop_flag is only
// ever set, never cleared (:631 is commented out), so it is sticky across evaluations.
void disfunction(const double* para, double* center, const SQ& A, const SQ& B, const double* position1,
                 const double* position2, double* p1, double* p2, bool& stop_flag, int& nev) {
  nev++;
  double contact[3] = {om::cos(para[0]) * om::cos(para[1]), om::sin(para[0]) * om::cos(para[1]),
                       om::sin(para[1])};
  double contact1[3], contact2[3], phi[2], t[3];
  matvec(A.Rl, contact, contact1);
  normal2phi(A, contact1, phi);
  get_surface(A, phi, t);
  matvec(A.Rg, t, p1);
  for (int i = 0; i < 3; i++) p1[i] = p1[i] + position1[i];
  matvec(B.Rl, contact, contact2);
  double mc2[3] = {(-1) * contact2[0], (-1) * contact2[1], (-1) * contact2[2]};
  normal2phi(B, mc2, phi);
  get_surface(B, phi, t);
  matvec(B.Rg, t, p2);
  for (int i = 0; i < 3; i++) p2[i] = p2[i] + position2[i];
  double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  for (int i = 0; i < 3; i++) center[i] = 0.5 * (p1[i] + p2[i]);
  if (dot3(contact, d) > 0) stop_flag = true;
}

// JacfunctionApp (Superquadrics_Ig2.cpp:1032-1054): forward differences with step paraDelta.
// QUIRK: p1/p2/center are left at the last perturbed evaluation.
void jacfunction_app(const double* para, const double* paraDelta, double* center, const SQ& A, const SQ& B,
                     const double* position1, const double* position2, double* p1, double* p2, bool& stop_flag,
                     double* J /*3x2 row-major*/, int& nev) {
  double p_f0[3], p_f1[3], p_f2[3], para0[2];
  disfunction(para, center, A, B, position1, position2, p1, p2, stop_flag, nev);
  for (int i = 0; i < 3; i++) p_f0[i] = p2[i] - p1[i];
  para0[0] = para[0] + paraDelta[0];
  para0[1] = para[1];
  disfunction(para0, center, A, B, position1, position2, p1, p2, stop_flag, nev);
  for (int i = 0; i < 3; i++) p_f1[i] = p2[i] - p1[i];
  para0[0] = para[0];
  para0[1] = para[1] + paraDelta[1];
  disfunction(para0, center, A, B, position1, position2, p1, p2, stop_flag, nev);
  for (int i = 0; i < 3; i++) p_f2[i] = p2[i] - p1[i];
  for (int i = 0; i < 3; i++) J[2 * i + 0] = (p_f1[i] - p_f0[i]) / paraDelta[0];
  for (int i = 0; i < 3; i++) J[2 * i + 1] = (p_f2[i] - p_f0[i]) / paraDelta[1];
}

// lmder2 (Superquadrics_Ig2.cpp:1056-1354).  Returns info[6] (stop code).
int lmder2(double* p, int itmax, const double* opts, const SQ& A, const SQ& B, const double* position1,
           const double* position2, double* p1, double* p2, double* center, bool& touching, double* contact,
           int& nev) {
  int k;
  touching = true;
  double p_f[3] = {0, 0, 0}, pDp_f[3] = {0, 0, 0};
  double jacTf[2], diag_jacTjac[2] = {0, 0};
  double jac[6];
  double jacTjac[4];
  double Dp[2] = {1e-7, 1e-7};
  double pDp[2];
  bool stop_flag = false;
  double mu = 0., tmp;
  double jacTf_inf = 0.;
  double p_L2, Dp_L2 = std::numeric_limits<double>::max(), dF, dL;
  double p_fL2 = 0., pDp_fL2 = 0.;
  double tau = opts[0], eps1 = opts[1], eps2 = opts[2], eps3 = opts[3];
  int nu = 2, nu2, stop = 0;

  disfunction(p, center, A, B, position1, position2, p1, p2, stop_flag, nev);
  for (int i = 0; i < 3; i++) p_f[i] = p2[i] - p1[i];
  p_fL2 = norm3(p_f);
  if (!std::isfinite(p_fL2)) stop = 7;

  for (k = 0; k < itmax && !stop; ++k) {
    if (p_fL2 <= eps3) { stop = 6; break; }
    if (stop_flag) { stop = 9; touching = false; break; }
    {
      contact[0] = om::cos(p[0]) * om::cos(p[1]);
      contact[1] = om::sin(p[0]) * om::cos(p[1]);
      contact[2] = om::sin(p[1]);
      double angle = dot3(p_f, contact) / (norm3(p_f) * norm3(contact));
      if (std::fabs(angle) > 0.95) { stop = 8; break; }
    }
    jacfunction_app(p, Dp, center, A, B, position1, position2, p1, p2, stop_flag, jac, nev);
    // J^T J and J^T (-f), coefficient-wise 3-term sums.
    jacTjac[0] = jac[0] * jac[0] + (jac[2] * jac[2] + jac[4] * jac[4]);
    jacTjac[1] = jac[0] * jac[1] + (jac[2] * jac[3] + jac[4] * jac[5]);
    jacTjac[2] = jac[1] * jac[0] + (jac[3] * jac[2] + jac[5] * jac[4]);
    jacTjac[3] = jac[1] * jac[1] + (jac[3] * jac[3] + jac[5] * jac[5]);
    double mf[3] = {-1.0 * p_f[0], -1.0 * p_f[1], -1.0 * p_f[2]};
    jacTf[0] = jac[0] * mf[0] + (jac[2] * mf[1] + jac[4] * mf[2]);
    jacTf[1] = jac[1] * mf[0] + (jac[3] * mf[1] + jac[5] * mf[2]);
    jacTf_inf = (std::fabs(jacTf[0]) > std::fabs(jacTf[1]) ? std::fabs(jacTf[0]) : std::fabs(jacTf[1]));
    p_L2 = std::sqrt(p[0] * p[0] + p[1] * p[1]);
    if (jacTf_inf <= eps1) { Dp_L2 = 0.0; stop = 1; break; }
    if (k == 0) {
      tmp = (jacTjac[0] > jacTjac[3] ? jacTjac[0] : jacTjac[3]);
      mu = tau * tmp;
    }
    diag_jacTjac[0] = jacTjac[0];
    diag_jacTjac[1] = jacTjac[3];
    while (1) {
      jacTjac[0] += mu;
      jacTjac[3] += mu;
      double n_a = jacTjac[0] * jacTjac[3] - jacTjac[1] * jacTjac[1];
      Dp[0] = (jacTf[0] * jacTjac[3] - jacTf[1] * jacTjac[1]) / n_a;
      Dp[1] = (jacTf[1] * jacTjac[0] - jacTf[0] * jacTjac[1]) / n_a;
      Dp_L2 = std::sqrt(Dp[0] * Dp[0] + Dp[1] * Dp[1]);
      if (Dp_L2 < 1e-5) { stop = 2; break; }
      pDp[0] = p[0] + Dp[0];
      pDp[1] = p[1] + Dp[1];
      if (Dp_L2 >= (p_L2 + eps2) / (1E-12 * 1E-12)) { stop = 4; break; }
      disfunction(pDp, center, A, B, position1, position2, p1, p2, stop_flag, nev);
      for (int i = 0; i < 3; i++) pDp_f[i] = p2[i] - p1[i];
      pDp_fL2 = norm3(pDp_f);
      if (!std::isfinite(pDp_fL2)) { stop = 7; break; }
      dL = Dp[0] * (mu * Dp[0] + jacTf[0]) + Dp[1] * (mu * Dp[1] + jacTf[1]);
      dF = p_fL2 - pDp_fL2;
      tmp = dF / dL;
      if (tmp > 0.0) {
        tmp = 1.0 - om::pow(2 * tmp - 1., 3.);
        mu = mu * ((tmp >= 0.3333333333333334) ? tmp : 0.3333333333333334);
        nu = 2;
        p[0] = pDp[0]; p[1] = pDp[1];
        copy3(p_f, pDp_f);
        p_fL2 = pDp_fL2;
        break;
      }
      mu *= nu;
      nu2 = int(unsigned(nu) << 1);  // 2*nu, wraps like the reference's int shift
      if (nu2 <= nu) { stop = 5; break; }
      nu = nu2;
      jacTjac[0] = diag_jacTjac[0];
      jacTjac[3] = diag_jacTjac[1];
    }
  }
  if (k >= itmax) stop = 3;
  return stop;
}

// ---------------------------------------------------------------------------------------
// Nelder-Mead narrowphase (Superquadrics_Ig2.cpp:295-319, 1624-1875, 2322-2517)
// ---------------------------------------------------------------------------------------
// CRotationMat (Superquadrics_Ig2.cpp:295-319).
void crotation_mat(const double* position1, const double* position2, double* R) {
  double e1[3] = {position2[0] - position1[0], position2[1] - position1[1], position2[2] - position1[2]};
  double x1[3] = {1., 0, 0};
  double cos_phi = dot3(x1, e1) / (norm3(x1) * norm3(e1));
  double cos_halfphi = std::sqrt(0.5 * (1 + cos_phi));
  double sin_halfphi = std::sqrt(0.5 * (1 - cos_phi));
  double q[4];
  if (std::fabs(cos_phi) > 1.0 - 1e-10) {
    q[0] = cos_halfphi; q[1] = 0; q[2] = 0; q[3] = sin_halfphi;
  } else {
    double n1[3];
    cross3(x1, e1, n1);
    normalize3(n1);
    for (int i = 0; i < 3; i++) n1[i] *= sin_halfphi;
    q[0] = cos_halfphi; q[1] = n1[0]; q[2] = n1[1]; q[3] = n1[2];
  }
  quat_to_rot(q, R);
}

// Disfun (Superquadrics_Ig2.cpp:1624-1678).
double disfun(const double* para, const SQ& A, const SQ& B, const double* position1, const double* position2,
              double* p1, double* p2, const double* gC, bool& touching, int& nev) {
  nev++;
  double c0[3] = {om::cos(para[0]) * om::cos(para[1]), om::sin(para[0]) * om::cos(para[1]), om::sin(para[1])};
  double contact[3], contact1[3], contact2[3], phi[2], t[3];
  matvec(gC, c0, contact);
  matvec(A.Rl, contact, contact1);
  normal2phi(A, contact1, phi);
  get_surface(A, phi, t);
  matvec(A.Rg, t, p1);
  for (int i = 0; i < 3; i++) p1[i] = p1[i] + position1[i];
  matvec(B.Rl, contact, contact2);
  double mc2[3] = {(-1) * contact2[0], (-1) * contact2[1], (-1) * contact2[2]};
  normal2phi(B, mc2, phi);
  get_surface(B, phi, t);
  matvec(B.Rg, t, p2);
  for (int i = 0; i < 3; i++) p2[i] = p2[i] + position2[i];
  double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  if (dot3(contact, d) > 0) touching = false;
  return norm3(d);
}

// amotry (Superquadrics_Ig2.cpp:1681-1699).
double amotry(double (&p)[3][2], double* y, double* psum, const int ihi, const double fac, const SQ& A, const SQ& B,
              const double* position1, const double* position2, double* pt1, double* pt2, const double* gC,
              bool& touching, int& nev) {
  double ptry[2];
  double fac1 = (1.0 - fac) / 2.0;
  double fac2 = fac1 - fac;
  ptry[0] = psum[0] * fac1 - p[ihi][0] * fac2;
  ptry[1] = psum[1] * fac1 - p[ihi][1] * fac2;
  double ytry = disfun(ptry, A, B, position1, position2, pt1, pt2, gC, touching, nev);
  if (ytry < y[ihi]) {
    y[ihi] = ytry;
    psum[0] += ptry[0] - p[ihi][0];
    psum[1] += ptry[1] - p[ihi][1];
    p[ihi][0] = ptry[0];
    p[ihi][1] = ptry[1];
  }
  return ytry;
}

// neldermead_simplex (Superquadrics_Ig2.cpp:1701-1875).  Returns the best vertex in point,
// info[0]=nfunc, info[1]=distance.  nfunc starts at 0 (see file header).
void neldermead_simplex(double* point, const double* dels, double* info, const SQ& A, const SQ& B,
                        const double* position1, const double* position2, double* pt1, double* pt2, const double* gC,
                        bool& touching, int& nev) {
  const double ftol = 1e-3;
  int nfunc = 0;
  double fmin = 0;
  const int NMAX = 5000;
  const double TINY = 1.0e-10;
  int ihi, ilo, inhi;
  double psum[2], y[3], p[3][2];
  for (int i = 0; i < 3; i++) { p[i][0] = point[0]; p[i][1] = point[1]; }
  p[1][0] += dels[0];
  p[2][1] += dels[1];
  double y_tmp = 1000000.0;
  for (int i = 0; i < 3; i++) {
    y[i] = disfun(p[i], A, B, position1, position2, pt1, pt2, gC, touching, nev);
    nfunc++;
    if (y[i] < y_tmp) y_tmp = y[i];
    if (!touching) {
      info[0] = nfunc; info[1] = y[0];
      point[0] = p[0][0]; point[1] = p[0][1];
      return;
    }
  }
  psum[0] = p[0][0] + p[1][0] + p[2][0];
  psum[1] = p[0][1] + p[1][1] + p[2][1];
  for (;;) {
    ilo = 0;
    if (!touching) {
      info[0] = nfunc; info[1] = fmin;
      point[0] = p[0][0]; point[1] = p[0][1];
      return;
    }
    ihi = y[0] > y[1] ? (inhi = 1, 0) : (inhi = 0, 1);
    for (int i = 0; i < 3; i++) {
      if (y[i] <= y[ilo]) ilo = i;
      if (y[i] > y[ihi]) {
        inhi = ihi;
        ihi = i;
      } else if (y[i] > y[inhi] && i != ihi)
        inhi = i;
    }
    double rtol = 2.0 * std::fabs(y[ihi] - y[ilo]) / (std::fabs(y[ihi]) + std::fabs(y[ilo]) + TINY);
    if (rtol < ftol) {
      double t = y[0]; y[0] = y[ilo]; y[ilo] = t;
      double t0 = p[ilo][0], t1 = p[ilo][1];
      p[ilo][0] = p[0][0]; p[ilo][1] = p[0][1];
      p[0][0] = t0; p[0][1] = t1;
      fmin = y[0];
      info[0] = nfunc; info[1] = fmin;
      point[0] = p[0][0]; point[1] = p[0][1];
      return;
    }
    if (nfunc >= NMAX) {
      info[0] = nfunc; info[1] = y[0];
      point[0] = p[0][0]; point[1] = p[0][1];
      return;
    }
    double ytry = amotry(p, y, psum, ihi, -1.0, A, B, position1, position2, pt1, pt2, gC, touching, nev);
    nfunc++;
    if (!touching) {
      info[0] = nfunc; info[1] = y[0];
      point[0] = p[0][0]; point[1] = p[0][1];
      return;
    }
    if (ytry <= y[ilo]) {
      ytry = amotry(p, y, psum, ihi, 2.0, A, B, position1, position2, pt1, pt2, gC, touching, nev);
      nfunc++;
      if (!touching) {
        info[0] = nfunc; info[1] = y[0];
        point[0] = p[0][0]; point[1] = p[0][1];
        return;
      }
    } else if (ytry >= y[inhi]) {
      double ysave = y[ihi];
      ytry = amotry(p, y, psum, ihi, 0.5, A, B, position1, position2, pt1, pt2, gC, touching, nev);
      nfunc++;
      if (!touching) {
        info[0] = nfunc; info[1] = y[0];
        point[0] = p[0][0]; point[1] = p[0][1];
        return;
      }
      if (ytry >= ysave) {
        for (int i = 0; i < 3; i++) {
          if (i != ilo) {
            psum[0] = 0.5 * (p[i][0] + p[ilo][0]);
            psum[1] = 0.5 * (p[i][1] + p[ilo][1]);
            p[i][0] = psum[0];
            p[i][1] = psum[1];
            y[i] = disfun(psum, A, B, position1, position2, pt1, pt2, gC, touching, nev);
            nfunc++;
            if (!touching) {
              info[0] = nfunc; info[1] = y[0];
              point[0] = p[0][0]; point[1] = p[0][1];
              return;
            }
          }
        }
        psum[0] = p[0][0] + p[1][0] + p[2][0];
        psum[1] = p[0][1] + p[1][1] + p[2][1];
      }
    }
  }
}

// SuperquadricsGeom2::HertzMindlin (Superquadrics.cpp:974-1026).  out = {curvSum, alpha, K, E}
void hertz_mindlin(const double* curvature1, const double* curvature2, double theta, double* out) {
  double rho11 = std::fabs(curvature1[0]);
  double rho12 = std::fabs(curvature1[1]);
  double rho21 = std::fabs(curvature2[0]);
  double rho22 = std::fabs(curvature2[1]);
  if (rho11 * rho12 < 1E-18) rho11 = rho12 = 0.0;
  if (rho21 * rho22 < 1E-18) rho21 = rho22 = 0.0;
  double sum_rho = rho11 + rho12 + rho21 + rho22;
  double a1 = rho11 - rho12, a2 = rho21 - rho22;
  double F_rho = -std::sqrt(a1 * a1 + a2 * a2 + 2 * (std::fabs(rho11 - rho12)) * std::fabs(rho21 - rho22) *
                                                  om::cos(2 * theta));
  double A = (sum_rho + F_rho) / 4.;
  double B = (sum_rho - F_rho) / 4.;
  double curvatures_sum = A + B;
  double lg = om::log10(B / A);
  double x_squ = lg * lg;
  double gamma1 = 2 * (1 + x_squ * (0.40227436 + x_squ * (3.7491752e-2 + x_squ * (7.4855761e-4 + x_squ * 2.1667028e-6))));
  double gamma2 = 3 * (1 + x_squ * (0.42678878 + x_squ * (4.2605401e-2 + x_squ * (9.0786922e-4 + x_squ * 2.7868927e-6))));
  double alpha = om::pow((B / A), gamma1 / gamma2);
  double m1 = 1 / (alpha * alpha);
  double K = (1.3862944 + m1 * (0.1119723 + 0.0725296 * m1)) - om::log(m1) * (0.5 + m1 * (0.1213478 + 0.0288729 * m1));
  double E = (1 + m1 * (0.4630151 + 0.1077812 * m1)) - om::log(m1) * (0.2452727 + 0.0412496 * m1) * m1;
  out[0] = curvatures_sum;
  out[1] = alpha;
  out[2] = K;
  out[3] = E;
}

// SuperquadricsGeom::precompute (Superquadrics.cpp:1030-1050).  `normal` holds the previous
// normal on entry and the current one on exit.
void precompute(Pair& g, const Body& b1, const Body& b2, double dt, const double* currentNormal, bool isNew,
                const double* shift2) {
  if (!isNew) {
    cross3(g.normal, currentNormal, g.orth);
    double w[3] = {b1.angVel[0] + b2.angVel[0], b1.angVel[1] + b2.angVel[1], b1.angVel[2] + b2.angVel[2]};
    double angle = dt * 0.5 * dot3(currentNormal, w);
    for (int i = 0; i < 3; i++) g.twist[i] = angle * currentNormal[i];
  } else {
    set3(g.twist, 0, 0, 0);
    set3(g.orth, 0, 0, 0);
  }
  copy3(g.normal, currentNormal);
  double c1x[3], c2x[3], t1[3], t2[3], rv[3];
  for (int i = 0; i < 3; i++) c1x[i] = g.contactPoint[i] - b1.pos[i];
  for (int i = 0; i < 3; i++) c2x[i] = g.contactPoint[i] - b2.pos[i] - shift2[i];
  cross3(b2.angVel, c2x, t2);
  cross3(b1.angVel, c1x, t1);
  for (int i = 0; i < 3; i++) rv[i] = (b2.vel[i] + t2[i]) - (b1.vel[i] + t1[i]);
  double vn = dot3(g.normal, rv);
  for (int i = 0; i < 3; i++) g.relVn[i] = vn * g.normal[i];
  for (int i = 0; i < 3; i++) g.relVs[i] = rv[i] - g.relVn[i];
  for (int i = 0; i < 3; i++) g.shearInc[i] = g.relVs[i] * dt;
}

// SuperquadricsGeom::rotate (Superquadrics.cpp:1053-1080).
void rotate_shear(const Pair& g, double* fs) {
  double t[3];
  cross3(fs, g.orth, t);
  for (int i = 0; i < 3; i++) fs[i] -= t[i];
  cross3(fs, g.twist, t);
  for (int i = 0; i < 3; i++) fs[i] -= t[i];
  double d = dot3(g.normal, fs);
  for (int i = 0; i < 3; i++) fs[i] -= d * g.normal[i];
}

const double zero3[3] = {0, 0, 0};

// shift2 = Cell::hSize * cellDist (InteractionLoop.cpp:103); the cell is axis-aligned, so
// component i is size_i * cellDist_i.  Zero outside a periodic scene (:101).
void pair_shift2(const sdm_ctx* c, const Pair& I, double* s) {
  for (int i = 0; i < 3; i++) s[i] = c->prm.periodic ? c->prm.cell[i] * (double)I.cellDist[i] : 0.0;
}

// Ig2_Superquadrics_Superquadrics_SuperquadricsGeom::go (Superquadrics_Ig2.cpp:1880-2317).
void ig2_lm(sdm_ctx* c, Pair& I, const Body& b1, const Body& b2) {
  double shift2[3];
  pair_shift2(c, I, shift2);
  const double* position1 = b1.pos;
  const double position2[3] = {b2.pos[0] + shift2[0], b2.pos[1] + shift2[1], b2.pos[2] + shift2[2]};
  SQ A = make_sq(b1), B = make_sq(b2);
  bool isNew = !I.hasGeom;
  double para[2] = {1e-5, 1e-5};
  int nev = 0;
  I.stopCode = 0;
  if (isNew) {
    geom_reset(I);
    I.hasGeom = 1;
    I.contactPoint[0] = I.contactPoint[1] = I.contactPoint[2] = 0;
    I.isShearNew = 1;
    para[0] = I.contactAngle[0];
    para[1] = I.contactAngle[1];
    double d0[3] = {position2[0] - position1[0], position2[1] - position1[1], position2[2] - position1[2]};
    para[0] = om::atan2(d0[1], d0[0]);
    para[1] = om::atan(d0[2] / std::sqrt(d0[0] * d0[0] + d0[1] * d0[1]));
    para[0] += 1e-5;
    para[1] += 1e-5;
  } else {
    I.isShearNew = I.depth <= 0;
    para[0] = I.contactAngle[0];
    para[1] = I.contactAngle[1];
  }
  double contactNormal[3] = {0, 0, 0}, center[3] = {0, 0, 0};
  double pDepth = 0.0;
  double p1[3], p2[3];
  if (A.isSphere && B.isSphere) {
    double dist[3] = {position2[0] - position1[0], position2[1] - position1[1], position2[2] - position1[2]};
    double r_sum = A.r_max + B.r_max;
    double depth = r_sum - norm3(dist);
    if (depth < 1E-18) {
      I.depth = 0;
      I.isShearNew = 1;
      I.nEval = nev;
      return;
    }
    normalize3(dist);
    for (int i = 0; i < 3; i++) p1[i] = position1[i] + A.r_max * dist[i];
    for (int i = 0; i < 3; i++) p2[i] = position2[i] - B.r_max * dist[i];
    for (int i = 0; i < 3; i++) contactNormal[i] = p1[i] - p2[i];
    normalize3(contactNormal);
    copy3(I.point1, p1);
    copy3(I.point2, p2);
    for (int i = 0; i < 3; i++) I.contactPoint[i] = 0.5 * (p1[i] + p2[i]);
    I.depth = depth;
  } else {
    double r_sum = A.r_max + B.r_max;
    double dd[3] = {position1[0] - position2[0], position1[1] - position2[1], position1[2] - position2[2]};
    if (norm3(dd) > r_sum) {
      I.depth = 0;
      I.isShearNew = 1;
      I.nEval = nev;
      return;
    }
    bool touching = true;
    const double opts[5] = {1.0, 1E-15, 1E-3, 1E-20, 1e-6};
    int stop = lmder2(para, 50, opts, A, B, position1, position2, p1, p2, center, touching, contactNormal, nev);
    if (std::isnan(para[0]) || std::isnan(para[1]) || 7 == stop) {
      if (touching) {
        double d0[3] = {position2[0] - position1[0], position2[1] - position1[1], position2[2] - position1[2]};
        para[0] = om::atan2(d0[1], d0[0]);
        para[1] = om::atan(d0[2] / std::sqrt(d0[0] * d0[0] + d0[1] * d0[1]));
        para[0] += 1e-5;
        para[1] += 1e-5;
        stop = lmder2(para, 50, opts, A, B, position1, position2, p1, p2, center, touching, contactNormal, nev) + 100;
      }
    }
    double p12[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
    pDepth = dot3(p12, contactNormal);
    if (touching) {
      double p21[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
      if (10.0 * norm3(p21) > A.r_max) {
        double q1[3] = {p2[0] - position1[0], p2[1] - position1[1], p2[2] - position1[2]};
        double q2[3] = {p1[0] - position2[0], p1[1] - position2[1], p1[2] - position2[2]};
        if (!(is_inside(A, q1)) || !(is_inside(B, q2))) {
          double d0[3] = {position2[0] - position1[0], position2[1] - position1[1], position2[2] - position1[2]};
          para[0] = om::atan2(d0[1], d0[0]);
          para[1] = om::atan(d0[2] / std::sqrt(d0[0] * d0[0] + d0[1] * d0[1]));
          para[0] += 1e-5;
          para[1] += 1e-5;
          stop = lmder2(para, 50, opts, A, B, position1, position2, p1, p2, center, touching, contactNormal, nev) + 200;
          double q[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
          pDepth = dot3(q, contactNormal);
        }
      }
    }
    I.stopCode = stop;
    if (!touching) {
      I.contactAngle[0] = para[0];
      I.contactAngle[1] = para[1];
      I.depth = 0;
      I.isShearNew = 1;
      I.nEval = nev;
      return;
    }
    copy3(I.point1, p1);
    copy3(I.point2, p2);
    copy3(I.contactPoint, center);
    I.depth = pDepth;
    normalize3(contactNormal);
  }
  precompute(I, b1, b2, c->prm.dt, contactNormal, I.isShearNew, shift2);
  copy3(I.normal, contactNormal);
  I.contactAngle[0] = para[0];
  I.contactAngle[1] = para[1];
  I.nEval = nev;
}

// Ig2_Superquadrics_Superquadrics_SuperquadricsGeom2::go (Superquadrics_Ig2.cpp:2322-2517).
void ig2_nm(sdm_ctx* c, Pair& I, const Body& b1, const Body& b2) {
  const double* position1 = b1.pos;
  const double* position2 = b2.pos;  // QUIRK: no shift2 (:2328); precompute gets it (:2511)
  double shift2[3];
  pair_shift2(c, I, shift2);
  SQ A = make_sq(b1), B = make_sq(b2);
  bool isNew = !I.hasGeom;
  double dels[2] = {0.01, 0.01};
  double para[2] = {0.0, 0.0};
  int nev = 0;
  I.stopCode = 0;
  if (isNew) {
    geom_reset(I);
    I.hasGeom = 1;
    I.contactPoint[0] = I.contactPoint[1] = I.contactPoint[2] = 0;
    I.isShearNew = 1;
    dels[0] = dels[1] = 0.1;
  } else {
    I.isShearNew = I.depth <= 0;
    dels[0] = dels[1] = 0.001;
  }
  double contactNormal[3] = {0, 0, 0};
  double p1[3], p2[3];
  if (A.isSphere && B.isSphere) {
    double dist[3] = {position2[0] - position1[0], position2[1] - position1[1], position2[2] - position1[2]};
    double r_sum = A.r_max + B.r_max;
    double depth = r_sum - norm3(dist);
    if (depth < 1E-18) {
      I.depth = 0;
      I.isShearNew = 1;
      I.nEval = nev;
      return;
    }
    normalize3(dist);
    for (int i = 0; i < 3; i++) p1[i] = position1[i] + A.r_max * dist[i];
    for (int i = 0; i < 3; i++) p2[i] = position2[i] - B.r_max * dist[i];
    for (int i = 0; i < 3; i++) contactNormal[i] = p1[i] - p2[i];
    normalize3(contactNormal);
    copy3(I.point1, p1);
    copy3(I.point2, p2);
    for (int i = 0; i < 3; i++) I.contactPoint[i] = 0.5 * (p1[i] + p2[i]);
    I.depth = depth;
    I.curvSum = r_sum / (A.r_max * B.r_max);
    I.geomIsSphere = 1;
  } else {
    double r_sum = A.r_max + B.r_max;
    double dd[3] = {position1[0] - position2[0], position1[1] - position2[1], position1[2] - position2[2]};
    if (norm3(dd) > r_sum) {
      I.depth = 0;
      I.isShearNew = 1;
      I.nEval = nev;
      return;
    }
    double gC[9];
    crotation_mat(position1, position2, gC);
    para[0] = I.contactAngle[0];
    para[1] = I.contactAngle[1];
    bool touching = true;
    double nm_info[2];
    neldermead_simplex(para, dels, nm_info, A, B, position1, position2, p1, p2, gC, touching, nev);
    I.stopCode = int(nm_info[0]);
    if (!touching) {
      I.contactAngle[0] = para[0];
      I.contactAngle[1] = para[1];
      I.depth = 0;
      I.isShearNew = 1;
      I.nEval = nev;
      return;
    }
    for (int i = 0; i < 3; i++) contactNormal[i] = p1[i] - p2[i];
    normalize3(contactNormal);
    copy3(I.point1, p1);
    copy3(I.point2, p2);
    for (int i = 0; i < 3; i++) I.contactPoint[i] = 0.5 * (p1[i] + p2[i]);
    I.depth = nm_info[1];
    // QUIRK: the curvature direction uses (cos a0 sin a1, sin a0 sin a1, cos a1) (:2463-2465).
    double c0[3] = {om::cos(para[0]) * om::sin(para[1]), om::sin(para[0]) * om::sin(para[1]), om::cos(para[1])};
    double contact[3], contact1[3], contact2[3];
    matvec(gC, c0, contact);
    matvec(A.Rl, contact, contact1);
    matvec(B.Rl, contact, contact2);
    double dmax1[3] = {0, 0, 0}, dmax2[3] = {0, 0, 0}, dmin1[3], dmin2[3], cd1[3], cd2[3];
    int f1, f2;
    double phi[2], k1[2], k2[2];
    normal2phi(A, contact1, phi);
    curvatures(A, phi, k1, dmax1, dmin1, &f1);
    double mc2[3] = {(-1) * contact2[0], (-1) * contact2[1], (-1) * contact2[2]};
    normal2phi(B, mc2, phi);
    curvatures(B, phi, k2, dmax2, dmin2, &f2);
    if (f1 && f2) {
      copy3(cd1, dmax1);
      copy3(cd2, dmax2);
    } else {
      copy3(cd1, dmin1);
      copy3(cd2, dmin2);
    }
    // QUIRK: component-wise products with the normal (:2483-2486).
    double dir1[3] = {cd1[0] * contactNormal[0], cd1[1] * contactNormal[1], cd1[2] * contactNormal[2]};
    double dir2[3] = {cd2[0] * contactNormal[0], cd2[1] * contactNormal[1], cd2[2] * contactNormal[2]};
    for (int i = 0; i < 3; i++) dir1[i] -= cd1[i];
    for (int i = 0; i < 3; i++) dir2[i] -= cd2[i];
    double angle = om::acos(dot3(dir1, dir2) / (norm3(dir1) * norm3(dir2)));
    if (std::isnan(angle)) angle = 0.0;
    I.curv1[0] = k1[0]; I.curv1[1] = k1[1];
    I.curv2[0] = k2[0]; I.curv2[1] = k2[1];
    I.curvAngle = angle;
    double hm[4];
    hertz_mindlin(k1, k2, angle, hm);
    I.curvSum = hm[0];
    I.alpha = hm[1];
    I.K = hm[2];
    I.E = hm[3];
  }
  precompute(I, b1, b2, c->prm.dt, contactNormal, I.isShearNew, shift2);
  I.contactAngle[0] = para[0];
  I.contactAngle[1] = para[1];
  I.nEval = nev;
}

// Ig2_Wall_Superquadrics_SuperquadricsGeom(2)::go (Superquadrics_Ig2.cpp:2522-2703).  The
// two variants differ only in the geom class they allocate.
void ig2_wall(sdm_ctx* c, Pair& I, const Body& w, const Body& b) {
  const int PA = w.wallAxis;
  const int sense = w.wallSense;
  SQ B = make_sq(b);
  bool isNew = !I.hasGeom;
  I.nEval = 0;
  I.stopCode = 0;
  if (isNew) {
    geom_reset(I);
    I.hasGeom = 1;
    I.contactAngle[0] = I.contactAngle[1] = 0;
    I.isShearNew = 1;
  } else {
    I.isShearNew = I.depth <= 0;
  }
  double normal[3] = {0, 0, 0};
  normal[PA] = sense;
  const double* wall_pos = w.pos;
  const double* B_pos = b.pos;
  if (B.isSphere) {
    double p1[3], point2[3];
    copy3(p1, B_pos);
    p1[PA] = wall_pos[PA];
    for (int i = 0; i < 3; i++) point2[i] = B_pos[i] - normal[i] * B.r_max;
    double dd[3] = {B_pos[0] - p1[0], B_pos[1] - p1[1], B_pos[2] - p1[2]};
    double depth = B.r_max - norm3(dd);
    if (depth < 0.0) {
      I.depth = 0;
      I.isShearNew = 1;
      return;
    }
    copy3(I.point1, p1);
    copy3(I.point2, point2);
    for (int i = 0; i < 3; i++) I.contactPoint[i] = 0.5 * (point2[i] + p1[i]);
    I.depth = depth;
  } else {
    double mn[3] = {-normal[0], -normal[1], -normal[2]};
    double ln[3], phi[2], t[3], point2[3];
    matvec(B.Rl, mn, ln);
    normal2phi(B, ln, phi);
    get_surface(B, phi, t);
    matvec(B.Rg, t, point2);
    for (int i = 0; i < 3; i++) point2[i] = point2[i] + B_pos[i];
    I.nEval = 1;
    double p1[3];
    copy3(p1, point2);
    p1[PA] = wall_pos[PA];
    double d[3] = {point2[0] - p1[0], point2[1] - p1[1], point2[2] - p1[2]};
    if (normal[PA] * d[PA] >= 0.0) {
      I.depth = 0;
      I.isShearNew = 1;
      return;
    }
    copy3(I.point1, p1);
    copy3(I.point2, point2);
    for (int i = 0; i < 3; i++) I.contactPoint[i] = 0.5 * (point2[i] + p1[i]);
    I.depth = norm3(d);
  }
  precompute(I, w, b, c->prm.dt, normal, I.isShearNew, zero3);
  copy3(I.normal, normal);
}

// Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys::go (Superquadrics.cpp:1086-1116)
// and Ip2_SuperquadricsMat2_SuperquadricsMat2_HertzMindlinPhys::go (:1118-1160).
// QUIRK: the wall branch is chosen by id1 <= 5 (:1100, :1134).
void ip2(sdm_ctx* c, Pair& I, const Body& b1, const Body& b2) {
  if (I.hasPhys) return;
  phys_reset(I);
  I.hasPhys = 1;
  const sdm_material& m1 = c->mats[b1.material];
  const sdm_material& m2 = c->mats[b2.material];
  const int id = I.id1;
  if (c->prm.law == SDM_LAW_LINEAR) {
    if (id > 5) {
      double frictionAngle = std::min(m1.frictionAngle, m2.frictionAngle);
      I.tanPhi = std::tan(frictionAngle);
      I.kn = 2.0 * m1.Kn * m2.Kn / (m1.Kn + m2.Kn);
      I.ks = 2.0 * m1.Ks * m2.Ks / (m1.Ks + m2.Ks);
    } else {
      I.tanPhi = std::tan(m1.frictionAngle);
      I.kn = m1.Kn;
      I.ks = m1.Ks;
    }
  } else {
    if (id > 5) {
      I.nuab = 0.5 * (m1.nu + m2.nu);
      I.Gab = 1.0 / ((1 - m1.nu) / m1.G + (1 - m2.nu) / m2.G);
      I.mu1 = 1. / m1.G + 1. / m2.G;
      I.mu2 = m1.nu / m1.G + m2.nu / m2.G;
      double frictionAngle = std::min(m1.frictionAngle, m2.frictionAngle);
      I.tanPhi = std::tan(frictionAngle);
    } else {
      I.tanPhi = std::tan(m1.frictionAngle);
      I.kn = m1.Kn;
      I.ks = m1.Ks;
    }
  }
  I.betan = std::max(m1.betan, m2.betan);
  I.betas = std::max(m1.betas, m2.betas);
}

// SuperquadricsLaw::go (Superquadrics.cpp:1181-1363) and SuperquadricsLaw2::go (:1367-1557).
// Returns false to request the erase.  `F/T1/T2/contributed` record the force scatter.
bool law(sdm_ctx* c, Pair& I, const Body& A, const Body& B) {
  const bool hertz = c->prm.law == SDM_LAW_HERTZ;
  I.contributed = 0;
  // AABB separation test, strict (:1203, :1381); SuperquadricsLaw shifts B's box by shift2
  // in a periodic scene (:1196-1201), QUIRK: SuperquadricsLaw2 does not (:1381)
  double bmn[3], bmx[3];
  copy3(bmn, B.bmin);
  copy3(bmx, B.bmax);
  if (c->prm.periodic && !hertz) {
    double shift2[3];
    pair_shift2(c, I, shift2);
    for (int i = 0; i < 3; i++) {
      bmn[i] += shift2[i];
      bmx[i] += shift2[i];
    }
  }
  if (A.bmin[0] >= bmx[0] || bmn[0] >= A.bmax[0] || A.bmin[1] >= bmx[1] || bmn[1] >= A.bmax[1] ||
      A.bmin[2] >= bmx[2] || bmn[2] >= A.bmax[2]) {
    set3(I.normalForce, 0., 0., 0.);
    set3(I.shearForce, 0., 0., 0.);
    return false;
  }
  if (!(I.depth > 1E-18)) {
    set3(I.normalForce, 0., 0., 0.);
    set3(I.shearForce, 0., 0., 0.);
    return true;
  }
  if (hertz && I.id1 > 5) {
    double depth = I.depth;
    double curvatures_sum = I.curvSum;
    if (I.geomIsSphere) {
      double tmp = std::sqrt(depth / curvatures_sum);
      I.kn = 8. / 3. * I.Gab * tmp;
      I.ks = 8. / (2.0 * I.mu1 - I.mu2) * tmp;
    } else {
      double E = I.E, K = I.K, alpha = I.alpha;
      double a = std::sqrt(depth * E / curvatures_sum / K);
      // QUIRK: M_PIl makes these two expressions long double (:1413, :1417).
      I.kn = 4. / 3. * M_PIl * a * I.Gab * alpha / K;
      double e2 = 1 - 1 / (alpha * alpha);
      double ksx = I.mu1 * K - I.mu2 * (K - E) / e2;
      double ksy = I.mu1 * K + I.mu2 * ((1 - e2) * K - E) / e2;
      I.ks = M_PIl * alpha * a * (1. / ksx + 1. / ksy);
    }
  }
  double normalForce[3];
  for (int i = 0; i < 3; i++) normalForce[i] = I.normal[i] * I.depth * I.kn;
  double shearForce1[3] = {0, 0, 0};
  bool useDamping = (I.betan > 1e-5 || I.betas > 1e-5);
  double cn = 0, cs = 0;
  double normalViscous[3] = {0, 0, 0}, shearViscous[3] = {0, 0, 0};
  if (useDamping) {
    bool Adyn = A.blockedDOFs != (int)DOF_ALL, Bdyn = B.blockedDOFs != (int)DOF_ALL;
    double mbar = (!Adyn && Bdyn) ? B.mass : ((!Bdyn && Adyn) ? A.mass : (A.mass * B.mass / (A.mass + B.mass)));
    double Cn_crit = 2. * std::sqrt(mbar * I.kn);
    double Cs_crit = 2. * std::sqrt(mbar * I.ks);
    cn = Cn_crit * I.betan;
    cs = Cs_crit * I.betas;
    for (int i = 0; i < 3; i++) normalViscous[i] = cn * I.relVn[i];
    double normTemp[3] = {normalForce[0] - normalViscous[0], normalForce[1] - normalViscous[1],
                          normalForce[2] - normalViscous[2]};
    if (dot3(normTemp, I.normal) < 0.0) {
      copy3(normalViscous, normalForce);
      set3(normalForce, 0, 0, 0);
    } else {
      for (int i = 0; i < 3; i++) normalForce[i] -= normalViscous[i];
    }
  }
  if (I.isShearNew) {
    set3(shearForce1, 0., 0., 0.);
  } else {
    rotate_shear(I, I.shearForce);
    copy3(shearForce1, I.shearForce);
  }
  for (int i = 0; i < 3; i++) shearForce1[i] -= I.ks * I.shearInc[i];
  double tp = I.tanPhi;
  double maxFs = sqn3(normalForce) * (tp * tp);
  if (sqn3(shearForce1) > maxFs) {
    double ratio = std::sqrt(maxFs) / norm3(shearForce1);
    for (int i = 0; i < 3; i++) shearForce1[i] *= ratio;
  } else if (useDamping) {
    for (int i = 0; i < 3; i++) shearViscous[i] = cs * I.relVs[i];
    if (hertz)
      for (int i = 0; i < 3; i++) shearForce1[i] -= shearViscous[i];
  }
  if (std::isnan(shearForce1[0]) || std::isnan(normalForce[0])) {
    set3(normalForce, 0, 0, 0);
    set3(shearForce1, 0, 0, 0);
  }
  double F[3];
  if (hertz) {
    for (int i = 0; i < 3; i++) F[i] = -normalForce[i] - shearForce1[i];
  } else {
    for (int i = 0; i < 3; i++) F[i] = -normalForce[i] - shearForce1[i] + shearViscous[i];
  }
  double ra[3] = {A.pos[0] - I.contactPoint[0], A.pos[1] - I.contactPoint[1], A.pos[2] - I.contactPoint[2]};
  double rb[3] = {B.pos[0] - I.contactPoint[0], B.pos[1] - I.contactPoint[1], B.pos[2] - I.contactPoint[2]};
  double ta[3], tb[3];
  cross3(ra, F, ta);
  cross3(rb, F, tb);
  copy3(I.F, F);
  for (int i = 0; i < 3; i++) I.T1[i] = -ta[i];
  copy3(I.T2, tb);
  I.contributed = 1;
  copy3(I.normalForce, normalForce);
  copy3(I.shearForce, shearForce1);
  return true;
}

// Interaction::reset via InteractionContainer::requestErase (core/Interaction.cpp:25-29).
void request_erase(Pair& I) {
  I.hasGeom = 0;
  I.hasPhys = 0;
  I.geomExists = 1;
  I.iterMadeReal = -1;
  geom_reset(I);
  phys_reset(I);
}

// ---------------------------------------------------------------------------------------
// InteractionLoop::action (pkg/common/InteractionLoop.cpp:23-149)
// ---------------------------------------------------------------------------------------
void process_pair(sdm_ctx* c, Pair& I, int64_t& contactEvals, int64_t& distEvals) {
  I.contributed = 0;
  if (!I.geomExists) return;
  Body* b1 = &c->bodies[I.id1];
  Body* b2 = &c->bodies[I.id2];
  if (!I.functorSet) {
    // IGeomDispatcher::getFunctor2D with DEFINE_FUNCTOR_ORDER_2D: (Wall,Superquadrics)
    // and (Superquadrics,Superquadrics) are declared; (Superquadrics,Wall) swaps.
    int k1 = b1->kind, k2 = b2->kind;
    if (k1 == SDM_BODY_WALL && k2 == SDM_BODY_WALL) {
      I.geomExists = 0;
      return;
    }
    I.functorSet = 1;
    if (k1 == SDM_BODY_SUPERQUADRIC && k2 == SDM_BODY_WALL) {
      std::swap(I.id1, I.id2);  // Interaction::swapOrder (core/Interaction.cpp:32-38)
      std::swap(b1, b2);
    }
  }
  bool wasReal = I.hasGeom && I.hasPhys;
  if (b1->kind == SDM_BODY_WALL) {
    ig2_wall(c, I, *b1, *b2);
  } else {
    contactEvals++;
    if (c->prm.law == SDM_LAW_LINEAR)
      ig2_lm(c, I, *b1, *b2);
    else
      ig2_nm(c, I, *b1, *b2);
    distEvals += I.nEval;
  }
  ip2(c, I, *b1, *b2);
  if (!wasReal) I.iterMadeReal = c->iter;
  if (!law(c, I, *b1, *b2)) request_erase(I);
}

void interaction_loop(sdm_ctx* c) {
  const long size = (long)c->intrs.size();
  const int N = (int)c->bodies.size();
  int64_t ce = 0, de = 0;
  if (c->nthreads <= 1) {
    for (long i = 0; i < size; i++) process_pair(c, c->intrs[i], ce, de);
    // serial ForceContainer: addForce/addTorque in linIntrs order (core/ForceContainer.hpp:218-319)
    for (long i = 0; i < size; i++) {
      const Pair& I = c->intrs[i];
      if (!I.contributed) continue;
      Body& A = c->bodies[I.id1];
      Body& B = c->bodies[I.id2];
      for (int k = 0; k < 3; k++) A.force[k] += I.F[k];
      for (int k = 0; k < 3; k++) B.force[k] += -I.F[k];
      for (int k = 0; k < 3; k++) A.torque[k] += I.T1[k];
      for (int k = 0; k < 3; k++) B.torque[k] += I.T2[k];
    }
  } else {
#ifdef _OPENMP
    const int nt = c->nthreads;
    c->thrForce.assign((size_t)nt * N * 6, 0.0);
#pragma omp parallel num_threads(nt) reduction(+ : ce, de)
    {
      double* buf = &c->thrForce[(size_t)omp_get_thread_num() * N * 6];
#pragma omp for schedule(guided)
      for (long i = 0; i < size; i++) {
        Pair& I = c->intrs[i];
        process_pair(c, I, ce, de);
        if (!I.contributed) continue;
        double* fa = buf + 6 * (size_t)I.id1;
        double* fb = buf + 6 * (size_t)I.id2;
        for (int k = 0; k < 3; k++) fa[k] += I.F[k];
        for (int k = 0; k < 3; k++) fb[k] += -I.F[k];
        for (int k = 0; k < 3; k++) fa[3 + k] += I.T1[k];
        for (int k = 0; k < 3; k++) fb[3 + k] += I.T2[k];
      }
    }
    // ForceContainer::sync: sum the thread buffers in thread order (ForceContainer.hpp:134-162)
#pragma omp parallel for num_threads(nt)
    for (int id = 0; id < N; id++) {
      double s[6] = {0, 0, 0, 0, 0, 0};
      for (int t = 0; t < nt; t++)
        for (int k = 0; k < 6; k++) s[k] += c->thrForce[((size_t)t * N + id) * 6 + k];
      for (int k = 0; k < 3; k++) c->bodies[id].force[k] = s[k];
      for (int k = 0; k < 3; k++) c->bodies[id].torque[k] = s[3 + k];
    }
#endif
  }
  c->cnt.contact_evals += ce;
  c->cnt.dist_evals += de;
}

// ---------------------------------------------------------------------------------------
// InsertionSortCollider (pkg/common/InsertionSortCollider.cpp:22-54, 177-400) with
// BoundDispatcher::processBody (pkg/common/Dispatching.cpp:30-62).
// ---------------------------------------------------------------------------------------
inline bool spatial_overlap(const sdm_ctx* c, int id1, int id2) {
  const double* mi = c->minima.data();
  const double* ma = c->maxima.data();
  return (mi[3 * id1 + 0] <= ma[3 * id2 + 0]) && (ma[3 * id1 + 0] >= mi[3 * id2 + 0]) &&
         (mi[3 * id1 + 1] <= ma[3 * id2 + 1]) && (ma[3 * id1 + 1] >= mi[3 * id2 + 1]) &&
         (mi[3 * id1 + 2] <= ma[3 * id2 + 2]) && (ma[3 * id1 + 2] >= mi[3 * id2 + 2]);
}

void insert_pair(sdm_ctx* c, int id1, int id2, const int32_t* cellDist = nullptr) {
  // InteractionContainer::insert (core/InteractionContainer.cpp:9-34)
  if (!c->found_set.insert(pair_key(id1, id2)).second) return;
  Pair p = new_pair(id1, id2);
  if (cellDist)
    for (int k = 0; k < 3; k++) p.cellDist[k] = cellDist[k];
  p.iterBorn = c->iter;
  c->intrs.push_back(p);
  c->cnt.pairs_inserted++;
}

// handleBoundInversion (InsertionSortCollider.cpp:22-27).  Collider::mayCollide
// (pkg/common/Collider.cpp:14-27) is true for every pair of standalone bodies with the
// default groupMask=1 and avoidSelfInteractionMask=0.
void handle_bound_inversion(sdm_ctx* c, int id1, int id2) {
  if (spatial_overlap(c, id1, id2) && !c->found_set.count(pair_key(id1, id2))) insert_pair(c, id1, id2);
}

// insertionSort (InsertionSortCollider.cpp:29-54), serial flavour.
void insertion_sort(sdm_ctx* c, std::vector<Bounds>& v) {
  const long size = (long)v.size();
  for (long i = 1; i < size; i++) {
    const Bounds viInit = v[i];
    long j = i - 1;
    const bool viInitBB = viInit.hasBB;
    const bool isMin = viInit.isMin;
    while (j >= 0 && v[j] > viInit) {
      v[j + 1] = v[j];
      if (isMin && !v[j].isMin && viInitBB && v[j].hasBB && (viInit.id != v[j].id))
        handle_bound_inversion(c, viInit.id, v[j].id);
      j--;
    }
    v[j + 1] = viInit;
  }
}

// ---- periodic cell (InsertionSortCollider.cpp:404-543) ----
// cellWrap: x wrapped into [x0, x1), the period number returned in `period`.
double cell_wrap(double x, double x0, double x1, int& period) {
  const double xNorm = (x - x0) / (x1 - x0);
  period = (int)std::floor(xNorm);
  return x0 + (xNorm - period) * (x1 - x0);
}
// cellWrapRel: x wrapped into [0, x1 - x0) relative to x0.
double cell_wrap_rel(double x, double x0, double x1) {
  const double xNorm = (x - x0) / (x1 - x0);
  return (xNorm - std::floor(xNorm)) * (x1 - x0);
}

// spatialOverlapPeri (:491-543) with allowBiggerThanPeriod = false: per axis, both boxes
// are wrapped into the period that starts at the minimum giving the smaller gap; a box
// whose maximum lands in another period than its minimum raises the error flag (the
// reference throws).  `periods` receives the period of id2 relative to id1.
bool spatial_overlap_peri(sdm_ctx* c, int id1, int id2, int32_t periods[3]) {
  const double* mi = c->minima.data();
  const double* ma = c->maxima.data();
  for (int axis = 0; axis < 3; axis++) {
    const double dim = c->prm.cell[axis];
    const double m1 = mi[3 * id1 + axis], m2 = mi[3 * id2 + axis];
    const double wMn = (cell_wrap_rel(m1, m2, m2 + dim) < cell_wrap_rel(m2, m1, m1 + dim)) ? m2 : m1;
    int pmn1, pmx1, pmn2, pmx2;
    const double mn1 = cell_wrap(m1, wMn, wMn + dim, pmn1), mx1 = cell_wrap(ma[3 * id1 + axis], wMn, wMn + dim, pmx1);
    const double mn2 = cell_wrap(m2, wMn, wMn + dim, pmn2), mx2 = cell_wrap(ma[3 * id2 + axis], wMn, wMn + dim, pmx2);
    if (pmn1 != pmx1 || pmn2 != pmx2) {
      c->periErr = true;
      return false;
    }
    periods[axis] = pmn1 - pmn2;
    if (!(mn1 <= mx2 && mx1 >= mn2)) return false;
  }
  return true;
}

// handleBoundInversionPeri (:466-475).
void handle_bound_inversion_peri(sdm_ctx* c, int id1, int id2) {
  int32_t periods[3] = {0, 0, 0};
  const bool overlap = spatial_overlap_peri(c, id1, id2, periods);
  if (overlap && !c->found_set.count(pair_key(id1, id2))) insert_pair(c, id1, id2, periods);
}

// insertionSortPeri (:416-463): the bounds of one axis form a ring sorted from loIdx; a
// bound that moves across the split changes period and coordinate by one cell.
void insertion_sort_peri(sdm_ctx* c, std::vector<Bounds>& v, long& loIdx, double cellDim, bool doCollide) {
  const long size = (long)v.size();
  auto norm = [size](long i) {
    if (i < 0) i += size;
    return i % size;
  };
  for (long ii = 0; ii < size; ii++) {
    const long i = norm(ii);
    const long i_1 = norm(i - 1);
    if (i == loIdx && v[i].coord < 0) {
      v[i].period -= 1;
      v[i].coord += cellDim;
      loIdx = norm(loIdx + 1);
    }
    const double iCmpCoord = v[i].coord + (i == loIdx ? cellDim : 0);
    if (v[i_1].coord <= iCmpCoord) continue;
    long j = i_1;
    Bounds vi = v[i];
    const bool viHasBB = vi.hasBB;
    const bool isMin = v[i].isMin;
    while (v[j].coord > vi.coord + (norm(j + 1) == loIdx ? cellDim : 0)) {
      const long j1 = norm(j + 1);
      Bounds& vNew = v[j1];
      vNew = v[j];
      if (j == loIdx && vi.coord < 0) {
        vi.period -= 1;
        vi.coord += cellDim;
        loIdx = norm(loIdx + 1);
      } else if (j1 == loIdx) {
        vNew.period += 1;
        vNew.coord -= cellDim;
        loIdx = norm(loIdx - 1);
      }
      if (isMin && !v[j].isMin && (doCollide && viHasBB && v[j].hasBB)) {
        if (vi.id != vNew.id) handle_bound_inversion_peri(c, vi.id, vNew.id);
      }
      j = norm(j - 1);
    }
    v[norm(j + 1)] = vi;
  }
}

// BoundDispatcher::processBody (Dispatching.cpp:30-62) with Bo1_Superquadrics_Aabb::go
// (Superquadrics.cpp:639-686) and Bo1_Wall_Aabb::go (pkg/common/Wall.cpp:54-62).
void process_body(sdm_ctx* c, Body& b, double sweepDist) {
  const double targetInterv = c->prm.target_interv;
  if (b.hasBound) {
    double& sweepLength = b.sweepLength;
    if (c->prm.target_interv >= 0) {
      double disp[3] = {b.pos[0] - b.refPos[0], b.pos[1] - b.refPos[1], b.pos[2] - b.refPos[2]};
      double dist = std::max(std::abs(disp[0]), std::max(std::abs(disp[1]), std::abs(disp[2])));
      if (dist) {
        double newLength = dist * targetInterv / (double)(c->iter - b.lastUpdateIter);
        newLength = std::max(0.9 * sweepLength, newLength);
        sweepLength = std::max(c->prm.min_sweep_dist_factor * sweepDist, std::min(newLength, sweepDist));
      } else
        sweepLength = 0;
    } else
      sweepLength = sweepDist;
  }
  if (b.kind == SDM_BODY_SUPERQUADRIC) {
    double r = b.r_max;
    for (int i = 0; i < 3; i++) b.bmin[i] = b.pos[i] + (-r);
    for (int i = 0; i < 3; i++) b.bmax[i] = b.pos[i] + r;
  } else {
    for (int i = 0; i < 3; i++) { b.bmin[i] = -Inf; b.bmax[i] = Inf; }
    b.bmin[b.wallAxis] = b.pos[b.wallAxis];
    b.bmax[b.wallAxis] = b.pos[b.wallAxis];
  }
  if (!b.hasBound) {  // the functor allocated the bound: default sweepLength 0 (Bound.hpp:27)
    b.hasBound = 1;
    b.sweepLength = 0;
  }
  copy3(b.refPos, b.pos);
  b.lastUpdateIter = (int32_t)c->iter;
  const double sl = b.sweepLength;
  if (sl > 0) {
    for (int i = 0; i < 3; i++) b.bmin[i] -= sl;
    for (int i = 0; i < 3; i++) b.bmax[i] += sl;
  }
}

// InsertionSortCollider::isActivated (InsertionSortCollider.cpp:177-189).
bool collider_is_activated(sdm_ctx* c) {
  if (!c->strideActive) return true;
  if (c->fastestBodyMaxDist < 0) {
    c->fastestBodyMaxDist = 0;
    return true;
  }
  c->fastestBodyMaxDist = c->maxVelocitySq;
  if (c->fastestBodyMaxDist >= 1 || c->fastestBodyMaxDist == 0) return true;
  if ((long)c->BB[0].size() != 2 * (long)c->bodies.size()) return true;
  return false;
}

// InsertionSortCollider::action (InsertionSortCollider.cpp:191-400).
void collider_action(sdm_ctx* c) {
  c->cnt.collider_runs++;
  if (c->found_dirty) build_found(c);
  const long nBodies = (long)c->bodies.size();
  const bool periodic = c->prm.periodic != 0;
  if (periodic != c->periodicState) {  // periodicity changed: reinitialise (:202-206)
    for (int i = 0; i < 3; i++) c->BB[i].clear();
    c->periodicState = periodic;
  }
  bool doInitSort = false;
  if ((long)c->BB[0].size() != 2 * nBodies) {
    long BBsize = (long)c->BB[0].size();
    if (2 * nBodies < BBsize)
      for (int i = 0; i < 3; i++) c->BB[i].clear();
    BBsize = (long)c->BB[0].size();
    if (2 * nBodies - BBsize > 200 || BBsize == 0) doInitSort = true;
    for (int i = 0; i < 3; i++) {
      for (long id = BBsize / 2; id < nBodies; id++) {
        c->BB[i].push_back(Bounds{0, (int32_t)id, 1, 0});
        c->BB[i].push_back(Bounds{0, (int32_t)id, 0, 0});
      }
    }
  }
  c->minima.resize(3 * nBodies);
  c->maxima.resize(3 * nBodies);
  const double verletDist = c->prm.verlet_dist;
  // bound dispatch (Dispatching.cpp:15-26), parallel over bodies
#pragma omp parallel for num_threads(c->nthreads) if (c->nthreads > 1)
  for (long id = 0; id < nBodies; id++) process_body(c, c->bodies[id], verletDist);
  if (!c->strideActive && verletDist > 0 && c->maxVelocitySq >= 0) c->strideActive = 1;
  // copy bounds (InsertionSortCollider.cpp:289-311)
  for (int j = 0; j < 3; j++) {
    std::vector<Bounds>& V = c->BB[j];
    for (long i = 0; i < 2 * nBodies; i++) {
      Bounds& bb = V[i];
      const Body& b = c->bodies[bb.id];
      bb.hasBB = b.hasBound;
      bb.coord = (b.hasBound ? (bb.isMin ? b.bmin[j] : b.bmax[j]) : b.pos[j]) -
                 (periodic ? c->prm.cell[j] * bb.period : 0.);
      if (doInitSort && periodic) bb.coord = cell_wrap(bb.coord, 0, c->prm.cell[j], bb.period);
    }
  }
  for (long id = 0; id < nBodies; id++) {
    const Body& b = c->bodies[id];
    if (!b.hasBound) continue;
    for (int k = 0; k < 3; k++) {
      c->minima[3 * id + k] = b.bmin[k];
      c->maxima[3 * id + k] = b.bmax[k];
    }
  }
  // conditionalyEraseNonReal (core/InteractionContainer.hpp:104-112)
  for (size_t linPos = 0; linPos < c->intrs.size();) {
    Pair& I = c->intrs[linPos];
    int32_t periods[3];
    if (!(I.hasGeom && I.hasPhys) &&
        !(periodic ? spatial_overlap_peri(c, I.id1, I.id2, periods) : spatial_overlap(c, I.id1, I.id2))) {
      c->found_set.erase(pair_key(I.id1, I.id2));
      if (linPos < c->intrs.size() - 1) c->intrs[linPos] = c->intrs.back();
      c->intrs.pop_back();
      c->cnt.pairs_erased++;
    } else
      linPos++;
  }
  if (!doInitSort) {
    if (!periodic)
      for (int i = 0; i < 3; i++) insertion_sort(c, c->BB[i]);
    else
      for (int i = 0; i < 3; i++) insertion_sort_peri(c, c->BB[i], c->loIdx[i], c->prm.cell[i], true);
  } else {
    for (int i = 0; i < 3; i++) {
      c->loIdx[i] = 0;
      if (c->initSortStable)
        std::stable_sort(c->BB[i].begin(), c->BB[i].end());
      else
        std::sort(c->BB[i].begin(), c->BB[i].end());
    }
    c->numReinit++;
    const std::vector<Bounds>& V = c->BB[c->prm.sort_axis];
    const long n2 = 2 * nBodies;
    if (periodic) {
      // the sweep wraps over the end of the ring (:384-397)
      for (long i = 0; i < n2; i++) {
        if (!(V[i].isMin && V[i].hasBB)) continue;
        const int iid = V[i].id;
        for (long j = (i + 1) % n2; V[j].id != iid; j = (j + 1) % n2) {
          if (!(V[j].isMin && V[j].hasBB)) continue;
          handle_bound_inversion_peri(c, iid, V[j].id);
        }
      }
      return;
    }
    for (long i = 0; i < 2 * nBodies; i++) {
      if (!(V[i].isMin && V[i].hasBB)) continue;
      const int iid = V[i].id;
      for (long j = i + 1; j < 2 * nBodies && V[j].id != iid; j++) {
        const int jid = V[j].id;
        if (!(V[j].isMin && V[j].hasBB)) continue;
        if (spatial_overlap(c, iid, jid) && !c->found_set.count(pair_key(iid, jid))) insert_pair(c, iid, jid);
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// NewtonIntegrator::action (pkg/dem/NewtonIntegrator.cpp:99-304) for isSuperquadrics=1.
// ---------------------------------------------------------------------------------------
void dotq(const double* w, const double* Q, double* o) {  // DotQ (:470-477)
  o[0] = (-Q[1] * w[0] - Q[2] * w[1] - Q[3] * w[2]) / 2;
  o[1] = (Q[0] * w[0] - Q[3] * w[1] + Q[2] * w[2]) / 2;
  o[2] = (Q[3] * w[0] + Q[0] * w[1] - Q[1] * w[2]) / 2;
  o[3] = (-Q[2] * w[0] + Q[1] * w[1] + Q[0] * w[2]) / 2;
}

double newton_body(sdm_ctx* c, Body& b) {
  const double dt = c->prm.dt;
  const double damping = c->prm.damping;
  const double* gravity = c->prm.gravity;
  double f[3], m[3];
  copy3(f, b.force);
  copy3(m, b.torque);
  const unsigned blocked = (unsigned)b.blockedDOFs;
  const bool aspherical = b.kind == SDM_BODY_SUPERQUADRIC;  // setAspherical(true) in the factories
  const bool useAspherical = c->prm.exact_aspherical_rot && aspherical && blocked != DOF_ALL;
  if (blocked != DOF_ALL) {
    double linAccel[3];
    if (blocked == 0) {  // computeAccel (:27-32)
      for (int i = 0; i < 3; i++) linAccel[i] = f[i] / b.mass + gravity[i];
    } else {
      for (int i = 0; i < 3; i++) {
        linAccel[i] = 0;
        if (!(blocked & axisDOF(i, false))) linAccel[i] += f[i] / b.mass + gravity[i];
      }
    }
    if (b.isDamped)  // cundallDamp2nd (:23-25)
      for (int i = 0; i < 3; i++) linAccel[i] *= 1 - damping * Sign(linAccel[i] * (b.vel[i] + 0.5 * dt * linAccel[i]));
    for (int i = 0; i < 3; i++) b.vel[i] += dt * linAccel[i];
    if (!useAspherical) {
      double angAccel[3];
      if (blocked == 0) {
        for (int i = 0; i < 3; i++) angAccel[i] = m[i] / b.inertia[i];
      } else {
        for (int i = 0; i < 3; i++) {
          angAccel[i] = 0;
          if (!(blocked & axisDOF(i, true))) angAccel[i] += m[i] / b.inertia[i];
        }
      }
      if (b.isDamped)
        for (int i = 0; i < 3; i++)
          angAccel[i] *= 1 - damping * Sign(angAccel[i] * (b.angVel[i] + 0.5 * dt * angAccel[i]));
      for (int i = 0; i < 3; i++) b.angVel[i] += dt * angAccel[i];
    } else {
      for (int i = 0; i < 3; i++)
        if (blocked & axisDOF(i, true)) m[i] = 0;
      if (b.isDamped) {
        if (b.isSphere) {
          double angAccel[3];
          if (blocked == 0) {
            for (int i = 0; i < 3; i++) angAccel[i] = m[i] / b.inertia[i];
          } else {
            for (int i = 0; i < 3; i++) {
              angAccel[i] = 0;
              if (!(blocked & axisDOF(i, true))) angAccel[i] += m[i] / b.inertia[i];
            }
          }
          for (int i = 0; i < 3; i++)
            angAccel[i] *= 1 - damping * Sign(angAccel[i] * (b.angVel[i] + 0.5 * dt * angAccel[i]));
          for (int i = 0; i < 3; i++) b.angVel[i] += dt * angAccel[i];
        } else {  // cundallDamp1st (:19-21)
          for (int i = 0; i < 3; i++) m[i] *= 1 - damping * Sign(m[i] * b.angVel[i]);
        }
      }
    }
  }
  if (c->servo.quiet) {  // quiet_system_flag (NewtonIntegrator.cpp:246-250): no move this step
    set3(b.vel, 0, 0, 0);
    set3(b.angVel, 0, 0, 0);
  } else {
  bool sphereRotate = !useAspherical || b.isSphere;
  if (!sphereRotate) {
    // leapfrogSuperquadricsRotate (:391-429) -> leapfrogAsphericalRotate (:306-335)
    const double* A = b.rot_local;
    double l_n[3], l_b_n[3], w_b_n[3], dQ_n[4], Q_half[4], l_b_half[3], w_b_half[3], dQ_half[4];
    for (int i = 0; i < 3; i++) l_n[i] = b.angMom[i] + dt / 2. * m[i];
    matvec(A, l_n, l_b_n);
    for (int i = 0; i < 3; i++) w_b_n[i] = l_b_n[i] / b.inertia[i];
    dotq(w_b_n, b.ori, dQ_n);
    for (int i = 0; i < 4; i++) Q_half[i] = b.ori[i] + dQ_n[i] * (dt / 2.);
    for (int i = 0; i < 3; i++) b.angMom[i] += dt * m[i];
    matvec(A, b.angMom, l_b_half);
    for (int i = 0; i < 3; i++) w_b_half[i] = l_b_half[i] / b.inertia[i];
    dotq(w_b_half, Q_half, dQ_half);
    for (int i = 0; i < 4; i++) b.ori[i] = b.ori[i] + dQ_half[i] * dt;
    quat_rotate(b.ori, w_b_half, b.angVel);
    quat_normalize(b.ori);
    double qc[4];
    quat_conj(b.ori, qc);
    quat_to_rot(qc, b.rot_local);
    quat_to_rot(b.ori, b.rot_global);
  } else {
    // leapfrogSphericalRotate (:319-334)
    double angle2 = sqn3(b.angVel);
    if (angle2 != 0) {
      double angle = std::sqrt(angle2);
      double ax[3] = {b.angVel[0] / angle, b.angVel[1] / angle, b.angVel[2] / angle};
      double ha = angle * dt * 0.5;
      double s = om::sin(ha);
      double q[4] = {om::cos(ha), s * ax[0], s * ax[1], s * ax[2]};
      double o[4];
      quat_mul(q, b.ori, o);
      for (int i = 0; i < 4; i++) b.ori[i] = o[i];
    }
    quat_normalize(b.ori);
  }
  for (int i = 0; i < 3; i++) b.pos[i] += b.vel[i] * dt;  // leapfrogTranslate (:306-317)
  }
  // saveMaximaDisplacement (:72-85)
  if (!b.hasBound) return 0;
  double disp[3] = {b.pos[0] - b.refPos[0], b.pos[1] - b.refPos[1], b.pos[2] - b.refPos[2]};
  double maxDisp = std::max(std::abs(disp[0]), std::max(std::abs(disp[1]), std::abs(disp[2])));
  if (!maxDisp || maxDisp < b.sweepLength)
    maxDisp = 0.5;
  else
    maxDisp = 2;
  return maxDisp;
}

void newton(sdm_ctx* c) {
  const long N = (long)c->bodies.size();
  double mx = 0;
#pragma omp parallel for num_threads(c->nthreads) if (c->nthreads > 1) reduction(max : mx)
  for (long i = 0; i < N; i++) {
    double v = newton_body(c, c->bodies[i]);
    mx = std::max(mx, v);
  }
  c->maxVelocitySq = mx;
}

// ---------------------------------------------------------------------------------------
// CompressionEngine, cubic specimen: consolidation, hydroConsolidation and triaxial shear
// (pkg/dem/Compression.cpp with SOLE_GAIN and NEW_SCHEME defined, Compression.hpp:26-29).  Walls: bodies 0-5 = left, right, front, back,
// bottom, top.
// ---------------------------------------------------------------------------------------
// updateBoxSize (:99-113)
void servo_box(sdm_ctx* c) {
  auto& S = c->servo;
  const auto& B = c->bodies;
  S.width = B[1].pos[0] - B[0].pos[0];
  S.depth = B[3].pos[1] - B[2].pos[1];
  S.height = B[5].pos[2] - B[4].pos[2];
  S.x_area = S.depth * S.height;
  S.y_area = S.width * S.height;
  S.z_area = S.depth * S.width;
}
// consol_ss, SOLE_GAIN branch (:1330-1345)
void servo_consol_ss(sdm_ctx* c) {
  auto& S = c->servo;
  const auto& B = c->bodies;
  servo_box(c);
  S.ws[1] = B[1].force[0] / S.x_area;
  S.ws[0] = -B[0].force[0] / S.x_area;
  S.wsxx = 0.5 * (S.ws[0] + S.ws[1]);
  S.ws[3] = B[3].force[1] / S.y_area;
  S.ws[2] = -B[2].force[1] / S.y_area;
  S.wsyy = 0.5 * (S.ws[3] + S.ws[2]);
  S.ws[5] = B[5].force[2] / S.z_area;
  S.ws[4] = -B[4].force[2] / S.z_area;
  S.wszz = 0.5 * (S.ws[5] + S.ws[4]);
}
// get_gain, SOLE_GAIN branch (:1125-1220): kn summed per wall over the real interactions in
// container order (the wall is the smaller id).
void servo_get_gain(sdm_ctx* c) {
  auto& S = c->servo;
  for (int j = 0; j < 6; j++) S.stiff[j] = 0.0;
  for (const Pair& I : c->intrs) {
    if (!(I.hasGeom && I.hasPhys)) continue;
    const int id = I.id1 < I.id2 ? I.id1 : I.id2;
    if (id < 6) S.stiff[id] += I.kn;
  }
  const double a = S.p.gain_alpha;
  S.gain[0] = a * S.x_area / S.stiff[0];
  S.gain[2] = a * S.y_area / S.stiff[2];
  S.gain[4] = a * S.z_area / S.stiff[4];
  S.gain[1] = a * S.x_area / S.stiff[1];
  S.gain[3] = a * S.y_area / S.stiff[3];
  S.gain[5] = a * S.z_area / S.stiff[5];
}
// Mathr::Sign(u) * std::min(fabs(u), lim)
double servo_clamp(double u, double lim) {
  const double a = std::abs(u);
  return Sign(u) * std::min(a, lim);
}
// servo, SOLE_GAIN branch (:1252-1316)
void servo_move(sdm_ctx* c, double dt) {
  auto& S = c->servo;
  auto& B = c->bodies;
  servo_consol_ss(c);
  const double lim = S.p.max_vel * dt;
  if (S.p.activated[0]) B[0].pos[0] += -servo_clamp(S.gain[0] * (S.ws[0] - S.p.goal[0]), lim);
  if (S.p.activated[1]) B[1].pos[0] += servo_clamp(S.gain[1] * (S.ws[1] - S.p.goal[0]), lim);
  if (S.p.activated[2]) B[2].pos[1] += -servo_clamp(S.gain[2] * (S.ws[2] - S.p.goal[1]), lim);
  if (S.p.activated[3]) B[3].pos[1] += servo_clamp(S.gain[3] * (S.ws[3] - S.p.goal[1]), lim);
  if (S.p.z_servo) {
    B[4].pos[2] += -servo_clamp(S.gain[4] * (S.ws[4] - S.p.goal[2]), lim);
    B[5].pos[2] += servo_clamp(S.gain[5] * (S.ws[5] - S.p.goal[2]), lim);
  }
}
// getStressTensor (:764-829) and getStress (:831-841): sum over real interactions with a
// non-zero normal force of (fn + fs) (pos2 - pos1)^T over the box volume; mean = trace / 3,
// deviator = |Sigma - mean I| sqrt(1.5).  Sums in the order of Eigen's unvectorised redux.
void servo_get_stress(sdm_ctx* c, double& mean, double& dev) {
  auto& S = c->servo;
  servo_box(c);
  const double vol = S.height * S.z_area;
  double Sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // row-major Sigma(r, c)
  for (const Pair& I : c->intrs) {
    if (!(I.hasGeom && I.hasPhys)) continue;
    if (norm3(I.normalForce) == 0) continue;
    const double* x1 = c->bodies[I.id1].pos;
    const double* x2 = c->bodies[I.id2].pos;
    double f[3], l[3];
    for (int k = 0; k < 3; k++) {
      f[k] = I.normalForce[k] + I.shearForce[k];
      l[k] = x2[k] - x1[k];
    }
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) Sg[3 * r + k] += f[r] * l[k];
  }
  for (int k = 0; k < 9; k++) Sg[k] /= vol;
  mean = (Sg[0] + (Sg[4] + Sg[8])) / 3.0;
  Sg[0] -= mean;
  Sg[4] -= mean;
  Sg[8] -= mean;
  double q[9];  // squares in column-major order: (0,0), (1,0), (2,0), (0,1), ...
  for (int col = 0; col < 3; col++)
    for (int row = 0; row < 3; row++) q[3 * col + row] = Sg[3 * row + col] * Sg[3 * row + col];
  const double sq = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + (q[7] + q[8])));
  dev = std::sqrt(sq) * std::sqrt(1.5);
}
// Shop::unbalancedForce(false) (pkg/dem/Shop_01.cpp:148-189): sum over bodies that are not
// fully blocked of |F + m g| over twice the sum over real interactions with a non-zero normal
// force of |fn + fs|.
double unbalanced_force(sdm_ctx* c) {
  const double* g = c->prm.gravity;
  double sumF = 0;
  for (const Body& b : c->bodies) {
    if ((unsigned)b.blockedDOFs == DOF_ALL) continue;
    const double v[3] = {b.force[0] + b.mass * g[0], b.force[1] + b.mass * g[1], b.force[2] + b.mass * g[2]};
    sumF += norm3(v);
  }
  double sumC = 0;
  for (const Pair& I : c->intrs) {
    if (!(I.hasGeom && I.hasPhys)) continue;
    if (norm3(I.normalForce) == 0) continue;
    const double v[3] = {I.normalForce[0] + I.shearForce[0], I.normalForce[1] + I.shearForce[1],
                         I.normalForce[2] + I.shearForce[2]};
    sumC += norm3(v);
  }
  sumC *= 2.0;
  return sumF / sumC;
}
// The servo's completion: stopAtIter = iter + 1 (sdmcpu_step returns after this step).
void servo_complete(sdm_ctx* c) {
  c->servo.completedIter = c->iter;
  c->servo.completedNow = true;
}
// generalConsolidation with NEW_SCHEME (:324-353)
void servo_general(sdm_ctx* c, double dt) {
  auto& S = c->servo;
  S.iterate_num += 1;
  S.solve_num += 1;
  servo_move(c, dt);
  if (S.iterate_num >= S.p.iterate_num_max) {
    S.iterate_num = 0;
    servo_get_gain(c);
  }
  if (S.solve_num >= S.p.solve_num_max) {
    S.solve_num = 0;
    servo_get_stress(c, S.mean, S.dev);
    if ((std::abs(S.mean - S.p.goal[0]) < S.p.goal[0] * S.p.f_threshold) && (S.dev / S.mean < S.p.fmin_threshold)) {
      S.unbf = unbalanced_force(c);
      if (S.unbf < S.p.unbf_tol) {  // and newton->quiet_system_flag = true
        servo_complete(c);
        S.quiet = true;
      }
    }
  }
}
// hydroConsolidation (:264-322): each wall below its goal moves in by the smaller of a tenth
// of its gain times the deficit and the strain-rate bound; the check uses the UnbalancedForce
// of the last echo.  The wall stresses here are locals of the engine (left_wall_s ...), not the
// exported wsxx_left ...
void servo_hydro(sdm_ctx* c, double dt) {
  auto& S = c->servo;
  auto& B = c->bodies;
  servo_box(c);
  const double s[6] = {B[0].force[0] / S.x_area, B[1].force[0] / S.x_area, B[2].force[1] / S.y_area,
                       B[3].force[1] / S.y_area, B[4].force[2] / S.z_area, B[5].force[2] / S.z_area};
  const double size[3] = {S.width, S.depth, S.height};
  const double rate = S.p.hydro_strain_rate;
  for (int w = 0; w < 6; w++) {
    const double delta = S.p.goal[w / 2] - std::abs(s[w]);
    if (delta > 0) {
      const double step = std::min(0.1 * S.gain[w] * delta, rate * size[w / 2] * dt);
      if (w % 2 == 0) B[w].pos[w / 2] += step;
      else B[w].pos[w / 2] -= step;
    }
  }
  S.iterate_num += 1;
  if (S.iterate_num >= S.p.iterate_num_max) {
    S.iterate_num = 1;
    servo_get_gain(c);
  }
  S.solve_num += 1;
  if (S.solve_num >= S.p.solve_num_max) {
    S.solve_num = 1;
    servo_get_stress(c, S.mean, S.dev);
    if ((std::abs(S.mean - S.p.goal[0]) / S.p.goal[0] < S.p.f_threshold) && (S.dev / S.mean < S.p.fmin_threshold)) {
      if (S.unbf < S.p.unbf_tol) servo_complete(c);
    }
  }
}
// action_cubic's shear branch (:498-603): the z walls close at the ramped axial rate, then at
// goalz; the lateral walls keep the stress (shearMode bit 1, servo without z) or the volume.
void servo_shear(sdm_ctx* c, double dt) {
  auto& S = c->servo;
  auto& B = c->bodies;
  const double goalz = S.p.goal[2];
  if (!S.p.continue_flag && S.rampNum < S.p.ramp_interval) {
    // (Flag_ForceTarget, which would re-take the sizes at the ramp's start, is only set by the
    // consolidation scheme the reference compiles out)
    S.rampNum++;
    const double ramp_inter = 0.5 * (double)S.p.ramp_interval;
    const double chunk = ramp_inter / (double)S.p.ramp_chunks;
    double vel;
    if ((double)S.rampNum < ramp_inter)
      vel = (double)((float)((double)S.rampNum / chunk + 1) / (float)S.p.ramp_chunks) * goalz * 2.0;
    else
      vel = ((double)((float)((double)(S.p.ramp_interval - S.rampNum) / chunk + 1) / (float)S.p.ramp_chunks) + 1.0) *
            goalz;
    B[4].pos[2] += vel * dt * S.height;
    B[5].pos[2] += -vel * dt * S.height;
  } else {
    if (S.p.shear_mode & 1) {
      S.iterate_num += 1;
      if (S.iterate_num >= S.p.iterate_num_max) {
        S.iterate_num = 1;
        servo_get_gain(c);
      }
    }
    B[4].pos[2] += goalz * dt * S.height;
    B[5].pos[2] += -goalz * dt * S.height;
  }
  if (S.p.shear_mode & 1) {
    servo_move(c, dt);  // z_servo is off: the x and y walls only
  } else {              // constant volume
    S.width = B[1].pos[0] - B[0].pos[0];
    S.depth = B[3].pos[1] - B[2].pos[1];
    S.height = B[5].pos[2] - B[4].pos[2];
    const double area = S.initVol / S.height;
    const double dd = 0.5 * (area / S.width - S.depth);
    const double dw = area / (S.depth + dd) - S.width;
    B[0].pos[0] += -0.5 * dw;
    B[1].pos[0] += 0.5 * dw;
    B[2].pos[1] += -0.5 * dd;
    B[3].pos[1] += 0.5 * dd;
  }
  if (std::log(S.height0 / S.height) > S.p.target_strain) servo_complete(c);
}
// action_cubic (:397-604): the first run (the parts the servo uses), the branch, the echo.
void servo_action(sdm_ctx* c) {
  auto& S = c->servo;
  S.completedNow = false;
  if (S.firstRun) {
    servo_box(c);
    if (!S.p.continue_flag) {
      S.width0 = S.width;
      S.depth0 = S.depth;
      S.height0 = S.height;
    }
    S.initVol = S.height0 * S.width0 * S.depth0;
    S.rampNum = 0;
    S.firstRun = false;
    S.iterate_num = 0;
    S.solve_num = 0;
    servo_get_gain(c);
  }
  const double dt = c->prm.dt;
  const bool echo = c->iter % S.p.echo_interval == 0;
  if (S.p.z_servo) {
    if (S.p.hydro_strain) servo_hydro(c, dt);
    else servo_general(c, dt);
    if (echo) {
      S.unbf = unbalanced_force(c);
      servo_consol_ss(c);
    }
  } else {
    servo_shear(c, dt);
    if (echo) {
      S.unbf = unbalanced_force(c);
      if (S.p.shear_mode & 2) servo_consol_ss(c);
    }
  }
}

// Scene::moveToNextTimeStep (core/Scene.cpp:70-138) with the canonical engine list.
void one_step(sdm_ctx* c) {
  for (auto& b : c->bodies) {  // ForceResetter (pkg/common/ForceResetter.hpp:9-12)
    set3(b.force, 0, 0, 0);
    set3(b.torque, 0, 0, 0);
  }
  if (collider_is_activated(c)) collider_action(c);
  interaction_loop(c);
  if (c->servo.on) servo_action(c);
  newton(c);
  c->servo.quiet = false;
  c->iter++;
  c->time += c->prm.dt;
  c->cnt.steps++;
}

// ---------------------------------------------------------------------------------------
// Attribute table
// ---------------------------------------------------------------------------------------
enum Scope { S_BODY, S_PAIR, S_SCALAR, S_BOUNDS };
struct Field {
  const char* name;
  Scope scope;
  int dtype;
  int comps;
  size_t off;
  bool writable;
};
#define BF(n, member, dt, k, w) {n, S_BODY, dt, k, offsetof(Body, member), w}
#define PF(n, member, dt, k, w) {n, S_PAIR, dt, k, offsetof(Pair, member), w}
const Field kFields[] = {
    BF("state.pos", pos, SDM_F64, 3, true), BF("state.ori", ori, SDM_F64, 4, true),
    BF("state.vel", vel, SDM_F64, 3, true), BF("state.angVel", angVel, SDM_F64, 3, true),
    BF("state.angMom", angMom, SDM_F64, 3, true), BF("state.mass", mass, SDM_F64, 1, true),
    BF("state.inertia", inertia, SDM_F64, 3, true), BF("state.blockedDOFs", blockedDOFs, SDM_I32, 1, true),
    BF("state.isDamped", isDamped, SDM_I32, 1, true), BF("shape.kind", kind, SDM_I32, 1, true),
    BF("shape.rxyz", rxyz, SDM_F64, 3, true), BF("shape.eps", eps, SDM_F64, 2, true),
    BF("shape.r_max", r_max, SDM_F64, 1, true), BF("shape.isSphere", isSphere, SDM_I32, 1, true),
    BF("shape.wallAxis", wallAxis, SDM_I32, 1, true), BF("shape.wallSense", wallSense, SDM_I32, 1, true),
    BF("shape.rot_mat2local", rot_local, SDM_F64, 9, true), BF("shape.rot_mat2global", rot_global, SDM_F64, 9, true),
    BF("body.material", material, SDM_I32, 1, true), BF("body.specimen", specimen, SDM_I32, 1, true), BF("bound.exists", hasBound, SDM_I32, 1, true),
    BF("bound.min", bmin, SDM_F64, 3, true), BF("bound.max", bmax, SDM_F64, 3, true),
    BF("bound.refPos", refPos, SDM_F64, 3, true), BF("bound.sweepLength", sweepLength, SDM_F64, 1, true),
    BF("bound.lastUpdateIter", lastUpdateIter, SDM_I32, 1, true), BF("forces.force", force, SDM_F64, 3, false),
    BF("forces.torque", torque, SDM_F64, 3, false),
    PF("intr.id1", id1, SDM_I32, 1, true), PF("intr.id2", id2, SDM_I32, 1, true),
    PF("intr.hasGeom", hasGeom, SDM_I32, 1, true), PF("intr.hasPhys", hasPhys, SDM_I32, 1, true),
    PF("intr.functorSet", functorSet, SDM_I32, 1, true), PF("intr.geomExists", geomExists, SDM_I32, 1, true),
    PF("intr.iterMadeReal", iterMadeReal, SDM_I64, 1, true), PF("intr.cellDist", cellDist, SDM_I32, 3, true),
    PF("geom.PenetrationDepth", depth, SDM_F64, 1, true), PF("geom.contactAngle", contactAngle, SDM_F64, 2, true),
    PF("geom.contactPoint", contactPoint, SDM_F64, 3, true), PF("geom.shearInc", shearInc, SDM_F64, 3, true),
    PF("geom.relativeVn", relVn, SDM_F64, 3, true), PF("geom.relativeVs", relVs, SDM_F64, 3, true),
    PF("geom.normal", normal, SDM_F64, 3, true), PF("geom.twist_axis", twist, SDM_F64, 3, true),
    PF("geom.orthonormal_axis", orth, SDM_F64, 3, true), PF("geom.point1", point1, SDM_F64, 3, true),
    PF("geom.point2", point2, SDM_F64, 3, true), PF("geom.isShearNew", isShearNew, SDM_I32, 1, true),
    PF("geom.curvatures1", curv1, SDM_F64, 2, true), PF("geom.curvatures2", curv2, SDM_F64, 2, true),
    PF("geom.curvatureAngle", curvAngle, SDM_F64, 1, true), PF("geom.alpha", alpha, SDM_F64, 1, true),
    PF("geom.K", K, SDM_F64, 1, true), PF("geom.E", E, SDM_F64, 1, true),
    PF("geom.curvatures_sum", curvSum, SDM_F64, 1, true), PF("geom.isSphere", geomIsSphere, SDM_I32, 1, true),
    PF("phys.kn", kn, SDM_F64, 1, true), PF("phys.ks", ks, SDM_F64, 1, true),
    PF("phys.tangensOfFrictionAngle", tanPhi, SDM_F64, 1, true), PF("phys.betan", betan, SDM_F64, 1, true),
    PF("phys.betas", betas, SDM_F64, 1, true), PF("phys.normalForce", normalForce, SDM_F64, 3, true),
    PF("phys.shearForce", shearForce, SDM_F64, 3, true), PF("phys.nuab", nuab, SDM_F64, 1, true),
    PF("phys.Gab", Gab, SDM_F64, 1, true), PF("phys.mu1", mu1, SDM_F64, 1, true),
    PF("phys.mu2", mu2, SDM_F64, 1, true), PF("contact.force", F, SDM_F64, 3, false),
    PF("contact.torque1", T1, SDM_F64, 3, false), PF("contact.torque2", T2, SDM_F64, 3, false),
    PF("contact.active", contributed, SDM_I32, 1, false), PF("diag.nEval", nEval, SDM_I32, 1, false),
    PF("diag.stop", stopCode, SDM_I32, 1, false),
    {"scene.iter", S_SCALAR, SDM_I64, 1, 0, true}, {"scene.time", S_SCALAR, SDM_F64, 1, 1, true},
    {"collider.strideActive", S_SCALAR, SDM_I32, 1, 2, true},
    {"collider.fastestBodyMaxDist", S_SCALAR, SDM_F64, 1, 3, true},
    {"newton.maxVelocitySq", S_SCALAR, SDM_F64, 1, 4, true}, {"collider.numReinit", S_SCALAR, SDM_I32, 1, 5, true},
    {"collider.bounds", S_BOUNDS, SDM_I32, 1, 0, true}, {"collider.periods", S_BOUNDS, SDM_I32, 1, 1, true},
    {"collider.loIdx", S_SCALAR, SDM_I64, 3, 9, true},
    {"scene.dt", S_SCALAR, SDM_F64, 1, 6, true}, {"newton.damping", S_SCALAR, SDM_F64, 1, 7, true},
    {"newton.gravity", S_SCALAR, SDM_F64, 3, 8, true},
};
#undef BF
#undef PF

const Field* find_field(const char* name) {
  for (const Field& f : kFields)
    if (std::strcmp(f.name, name) == 0) return &f;
  return nullptr;
}
size_t dtype_size(int dt) { return dt == SDM_F64 ? 8 : (dt == SDM_I32 ? 4 : 8); }

int64_t field_count(sdm_ctx* c, const Field* f) {
  switch (f->scope) {
    case S_BODY: return (int64_t)c->bodies.size() * f->comps;
    case S_PAIR: return (int64_t)c->intrs.size() * f->comps;
    case S_SCALAR: return f->comps;
    case S_BOUNDS: return (int64_t)(c->BB[0].size() * 3);
  }
  return 0;
}

int fail(sdm_ctx* c, const std::string& msg) {
  if (c) c->err = msg;
  g_err = msg;
  return -1;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

const char* sdmcpu_version(void) { return "sdm-oracle 1.0 (FP64 CPU restatement)"; }

sdm_ctx* sdmcpu_create(const sdm_params* params, int32_t n_bodies, int32_t n_materials) {
  if (!params || n_bodies < 0 || n_materials < 0) {
    g_err = "sdmcpu_create: invalid arguments";
    return nullptr;
  }
  if (!(params->verlet_dist >= 0)) {
    g_err = "verletDist must be set explicitly (>= 0): the automatic value needs spheres";
    return nullptr;
  }
  if (params->law != SDM_LAW_LINEAR && params->law != SDM_LAW_HERTZ) {
    g_err = "sdmcpu_create: unknown law";
    return nullptr;
  }
  if (params->precision != SDM_FP64 && params->precision != SDM_MIXED) {
    g_err = "sdmcpu_create: unknown precision";
    return nullptr;
  }
  if (params->sort_axis < 0 || params->sort_axis > 2) {
    g_err = "sdmcpu_create: sortAxis must be 0, 1 or 2";
    return nullptr;
  }
  if (params->periodic && !(params->cell[0] > 0 && params->cell[1] > 0 && params->cell[2] > 0 &&
                            std::isfinite(params->cell[0] + params->cell[1] + params->cell[2]))) {
    g_err = "sdmcpu_create: a periodic cell needs three positive finite sizes";
    return nullptr;
  }
  sdm_ctx* c = new sdm_ctx();
  c->prm = *params;
  c->bodies.resize(n_bodies);
  for (auto& b : c->bodies) {
    std::memset(&b, 0, sizeof(Body));
    b.ori[0] = 1;
    b.isDamped = 1;
    b.refPos[0] = b.refPos[1] = b.refPos[2] = NaN;
    b.bmin[0] = b.bmin[1] = b.bmin[2] = NaN;
    b.bmax[0] = b.bmax[1] = b.bmax[2] = NaN;
    b.rot_local[0] = b.rot_local[4] = b.rot_local[8] = 1;
    b.rot_global[0] = b.rot_global[4] = b.rot_global[8] = 1;
  }
  c->mats.resize(n_materials);
  std::memset(&c->cnt, 0, sizeof(c->cnt));
  return c;
}

void sdmcpu_destroy(sdm_ctx* c) { delete c; }

const char* sdmcpu_last_error(const sdm_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int sdmcpu_set_materials(sdm_ctx* c, const sdm_material* mats, int32_t n) {
  if (!c || (n > 0 && !mats)) return fail(c, "sdmcpu_set_materials: invalid arguments");
  c->mats.assign(mats, mats + n);
  return 0;
}

int sdmcpu_superquadric_props(double rx, double ry, double rz, double eps1, double eps2, double* r_max,
                              double* volume, double inertia[3]) {
  sq_initial(rx, ry, rz, eps1, eps2, r_max, volume, inertia);
  return 0;
}

int sdmcpu_superquadric_props_n(int64_t n, const double* rxyz, const double* eps, double* r_max, double* volume,
                                double* inertia) {
  if (n < 0 || (n > 0 && (!rxyz || !eps || !r_max || !volume || !inertia)))
    return fail(nullptr, "sdm_superquadric_props_n: bad arguments");
  for (int64_t i = 0; i < n; i++)
    sq_initial(rxyz[3 * i], rxyz[3 * i + 1], rxyz[3 * i + 2], eps[2 * i], eps[2 * i + 1], &r_max[i], &volume[i],
               &inertia[3 * i]);
  return 0;
}

const char* sdmcpu_field_list(void) {
  static std::string s;
  if (s.empty())
    for (const Field& f : kFields) {
      s += f.name;
      s += "\n";
    }
  return s.c_str();
}

int sdmcpu_field_info(sdm_ctx* c, const char* name, int32_t* dtype, int64_t* count) {
  const Field* f = find_field(name);
  if (!c || !f) return fail(c, std::string("unknown attribute ") + (name ? name : "(null)"));
  if (dtype) *dtype = f->dtype;
  if (count) *count = field_count(c, f);
  return 0;
}

int sdmcpu_get(sdm_ctx* c, const char* name, void* dst, int64_t count) {
  const Field* f = find_field(name);
  if (!c || !f || !dst) return fail(c, std::string("sdmcpu_get: bad attribute ") + (name ? name : "(null)"));
  int64_t n = field_count(c, f);
  if (count != n) return fail(c, std::string("sdmcpu_get: count mismatch for ") + name);
  const size_t es = dtype_size(f->dtype);
  char* out = (char*)dst;
  if (f->scope == S_BODY) {
    for (size_t i = 0; i < c->bodies.size(); i++)
      std::memcpy(out + i * es * f->comps, (const char*)&c->bodies[i] + f->off, es * f->comps);
  } else if (f->scope == S_PAIR) {
    for (size_t i = 0; i < c->intrs.size(); i++)
      std::memcpy(out + i * es * f->comps, (const char*)&c->intrs[i] + f->off, es * f->comps);
  } else if (f->scope == S_SCALAR) {
    switch (f->off) {
      case 0: *(int64_t*)dst = c->iter; break;
      case 1: *(double*)dst = c->time; break;
      case 2: *(int32_t*)dst = c->strideActive; break;
      case 3: *(double*)dst = c->fastestBodyMaxDist; break;
      case 4: *(double*)dst = c->maxVelocitySq; break;
      case 5: *(int32_t*)dst = c->numReinit; break;
      case 6: *(double*)dst = c->prm.dt; break;
      case 7: *(double*)dst = c->prm.damping; break;
      case 8: std::memcpy(dst, c->prm.gravity, 3 * sizeof(double)); break;
      case 9:
        for (int a = 0; a < 3; a++) ((int64_t*)dst)[a] = c->loIdx[a];
        break;
    }
  } else {
    int32_t* o = (int32_t*)dst;
    size_t m = c->BB[0].size();
    for (int a = 0; a < 3; a++)
      for (size_t i = 0; i < m; i++)
        o[a * m + i] = f->off == 0 ? c->BB[a][i].id * 2 + (c->BB[a][i].isMin ? 0 : 1) : c->BB[a][i].period;
  }
  return 0;
}

int sdmcpu_set(sdm_ctx* c, const char* name, const void* src, int64_t count) {
  const Field* f = find_field(name);
  if (!c || !f || (!src && count > 0)) return fail(c, std::string("sdmcpu_set: bad attribute ") + (name ? name : "(null)"));
  if (!f->writable) return fail(c, std::string("sdmcpu_set: read-only attribute ") + name);
  if (count < 0) return fail(c, std::string("sdmcpu_set: negative count for ") + name);
  const size_t es = dtype_size(f->dtype);
  const char* in = (const char*)src;
  if (f->scope == S_PAIR && std::strcmp(name, "intr.id1") == 0) {
    // setting id1 defines the container size; new entries are fresh potential interactions
    c->intrs.resize((size_t)count, new_pair(-1, -1));
    c->found_dirty = true;
  }
  int64_t n = field_count(c, f);
  if (f->scope == S_BOUNDS && f->off == 1) {
    // the periods of the current bound vectors
    if (count != (int64_t)(3 * c->BB[0].size())) return fail(c, "collider.periods: count must be 3 x the bound vector");
    const int32_t* h = (const int32_t*)src;
    const size_t m = c->BB[0].size();
    for (int a = 0; a < 3; a++)
      for (size_t i = 0; i < m; i++) c->BB[a][i].period = h[a * m + i];
    return 0;
  }
  if (f->scope == S_BOUNDS) {
    if (count % 3) return fail(c, "collider.bounds: count must be 3*2N");
    size_t m = (size_t)(count / 3);
    const int32_t* h = (const int32_t*)src;
    for (int a = 0; a < 3; a++) {
      c->BB[a].resize(m);
      for (size_t i = 0; i < m; i++) {
        int32_t v = h[a * m + i];
        c->BB[a][i] = Bounds{0, v >> 1, (v & 1) ? 0 : 1, 0};
      }
    }
    return 0;
  }
  if (count != n) return fail(c, std::string("sdmcpu_set: count mismatch for ") + name);
  if (f->scope == S_BODY) {
    for (size_t i = 0; i < c->bodies.size(); i++)
      std::memcpy((char*)&c->bodies[i] + f->off, in + i * es * f->comps, es * f->comps);
  } else if (f->scope == S_PAIR) {
    for (size_t i = 0; i < c->intrs.size(); i++)
      std::memcpy((char*)&c->intrs[i] + f->off, in + i * es * f->comps, es * f->comps);
    if (std::strcmp(name, "intr.id2") == 0 || std::strcmp(name, "intr.id1") == 0) c->found_dirty = true;
  } else {
    switch (f->off) {
      case 0: c->iter = *(const int64_t*)src; break;
      case 1: c->time = *(const double*)src; break;
      case 2: c->strideActive = *(const int32_t*)src; break;
      case 3: c->fastestBodyMaxDist = *(const double*)src; break;
      case 4: c->maxVelocitySq = *(const double*)src; break;
      case 5: c->numReinit = *(const int32_t*)src; break;
      case 6: c->prm.dt = *(const double*)src; break;
      case 7: c->prm.damping = *(const double*)src; break;
      case 8: std::memcpy(c->prm.gravity, src, 3 * sizeof(double)); break;
      case 9:
        for (int a = 0; a < 3; a++) {
          const int64_t v = ((const int64_t*)src)[a];
          if (v < 0 || (v > 0 && v >= (int64_t)c->BB[a].size())) return fail(c, "collider.loIdx: outside the bound vector");
          c->loIdx[a] = v;
        }
        break;
    }
  }
  return 0;
}

int sdmcpu_step(sdm_ctx* c, int32_t nsteps) {
  if (!c) return fail(c, "null context");
  for (auto& b : c->bodies) {
    if (b.material < 0 || b.material >= (int)c->mats.size()) return fail(c, "body material index out of range");
    if (b.specimen != 0) return fail(c, "the oracle steps one specimen (body.specimen must be 0)");
    if (c->prm.periodic && b.kind == SDM_BODY_WALL) return fail(c, "walls are not supported in a periodic cell");
  }
  if (c->servo.on) {
    if (c->prm.periodic) return fail(c, "the wall servo needs the six walls of a cubic specimen, not a periodic cell");
    if (c->bodies.size() < 6) return fail(c, "the wall servo needs bodies 0-5 to be the walls");
    for (int i = 0; i < 6; i++)
      if (c->bodies[i].kind != SDM_BODY_WALL) return fail(c, "the wall servo needs bodies 0-5 to be the walls");
  }
  for (int s = 0; s < nsteps; s++) {
    one_step(c);
    if (c->periErr) {
      c->periErr = false;
      return fail(c, "body larger than half of the cell size encountered");
    }
    if (c->servo.completedNow) break;  // stopAtIter = iter + 1
  }
  return 0;
}

int sdmcpu_set_servo(sdm_ctx* c, const sdm_servo* s) {
  if (!c) return -1;
  if (!s) {
    c->servo.on = false;
    return 0;
  }
  if (s->iterate_num_max < 1 || s->solve_num_max < 1 || s->echo_interval < 1)
    return fail(c, "sdm_set_servo: intervals must be >= 1");
  if (s->ramp_interval < 0 || s->ramp_chunks < 1 || s->shear_mode < 0)
    return fail(c, "sdm_set_servo: ramp_interval >= 0, ramp_chunks >= 1 and shear_mode >= 0 are needed");
  if (s->keep_state) {
    if (!c->servoEverSet) return fail(c, "sdm_set_servo: keep_state needs a servo that was set before");
  } else {
    c->servo = sdm_ctx::Servo();
    if (s->continue_flag) {
      c->servo.width0 = s->size0[0];
      c->servo.depth0 = s->size0[1];
      c->servo.height0 = s->size0[2];
    }
  }
  c->servo.p = *s;
  c->servo.on = true;
  c->servoEverSet = true;
  return 0;
}

int sdmcpu_get_servo(sdm_ctx* c, sdm_servo_state* o) {
  if (!c || !o) return fail(c, "bad arguments");
  const auto& S = c->servo;
  for (int k = 0; k < 6; k++) {
    o->wall_stress[k] = S.ws[k];
    o->gain[k] = S.gain[k];
    o->stiffness[k] = S.stiff[k];
  }
  o->stress[0] = S.wsxx;
  o->stress[1] = S.wsyy;
  o->stress[2] = S.wszz;
  o->size[0] = S.width;
  o->size[1] = S.depth;
  o->size[2] = S.height;
  o->mean_stress = S.mean;
  o->deviator_stress = S.dev;
  o->unbalanced_force = S.unbf;
  o->iterate_num = S.iterate_num;
  o->solve_num = S.solve_num;
  o->completed_iter = S.completedIter;
  o->size0[0] = S.width0;
  o->size0[1] = S.depth0;
  o->size0[2] = S.height0;
  o->ramp_num = S.rampNum;
  o->pad_ = 0;
  return 0;
}

int sdmcpu_sync(sdm_ctx*) { return 0; }

// Shop__meanCoordinationNumber (py/_utils.cpp:522-539).
int sdmcpu_mean_coordination(sdm_ctx* c, double* cn) {
  if (!c || !cn) return fail(c, "bad arguments");
  int num_contacts = 0, num_particles = 0;
  for (const Pair& I : c->intrs) {
    if (!(I.hasGeom && I.hasPhys)) continue;
    const double* f = I.normalForce;
    if (std::sqrt(f[0] * f[0] + (f[1] * f[1] + f[2] * f[2])) == 0) continue;  // normalForce.norm() == 0
    num_contacts += 1;
  }
  for (const Body& b : c->bodies) {
    if (b.kind == SDM_BODY_WALL) continue;
    num_particles += 1;
  }
  if (num_particles == 0) return fail(c, "sdm_mean_coordination: no particles (every body is a wall)");
  *cn = 2.0 * num_contacts / num_particles;
  return 0;
}

// Shop::getStress (pkg/dem/Shop_02.cpp:350-377), serial over the container.
int sdmcpu_stress(sdm_ctx* c, double volume, double sigma[9]) {
  if (!c || !sigma) return fail(c, "bad arguments");
  const bool isPeriodic = c->prm.periodic != 0;
  if (volume == 0) volume = isPeriodic ? c->prm.cell[0] * c->prm.cell[1] * c->prm.cell[2] : 1;
  double st[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (const Pair& I : c->intrs) {
    if (!(I.hasGeom && I.hasPhys)) continue;
    const Body& b1 = c->bodies[I.id1];
    const Body& b2 = c->bodies[I.id2];
    double branch[3], f[3];
    for (int k = 0; k < 3; k++) branch[k] = b1.pos[k] - b2.pos[k];
    if (isPeriodic) {
      double shift2[3];
      pair_shift2(c, I, shift2);
      for (int k = 0; k < 3; k++) branch[k] -= shift2[k];
    }
    for (int k = 0; k < 3; k++) f[k] = I.normalForce[k] + I.shearForce[k];
    for (int r = 0; r < 3; r++)
      for (int q = 0; q < 3; q++) st[3 * r + q] += f[r] * branch[q];
  }
  for (int k = 0; k < 9; k++) sigma[k] = st[k] / volume;
  return 0;
}

int sdmcpu_body_forces(sdm_ctx* c, const int32_t* ids, int32_t n, double* forces) {
  if (!c || n < 0 || (n > 0 && (!ids || !forces))) return fail(c, "bad arguments");
  for (int32_t i = 0; i < n; i++) {
    if (ids[i] < 0 || ids[i] >= (int32_t)c->bodies.size()) return fail(c, "body id out of range");
    for (int k = 0; k < 3; k++) forces[3 * i + k] = c->bodies[ids[i]].force[k];
  }
  return 0;
}

int sdmcpu_get_counters(sdm_ctx* c, sdm_counters* out) {
  if (!c || !out) return fail(c, "bad arguments");
  *out = c->cnt;
  return 0;
}
int sdmcpu_reset_counters(sdm_ctx* c) {
  if (!c) return -1;
  std::memset(&c->cnt, 0, sizeof(c->cnt));
  return 0;
}

int sdmcpu_set_timing(sdm_ctx* c, int32_t) { return c ? 0 : -1; }
int sdmcpu_set_graph(sdm_ctx* c, int32_t) { return c ? 0 : -1; }
// Launch shapes have no meaning on the CPU: accepted and reported as the device defaults.
int sdmcpu_set_tuning(sdm_ctx* c, const sdm_tuning* t) { return (c && t) ? 0 : -1; }
int sdmcpu_get_tuning(sdm_ctx* c, sdm_tuning* t) {
  if (!c || !t) return -1;
  t->solver_blocks_per_cu = 0;
  t->sort_by_cost = 1;
  t->misc_blocks_per_cu = 2;
  t->graph = 1;
  return 0;
}
int sdmcpu_get_timers(sdm_ctx* c, double out[8]) {
  if (!c || !out) return -1;
  for (int k = 0; k < 8; k++) out[k] = 0;
  return 0;
}

int sdmcpu_get_kernel_timers(sdm_ctx* c, double out[32]) {
  if (!c || !out) return -1;
  for (int k = 0; k < 32; k++) out[k] = 0;
  return 0;
}

int sdmcpu_set_threads(sdm_ctx* c, int32_t n) {
  if (!c || n < 1) return fail(c, "bad thread count");
#ifndef _OPENMP
  if (n != 1) return fail(c, "oracle built without OpenMP");
#endif
  c->nthreads = n;
  return 0;
}

// The collider's bound ring of one axis (coordinates, ids, min flags, periods) and loIdx.
int sdmcpu_dump_bounds(sdm_ctx* c, int32_t axis, double* coord, int32_t* id, int32_t* isMin, int32_t* period,
                       int64_t* loIdx) {
  if (!c || axis < 0 || axis > 2) return -1;
  const std::vector<Bounds>& v = c->BB[axis];
  for (size_t i = 0; i < v.size(); i++) {
    coord[i] = v[i].coord;
    id[i] = v[i].id;
    isMin[i] = v[i].isMin;
    period[i] = v[i].period;
  }
  *loIdx = c->loIdx[axis];
  return 0;
}

int sdmcpu_set_init_sort(sdm_ctx* c, int32_t stable) {
  if (!c) return -1;
  c->initSortStable = stable ? 1 : 0;
  return 0;
}

int sdmcpu_support_point(double rx, double ry, double rz, double eps1, double eps2, const double n_local[3],
                         double phi_out[2], double p_out[3]) {
  SQ s;
  s.rx = rx; s.ry = ry; s.rz = rz; s.eps1 = eps1; s.eps2 = eps2;
  normal2phi(s, n_local, phi_out);
  get_surface(s, phi_out, p_out);
  return 0;
}

int sdmcpu_curvatures(double rx, double ry, double rz, double eps1, double eps2, const double rot_global[9],
                      const double phi[2], double k_out[2], double dir_max[3], double dir_min[3], int32_t* max_flag) {
  SQ s;
  s.rx = rx; s.ry = ry; s.rz = rz; s.eps1 = eps1; s.eps2 = eps2;
  s.Rg = rot_global;
  s.Rl = rot_global;
  double ph[2] = {phi[0], phi[1]};
  int f = 0;
  curvatures(s, ph, k_out, dir_max, dir_min, &f);
  *max_flag = f;
  return 0;
}

int sdmcpu_hertz_mindlin(const double curv1[2], const double curv2[2], double theta, double out[4]) {
  hertz_mindlin(curv1, curv2, theta, out);
  return 0;
}

}  // extern "C"
