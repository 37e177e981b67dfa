// frozen.h -- differentiable rollout over a frozen contact set (the differentiation form).
//
// Over a window of K steps the contact set of the linear law and each contact's geometry
// (normal n from body a to body b, contact point, start depth) are those of the window's
// first step.  Per step, for every contact with depth > 0:
//   depth  = depth0 - n.((x_b - x_b0) - (x_a - x_a0))
//   Fn     = kn depth n
//   vt     = tangential part of (v_b + w_b x r_b) - (v_a + w_a x r_a)
//   Fs     = cap(Fs_prev - ks dt vt),  cap: |Fs| <= mu |Fn|  (SuperquadricsLaw::go,
//            pkg/dem/Superquadrics.cpp:1221-1274, with the normal frozen so `rotate` is the identity)
//   F      = -Fn - Fs;  body a gets F and r_a x F, body b gets -F and -r_b x F (:1351-1356)
// and a contact with depth <= 0 carries no force and forgets its shear force.  Bodies follow
// NewtonIntegrator's leapfrog with Cundall damping (pkg/dem/NewtonIntegrator.cpp:19-38, 189, 232,
// 315); rotation integrates the angular velocity with the world-frame inertia of the window
// start (orientation changes are not fed back into the frozen geometry).  Non-dynamic bodies
// (walls) move with their prescribed velocity.  Loss = mean over the last `loss_steps` steps of
// the normal-force magnitudes (kn depth) of the flagged contacts.
//
// The adjoint kernels below are the exact reverse of these steps (the Cundall factors, the
// depth > 0 test and the cap branch are held at their forward values, so the gradient is that of
// the piecewise-smooth map away from those transitions).  States are checkpointed every
// `checkpoint_every` steps; the backward pass recomputes one segment at a time.
#pragma once
#include "dev_math.h"
#include "../../include/sdm.h"

namespace fz {

__device__ __forceinline__ V3<double> ld(const double* a, int64_t i) { return mk3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
__device__ __forceinline__ void st(double* a, int64_t i, V3<double> v) {
  a[3 * i] = v.x;
  a[3 * i + 1] = v.y;
  a[3 * i + 2] = v.z;
}
__device__ __forceinline__ V3<double> mat_vec(const double* M, V3<double> v) {
  return mk3(M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z,
             M[6] * v.x + M[7] * v.y + M[8] * v.z);
}
__device__ __forceinline__ double sgn0(double a) { return a < 0 ? -1.0 : (a > 0 ? 1.0 : 0.0); }

// Per-block sum of doubles into c[0..K) (one atomic per value and block).
template <int K>
__device__ __forceinline__ void block_sum(double* c, const double* v) {
  __shared__ double part[K][16];
  double x[K];
  for (int k = 0; k < K; k++) {
    x[k] = v[k];
    for (int off = 32; off > 0; off >>= 1) x[k] += __shfl_down(x[k], off);
  }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < K; k++) part[k][threadIdx.x >> 6] = x[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < K; k++) {
      double s = 0;
      for (int w = 0; w < (int)((blockDim.x + 63) >> 6); w++) s += part[k][w];
      if (s != 0) atomicAdd(&c[k], s);
    }
  }
  __syncthreads();
}

struct Law {
  bool active, capped;
  double depth, fnMag, rho;
  V3<double> n, vt, fsTr, fsNew, shat, F;
};

// The contact law of one contact in one step (forward, and recomputed by the adjoint).
__device__ __forceinline__ Law contact_law(const sdm_frozen_problem& P, int64_t c, const double* par,
                                           const double* x, const double* v, const double* w, const double* fsIn) {
  Law L;
  const int a = P.id_a[c], b = P.id_b[c];
  L.n = ld(P.normal, c);
  const V3<double> ua = ld(x, a) - ld(P.pos0, a), ub = ld(x, b) - ld(P.pos0, b);
  L.depth = P.depth0[c] - dot(L.n, ub - ua);
  L.active = L.depth > 0;
  L.capped = false;
  L.fnMag = 0;
  L.rho = 1;
  L.F = mk3(0.0, 0.0, 0.0);
  L.fsNew = L.F;
  L.fsTr = L.F;
  L.vt = L.F;
  L.shat = L.F;
  if (!L.active) return L;
  const double kn = par[0], ks = par[1], mu = par[2];
  const V3<double> ra = ld(P.ra, c), rb = ld(P.rb, c);
  const V3<double> vrel = (ld(v, b) + cross(ld(w, b), rb)) - (ld(v, a) + cross(ld(w, a), ra));
  L.vt = vrel - dot(L.n, vrel) * L.n;
  L.fnMag = kn * L.depth;
  L.fsTr = ld(fsIn, c) - (ks * P.dt) * L.vt;
  const double ft = norm(L.fsTr);
  if (ft > mu * L.fnMag && ft > 0) {
    L.capped = true;
    L.rho = mu * L.fnMag / ft;
    L.shat = (1.0 / ft) * L.fsTr;
    L.fsNew = L.rho * L.fsTr;
  } else {
    L.fsNew = L.fsTr;
  }
  L.F = -(L.fnMag * L.n) - L.fsNew;
  return L;
}

__device__ __forceinline__ bool in_loss_window(const sdm_frozen_problem& P, int t) {
  return t >= P.steps - P.loss_steps;
}

__global__ void k_contacts(sdm_frozen_problem P, const double* par, const double* x, const double* v, const double* w,
                           const double* fsIn, double* fsOut, double* fc, int t, double* loss) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0;
  if (c < P.n_contacts) {
    const Law L = contact_law(P, c, par, x, v, w, fsIn);
    st(fc, c, L.F);
    st(fsOut, c, L.fsNew);
    if (L.active && P.in_loss[c] && in_loss_window(P, t)) l = L.fnMag / P.loss_steps;
  }
  block_sum<1>(loss, &l);
}

// Force and torque on body i from its contacts, in list order.
__device__ __forceinline__ void gather(const sdm_frozen_problem& P, int64_t i, const double* fc, V3<double>& F,
                                       V3<double>& T) {
  F = mk3(0.0, 0.0, 0.0);
  T = F;
  for (int e = P.csr_off[i]; e < P.csr_off[i + 1]; e++) {
    const int32_t ent = P.csr_ent[e];
    const int64_t c = ent >> 1;
    const V3<double> f = ld(fc, c);
    if ((ent & 1) == 0) {
      F = F + f;
      T = T + cross(ld(P.ra, c), f);
    } else {
      F = F - f;
      T = T - cross(ld(P.rb, c), f);
    }
  }
}

// Cundall factors 1 - d sgn(a (v + dt/2 a)) and 1 - d sgn(T w).
__device__ __forceinline__ void newton_factors(const sdm_frozen_problem& P, int64_t i, V3<double> F, V3<double> T,
                                               V3<double> vi, V3<double> wi, V3<double>& a, V3<double>& ca,
                                               V3<double>& cT) {
  const double m = P.mass[i];
  a = mk3(F.x / m + P.gravity[0], F.y / m + P.gravity[1], F.z / m + P.gravity[2]);
  const double d = P.damping, h = 0.5 * P.dt;
  ca = mk3(1 - d * sgn0(a.x * (vi.x + h * a.x)), 1 - d * sgn0(a.y * (vi.y + h * a.y)),
           1 - d * sgn0(a.z * (vi.z + h * a.z)));
  cT = mk3(1 - d * sgn0(T.x * wi.x), 1 - d * sgn0(T.y * wi.y), 1 - d * sgn0(T.z * wi.z));
}

__global__ void k_newton(sdm_frozen_problem P, const double* x, const double* v, const double* w, const double* fc,
                         double* x1, double* v1, double* w1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n_bodies) return;
  const V3<double> xi = ld(x, i), vi = ld(v, i), wi = ld(w, i);
  if (!P.dynamic[i]) {
    st(v1, i, vi);
    st(x1, i, xi + P.dt * vi);
    st(w1, i, wi);
    return;
  }
  V3<double> F, T, a, ca, cT;
  gather(P, i, fc, F, T);
  newton_factors(P, i, F, T, vi, wi, a, ca, cT);
  const V3<double> vn = vi + P.dt * mk3(ca.x * a.x, ca.y * a.y, ca.z * a.z);
  st(v1, i, vn);
  st(x1, i, xi + P.dt * vn);
  st(w1, i, wi + P.dt * mat_vec(P.inv_inertia + 9 * i, mk3(cT.x * T.x, cT.y * T.y, cT.z * T.z)));
}

// ---- adjoint ----
// Reverse of k_newton for the dynamic bodies: in: adjoints of (x, v, w) after the step;
// out: the identity parts of the adjoints before it, and the adjoints of each body's force
// and torque.
__global__ void k_back_newton(sdm_frozen_problem P, const double* x, const double* v, const double* w,
                              const double* fc, double* lx, double* lv, double* lw, double* lF, double* lT) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n_bodies) return;
  if (!P.dynamic[i]) {
    st(lF, i, mk3(0.0, 0.0, 0.0));
    st(lT, i, mk3(0.0, 0.0, 0.0));
    st(lx, i, mk3(0.0, 0.0, 0.0));
    st(lv, i, mk3(0.0, 0.0, 0.0));
    st(lw, i, mk3(0.0, 0.0, 0.0));
    return;
  }
  const V3<double> vi = ld(v, i), wi = ld(w, i);
  V3<double> F, T, a, ca, cT;
  gather(P, i, fc, F, T);
  newton_factors(P, i, F, T, vi, wi, a, ca, cT);
  // x1 = x + dt v1, v1 = v + dt ca a, w1 = w + dt J^-1 (cT T)
  const V3<double> lv1 = ld(lv, i) + P.dt * ld(lx, i);
  const double m = P.mass[i];
  st(lF, i, mk3(P.dt * ca.x * lv1.x / m, P.dt * ca.y * lv1.y / m, P.dt * ca.z * lv1.z / m));
  const double* J = P.inv_inertia + 9 * i;  // symmetric
  const V3<double> jl = mat_vec(J, ld(lw, i));
  st(lT, i, mk3(P.dt * cT.x * jl.x, P.dt * cT.y * jl.y, P.dt * cT.z * jl.z));
  st(lv, i, lv1);
  // lx and lw: identity parts unchanged
}

// Reverse of the contact law: adjoints of the contact's inputs (positions, velocities, angular
// velocities of both bodies, written per contact), of the previous shear force (lfs, in: the
// adjoint of the shear force the step leaves; out: of the one it received) and of (kn, ks, mu).
__global__ void k_back_contacts(sdm_frozen_problem P, const double* par, const double* x, const double* v,
                                const double* w, const double* fsIn, const double* lF, const double* lT, double* lfs,
                                double* contrib, int t, double* gpar) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double g[3] = {0, 0, 0};
  if (c < P.n_contacts) {
    const Law L = contact_law(P, c, par, x, v, w, fsIn);
    double* cb = contrib + 18 * c;
    for (int k = 0; k < 18; k++) cb[k] = 0;
    if (!L.active) {
      st(lfs, c, mk3(0.0, 0.0, 0.0));
    } else {
      const double kn = par[0], ks = par[1], mu = par[2];
      const int a = P.id_a[c], b = P.id_b[c];
      const V3<double> ra = ld(P.ra, c), rb = ld(P.rb, c);
      const V3<double> lFc = (ld(lF, a) - ld(lF, b)) + (cross(ld(lT, a), ra) - cross(ld(lT, b), rb));
      const V3<double> lFn = -lFc;
      const V3<double> lfsNew = ld(lfs, c) - lFc;
      V3<double> lfsTr;
      double lFnMag = 0;
      if (L.capped) {
        const double s = dot(L.shat, lfsNew);
        lfsTr = L.rho * (lfsNew - s * L.shat);
        g[2] += L.fnMag * s;
        lFnMag = mu * s;
      } else {
        lfsTr = lfsNew;
      }
      st(lfs, c, lfsTr);
      g[1] += -P.dt * dot(L.vt, lfsTr);
      const V3<double> lvt = (-ks * P.dt) * lfsTr;
      const V3<double> lvrel = lvt - dot(L.n, lvt) * L.n;
      double ldepth = kn * dot(L.n, lFn) + kn * lFnMag;
      g[0] += L.depth * dot(L.n, lFn) + L.depth * lFnMag;
      if (P.in_loss[c] && in_loss_window(P, t)) {
        ldepth += kn / P.loss_steps;
        g[0] += L.depth / P.loss_steps;
      }
      // depth = depth0 - n.((x_b - x_b0) - (x_a - x_a0)); vrel = (v_b + w_b x r_b) - (v_a + w_a x r_a)
      const V3<double> lxa = ldepth * L.n, lxb = -(ldepth * L.n);
      const V3<double> lva = -lvrel, lvb = lvrel;
      const V3<double> lwa = -cross(ra, lvrel), lwb = cross(rb, lvrel);
      const V3<double> q[6] = {lxa, lva, lwa, lxb, lvb, lwb};
      for (int k = 0; k < 6; k++) {
        cb[3 * k] = q[k].x;
        cb[3 * k + 1] = q[k].y;
        cb[3 * k + 2] = q[k].z;
      }
    }
  }
  block_sum<3>(gpar, g);
}

// Adjoints of each dynamic body's x, v, w: add its contacts' contributions in list order.
__global__ void k_back_gather(sdm_frozen_problem P, const double* contrib, double* lx, double* lv, double* lw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n_bodies || !P.dynamic[i]) return;
  V3<double> ax = ld(lx, i), av = ld(lv, i), aw = ld(lw, i);
  for (int e = P.csr_off[i]; e < P.csr_off[i + 1]; e++) {
    const int32_t ent = P.csr_ent[e];
    const double* cb = contrib + 18 * (int64_t)(ent >> 1) + ((ent & 1) ? 9 : 0);
    ax = ax + mk3(cb[0], cb[1], cb[2]);
    av = av + mk3(cb[3], cb[4], cb[5]);
    aw = aw + mk3(cb[6], cb[7], cb[8]);
  }
  st(lx, i, ax);
  st(lv, i, av);
  st(lw, i, aw);
}

}  // namespace fz
