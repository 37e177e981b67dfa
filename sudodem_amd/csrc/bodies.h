// bodies.h -- per-body kernels: bound update (Bo1 functors + adaptive sweep) and the
// Newton integrator with the per-body force reduction.
#pragma once
#include "pairs.h"

// BoundDispatcher::processBody (pkg/common/Dispatching.cpp:30-62) with
// Bo1_Superquadrics_Aabb::go (pkg/dem/Superquadrics.cpp:639-686: an orientation-free cube of
// half-size r_max) and Bo1_Wall_Aabb::go (pkg/common/Wall.cpp:54-62).
__global__ __launch_bounds__(256) void k_bounds(BodyArrays B, int64_t iter, double sweepDist, double minSweepDistFactor,
                                                int targetInterv) {
  const int64_t N = B.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const V3<double> pos = ld3(B.pos, i, N);
  double sweep = B.sweep[i];
  const bool had = B.hasBound[i] != 0;
  if (had) {
    if (targetInterv >= 0) {
      const V3<double> r = ld3(B.refPos, i, N);
      const V3<double> disp = pos - r;
      // std::max(a,b) = (a<b)?b:a, which keeps a when b is NaN (first call: refPos is NaN)
      const double ax = fabs(disp.x), ay = fabs(disp.y), az = fabs(disp.z);
      const double myz = (ay < az) ? az : ay;
      const double dist = (ax < myz) ? myz : ax;
      if (dist != 0.0) {  // NaN counts as true
        double newLength = dist * (double)targetInterv / (double)(iter - (int64_t)B.lastUpd[i]);
        const double a = 0.9 * sweep;
        newLength = (a < newLength) ? newLength : a;
        const double m = (sweepDist < newLength) ? sweepDist : newLength;
        const double lo = minSweepDistFactor * sweepDist;
        sweep = (lo < m) ? m : lo;
      } else {
        sweep = 0;
      }
    } else {
      sweep = sweepDist;
    }
  } else {
    sweep = 0;  // the functor allocates the bound: Bound::sweepLength defaults to 0
  }
  V3<double> mn, mx;
  if (B.kind[i] == SDM_KIND_SQ) {
    const double r = B.rmax[i];
    mn = pos + mk3(-r, -r, -r);
    mx = pos + mk3(r, r, r);
  } else {
    const double inf = __builtin_huge_val();
    mn = mk3(-inf, -inf, -inf);
    mx = mk3(inf, inf, inf);
    const int ax = B.wallAxis[i];
    if (ax == 0) { mn.x = pos.x; mx.x = pos.x; }
    else if (ax == 1) { mn.y = pos.y; mx.y = pos.y; }
    else { mn.z = pos.z; mx.z = pos.z; }
  }
  if (sweep > 0) {
    mn = mn - mk3(sweep, sweep, sweep);
    mx = mx + mk3(sweep, sweep, sweep);
  }
  st3(B.bmin, i, N, mn);
  st3(B.bmax, i, N, mx);
  double* gr = B.geo + i * GEO_STRIDE;
  gr[GEO_BMIN] = mn.x; gr[GEO_BMIN + 1] = mn.y; gr[GEO_BMIN + 2] = mn.z;
  gr[GEO_BMAX] = mx.x; gr[GEO_BMAX + 1] = mx.y; gr[GEO_BMAX + 2] = mx.z;
  st3(B.refPos, i, N, pos);
  B.sweep[i] = sweep;
  B.lastUpd[i] = (int32_t)iter;
  B.hasBound[i] = 1;
  B.disp[i] = 0.0f;
}

// |d| rounded up to float (NaN stays NaN: no bound).
__device__ __forceinline__ float disp_up(V3<double> d) { return __double2float_ru(norm(d)); }

// Rebuild the packed records from the attribute arrays (after the host wrote attributes).
__global__ __launch_bounds__(256) void k_geo_pack(BodyArrays B) {
  const int64_t N = B.n;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double* g = B.geo + i * GEO_STRIDE;
  for (int k = 0; k < 3; k++) g[GEO_POS + k] = B.pos[k * N + i];
  g[GEO_RMAX] = B.rmax[i];
  for (int k = 0; k < 9; k++) g[GEO_G + k] = B.rotG[k * N + i];
  for (int k = 0; k < 3; k++) g[GEO_R + k] = B.rxyz[k * N + i];
  for (int k = 0; k < 2; k++) g[GEO_E + k] = B.eps[k * N + i];
  for (int k = 0; k < 3; k++) g[GEO_BMIN + k] = B.bmin[k * N + i];
  for (int k = 0; k < 3; k++) g[GEO_BMAX + k] = B.bmax[k * N + i];
  for (int k = 0; k < 3; k++) g[GEO_VEL + k] = B.vel[k * N + i];
  for (int k = 0; k < 3; k++) g[GEO_ANGVEL + k] = B.angVel[k * N + i];
  g[GEO_MASS] = B.mass[i];
  const uint64_t fl =
      geo_flags(B.material[i], B.blocked[i], B.kind[i], B.isSphere[i], B.wallAxis[i], B.wallSense[i], B.lid[i]);
  g[GEO_FLAGS] = __longlong_as_double((long long)fl);
  B.disp[i] = B.hasBound[i] ? disp_up(ld3(B.pos, i, N) - ld3(B.refPos, i, N)) : __int_as_float(0x7fc00000);
}

// Per-body force and torque: the contributions of the body's interactions summed in
// interaction order, i.e. in the order of the reference's serial addForce/addTorque calls
// (core/ForceContainer.hpp:218-319), which makes the sums bitwise reproducible.  A group of
// 16 lanes serves one body: the lanes read 16 list entries at a time (coalesced), and the
// active ones are added in entry order through lane shuffles.
__global__ __launch_bounds__(256) void k_force_sum(BodyArrays B, PairArrays P, const int32_t* csrOff,
                                                   const int32_t* csrPair) {
  const int64_t N = B.n;
  const int64_t S = P.cap;
  const int lane = threadIdx.x & 63;
  const int gl = lane & 15;                 // lane in the group
  const int gbase = lane & ~15;             // first lane of the group in the wave
  const int64_t b = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  const bool valid = b < N;
  const int e0 = valid ? csrOff[b] : 0, e1 = valid ? csrOff[b + 1] : 0;
  V3<double> f = mk3(0.0, 0.0, 0.0), m = f;
  // every lane of the wave runs the same number of chunks so that shuffles stay convergent
  int chunks = (e1 - e0 + 15) >> 4;
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(chunks, off);
    chunks = o > chunks ? o : chunks;
  }
  for (int c = 0; c < chunks; c++) {
    const int e = e0 + c * 16 + gl;
    bool act = false;
    V3<double> F = mk3(0.0, 0.0, 0.0), T = F;
    if (e < e1) {
      // the entry says whether the interaction contributed this step and which side the
      // body is on; only contributing interactions are read
      const int32_t v = csrPair[e];
      if (v & 1) {
        act = true;
        const PairRec& R = P.r[v >> 2];
        F = ldv3(R.cF);
        if (((v >> 1) & 1) == 0) {
          T = ldv3(R.cT1);
        } else {
          F = -F;
          T = ldv3(R.cT2);
        }
      }
    }
    // positions (0..15) active in any group of the wave; an inactive lane contributes +0,
    // which leaves a sum unchanged, so the groups can share one loop over positions
    uint64_t m64 = __ballot(act);
    uint32_t pos = 0;
    for (int g = 0; g < 64; g += 16) pos |= (uint32_t)(m64 >> g) & 0xFFFFu;
    while (pos) {
      const int k = __ffs(pos) - 1;
      pos &= pos - 1;
      const int src = gbase + k;
      f = f + mk3(__shfl(F.x, src), __shfl(F.y, src), __shfl(F.z, src));
      m = m + mk3(__shfl(T.x, src), __shfl(T.y, src), __shfl(T.z, src));
    }
  }
  if (valid && gl == 0) {
    st3(B.force, b, N, f);
    st3(B.torque, b, N, m);
  }
}

struct NewtonArgs {
  BodyArrays B;
  PairArrays P;
  const int32_t* csrOff;
  const int32_t* csrPair;
  double dt, damping;
  double g[3];
  int exactAsph;
  int* maxDisp;  // 0: no bounded body, 1: all inside their sweep (0.5), 2: some outside (2)
  const int32_t* quiet;  // quiet_system_flag (set by the wall servo), nullptr: never
};

__device__ __forceinline__ Q4 dotq(V3<double> w, Q4 Q) {  // NewtonIntegrator::DotQ (:470-477)
  Q4 o;
  o.w = (-Q.x * w.x - Q.y * w.y - Q.z * w.z) / 2;
  o.x = (Q.w * w.x - Q.z * w.y + Q.y * w.z) / 2;
  o.y = (Q.z * w.x + Q.w * w.y - Q.x * w.z) / 2;
  o.z = (-Q.y * w.x + Q.x * w.y + Q.w * w.z) / 2;
  return o;
}

__device__ __forceinline__ double cundall(double damping, double a, double b) {
  return 1 - damping * sgn(a * b);
}

// NewtonIntegrator::action (pkg/dem/NewtonIntegrator.cpp:99-304) for isSuperquadrics=1, one
// body per thread.  The force and torque are first summed over the body's interactions in
// interaction order, which is the order of the reference's serial addForce/addTorque calls.
__global__ __launch_bounds__(256) void k_newton(NewtonArgs A) {
  const BodyArrays& B = A.B;
  const int64_t N = B.n;
  const int64_t S = A.P.cap;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int code = 0;
  if (i < N) {
    (void)S;
    V3<double> f = ld3(B.force, i, N), m = ld3(B.torque, i, N);
    const double dt = A.dt, damping = A.damping;
    const unsigned blocked = (unsigned)B.blocked[i];
    const bool isSQ = B.kind[i] == SDM_KIND_SQ;
    const bool isSphere = B.isSphere[i] != 0;
    const bool useAsph = A.exactAsph && isSQ && blocked != 63u;
    const bool damped = B.damped[i] != 0;
    V3<double> vel = ld3(B.vel, i, N);
    V3<double> angVel = ld3(B.angVel, i, N);
    const V3<double> inertia = ld3(B.inertia, i, N);
    if (blocked != 63u) {
      const double mass = B.mass[i];
      V3<double> a;
      if (blocked == 0) {
        a = mk3(f.x / mass + A.g[0], f.y / mass + A.g[1], f.z / mass + A.g[2]);
      } else {
        a.x = (blocked & 1u) ? 0.0 : 0.0 + (f.x / mass + A.g[0]);
        a.y = (blocked & 2u) ? 0.0 : 0.0 + (f.y / mass + A.g[1]);
        a.z = (blocked & 4u) ? 0.0 : 0.0 + (f.z / mass + A.g[2]);
      }
      if (damped) {
        a.x *= cundall(damping, a.x, vel.x + 0.5 * dt * a.x);
        a.y *= cundall(damping, a.y, vel.y + 0.5 * dt * a.y);
        a.z *= cundall(damping, a.z, vel.z + 0.5 * dt * a.z);
      }
      vel = vel + dt * a;
      if (!useAsph || (damped && isSphere)) {
        V3<double> aa;
        if (blocked == 0) {
          aa = mk3(m.x / inertia.x, m.y / inertia.y, m.z / inertia.z);
        } else {
          aa.x = (blocked & 8u) ? 0.0 : 0.0 + m.x / inertia.x;
          aa.y = (blocked & 16u) ? 0.0 : 0.0 + m.y / inertia.y;
          aa.z = (blocked & 32u) ? 0.0 : 0.0 + m.z / inertia.z;
        }
        if (damped) {
          aa.x *= cundall(damping, aa.x, angVel.x + 0.5 * dt * aa.x);
          aa.y *= cundall(damping, aa.y, angVel.y + 0.5 * dt * aa.y);
          aa.z *= cundall(damping, aa.z, angVel.z + 0.5 * dt * aa.z);
        }
        angVel = angVel + dt * aa;
      }
      if (useAsph) {
        if (blocked & 8u) m.x = 0;
        if (blocked & 16u) m.y = 0;
        if (blocked & 32u) m.z = 0;
        if (damped && !isSphere) {
          m.x *= cundall(damping, m.x, angVel.x);
          m.y *= cundall(damping, m.y, angVel.y);
          m.z *= cundall(damping, m.z, angVel.z);
        }
      }
    }
    Q4 ori;
    ori.w = B.ori[i]; ori.x = B.ori[N + i]; ori.y = B.ori[2 * N + i]; ori.z = B.ori[3 * N + i];
    // quiet_system_flag (NewtonIntegrator.cpp:246-250): velocities zeroed, no rotation or move
    const bool quiet = A.quiet && *A.quiet;
    if (quiet) {
      vel = mk3(0.0, 0.0, 0.0);
      angVel = vel;
    } else if (useAsph && !isSphere) {
      // leapfrogSuperquadricsRotate (:391-429) -> leapfrogAsphericalRotate (:306-335), with the
      // rotation matrix cached from the previous orientation.
      const M3<double> Am = ld9(B.rotL, i, N);
      V3<double> angMom = ld3(B.angMom, i, N);
      const V3<double> l_n = angMom + (dt / 2.) * m;
      const V3<double> l_b_n = mul(Am, l_n);
      const V3<double> w_b_n = mk3(l_b_n.x / inertia.x, l_b_n.y / inertia.y, l_b_n.z / inertia.z);
      const Q4 dQ_n = dotq(w_b_n, ori);
      const double h = dt / 2.;
      Q4 Q_half;
      Q_half.w = ori.w + dQ_n.w * h; Q_half.x = ori.x + dQ_n.x * h;
      Q_half.y = ori.y + dQ_n.y * h; Q_half.z = ori.z + dQ_n.z * h;
      angMom = angMom + dt * m;
      const V3<double> l_b_half = mul(Am, angMom);
      const V3<double> w_b_half = mk3(l_b_half.x / inertia.x, l_b_half.y / inertia.y, l_b_half.z / inertia.z);
      const Q4 dQ_half = dotq(w_b_half, Q_half);
      ori.w = ori.w + dQ_half.w * dt; ori.x = ori.x + dQ_half.x * dt;
      ori.y = ori.y + dQ_half.y * dt; ori.z = ori.z + dQ_half.z * dt;
      angVel = quat_apply(ori, w_b_half);
      ori = quat_normalized(ori);
      const M3<double> RL = quat_rot(quat_conj(ori)), RG = quat_rot(ori);
      double* gr = B.geo + i * GEO_STRIDE;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        B.rotL[k * N + i] = RL.m[k];
        B.rotG[k * N + i] = RG.m[k];
        gr[GEO_G + k] = RG.m[k];
      }
      st3(B.angMom, i, N, angMom);
    } else {
      // leapfrogSphericalRotate (:319-334)
      const double angle2 = sqn(angVel);
      if (angle2 != 0) {
        const double angle = sqrt(angle2);
        const V3<double> ax = mk3(angVel.x / angle, angVel.y / angle, angVel.z / angle);
        const double ha = 0.5 * (angle * dt);
        const double sh = xsin(ha);
        Q4 q;
        q.w = xcos(ha); q.x = sh * ax.x; q.y = sh * ax.y; q.z = sh * ax.z;
        ori = quat_mul(q, ori);
      }
      ori = quat_normalized(ori);
    }
    B.ori[i] = ori.w; B.ori[N + i] = ori.x; B.ori[2 * N + i] = ori.y; B.ori[3 * N + i] = ori.z;
    st3(B.vel, i, N, vel);
    st3(B.angVel, i, N, angVel);
    V3<double> pos = ld3(B.pos, i, N);
    if (!quiet) pos = pos + vel * dt;  // leapfrogTranslate (:306-317)
    st3(B.pos, i, N, pos);
    double* gp = B.geo + i * GEO_STRIDE;
    gp[GEO_POS] = pos.x; gp[GEO_POS + 1] = pos.y; gp[GEO_POS + 2] = pos.z;
    gp[GEO_VEL] = vel.x; gp[GEO_VEL + 1] = vel.y; gp[GEO_VEL + 2] = vel.z;
    gp[GEO_ANGVEL] = angVel.x; gp[GEO_ANGVEL + 1] = angVel.y; gp[GEO_ANGVEL + 2] = angVel.z;
    // saveMaximaDisplacement (:72-85)
    B.disp[i] = __int_as_float(0x7fc00000);
    if (B.hasBound[i]) {
      const V3<double> d = pos - ld3(B.refPos, i, N);
      B.disp[i] = disp_up(d);
      const double ax = fabs(d.x), ay = fabs(d.y), az = fabs(d.z);
      const double myz = (ay < az) ? az : ay;
      const double md = (ax < myz) ? myz : ax;
      code = (!(md != 0.0) || md < B.sweep[i]) ? 1 : 2;
    }
  }
  // wave-level max, then one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_down(code, off);
    code = o > code ? o : code;
  }
  if ((threadIdx.x & 63) == 0 && code) atomicMax(A.maxDisp, code);
}
