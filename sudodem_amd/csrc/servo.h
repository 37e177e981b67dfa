// servo.h -- wall stress servo of a cubic specimen, inside the step.
//
// CompressionEngine (pkg/dem/Compression.cpp) as the reference compiles it (SOLE_GAIN and
// NEW_SCHEME, Compression.hpp:26-29), action_cubic (:395-604): generalConsolidation (:324-353)
// or hydroConsolidation (:264-322) when z_servo, the triaxial shear branch (:498-603)
// otherwise; servo (:1252-1316), get_gain (:1125-1220), consol_ss (:1330-1345), getStress
// (:764-841), Shop::unbalancedForce (pkg/dem/Shop_01.cpp:148-189).
// Walls are bodies 0-5: left, right, front, back, bottom, top.  The controller state lives on
// the device and every decision is taken there, so a step with the servo has the same launch
// sequence as any other (graph capture) and no host round trip:
//   k_servo_begin   first run (box size, reference sizes, counters), the echo decision, and in
//                   shear the gain-update counter, which the reference advances before it
//                   moves any wall
//   k_servo_gain    (one wave per wall) kn of the wall's real contacts summed in interaction
//                   order (the wall's per-body list is in that order), gain = alpha area / kn
//   k_servo_step    the branch's wall moves; requests the gain update and the stability check
//                   when their intervals are due; the end of shear
//   k_servo_gain    again (the update after the move, with the areas before it)
//   k_servo_check_partial / k_servo_check_final
//                   stress tensor over the box volume, mean and deviator, the unbalanced force,
//                   completion (which quiets the integrator in consolidation), and the echo
// Sums over interactions and bodies in the check are block-tree sums in block order
// (deterministic; equal to the reference's serial sums to rounding).
#pragma once
#include "state.h"

struct ServoDev {
  // parameters (sdm_servo)
  double goal[3], max_vel, gain_alpha, f_threshold, fmin_threshold, unbf_tol, hydroRate, targetStrain;
  int32_t iterate_num_max, solve_num_max, activated[4], z_servo, hydro, shearMode, rampInterval, rampChunks,
      echoInterval, cont;
  // controller state
  int32_t first, iterate_num, solve_num, rampNum, needGain, needCheck, needEcho, quiet, completedNow, pad_;
  int64_t completedIter;
  double width, depth, height, x_area, y_area, z_area;
  double width0, depth0, height0, initVol;
  double ws[6];  // wsxx_left, wsxx_right, wsyy_front, wsyy_back, wszz_bottom, wszz_top
  double wsxx, wsyy, wszz;
  double stiff[6], gain[6];
  double mean, dev, unbf;
};

// updateBoxSize (:99-113)
__device__ __forceinline__ void servo_box(ServoDev& S, const BodyArrays& B) {
  const int64_t N = B.n;
  S.width = B.pos[1] - B.pos[0];
  S.depth = B.pos[N + 3] - B.pos[N + 2];
  S.height = B.pos[2 * N + 5] - B.pos[2 * N + 4];
  S.x_area = S.depth * S.height;
  S.y_area = S.width * S.height;
  S.z_area = S.depth * S.width;
}

// consol_ss, SOLE_GAIN (:1330-1345): box size, then each wall's stress from its force
__device__ __forceinline__ void servo_consol_ss(ServoDev& s, const BodyArrays& B) {
  const int64_t N = B.n;
  servo_box(s, B);
  s.ws[1] = B.force[1] / s.x_area;
  s.ws[0] = -B.force[0] / s.x_area;
  s.wsxx = 0.5 * (s.ws[0] + s.ws[1]);
  s.ws[3] = B.force[N + 3] / s.y_area;
  s.ws[2] = -B.force[N + 2] / s.y_area;
  s.wsyy = 0.5 * (s.ws[3] + s.ws[2]);
  s.ws[5] = B.force[2 * N + 5] / s.z_area;
  s.ws[4] = -B.force[2 * N + 4] / s.z_area;
  s.wszz = 0.5 * (s.ws[5] + s.ws[4]);
}

// The shear branch ramps the axial rate while rampNum < ramp_interval (and not continueFlag).
__device__ __forceinline__ bool servo_ramping(const ServoDev& s) { return !s.cont && s.rampNum < s.rampInterval; }

__global__ void k_servo_begin(ServoDev* S, BodyArrays B, const unsigned long long* iterDev) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  ServoDev& s = *S;
  s.completedNow = 0;
  s.needGain = 0;
  s.needCheck = 0;
  if (s.first) {  // action_cubic's first run (:397-480)
    servo_box(s, B);
    if (!s.cont) {
      s.width0 = s.width;
      s.depth0 = s.depth;
      s.height0 = s.height;
    }
    s.initVol = s.height0 * s.width0 * s.depth0;
    s.rampNum = 0;
    s.iterate_num = 0;
    s.solve_num = 0;
    s.needGain = 1;
    s.first = 0;
  }
  s.needEcho = (*iterDev % (unsigned long long)s.echoInterval) == 0ull;
  if (!s.z_servo && !servo_ramping(s) && (s.shearMode & 1)) {  // shear: gains before the moves
    s.iterate_num += 1;
    if (s.iterate_num >= s.iterate_num_max) {
      s.iterate_num = 1;
      s.needGain = 1;
    }
  }
}

// get_gain: one block of one wave per wall.  The wall's real contacts are the real entries of
// its list (another wall is never a real partner), summed in list = interaction order: 64
// entries per chunk, added lane by lane (a non-real entry adds +0, which changes no sum).
__global__ __launch_bounds__(64) void k_servo_gain(ServoDev* S, PairArrays P, const int32_t* csrOff,
                                                   const int32_t* csrPair) {
  const int w = blockIdx.x;
  if (w >= 6 || !S->needGain) return;
  const int lane = threadIdx.x;
  const int e0 = csrOff[w], e1 = csrOff[w + 1];
  double s = 0.0;
  for (int base = e0; base < e1; base += 64) {
    const int e = base + lane;
    double kn = 0.0;
    if (e < e1) {
      const PairRec& R = P.r[csrPair[e] >> 2];
      if (R.hasGeom[0] && R.hasPhys[0]) kn = R.kn[0];
    }
    for (int k = 0; k < 64; k++) s += __shfl(kn, k);
  }
  if (lane == 0) {
    S->stiff[w] = s;
    const double area = w < 2 ? S->x_area : (w < 4 ? S->y_area : S->z_area);
    S->gain[w] = S->gain_alpha * area / s;
  }
}

// Mathr::Sign(u) * std::min(fabs(u), lim)
__device__ __forceinline__ double servo_clamp(double u, double lim) {
  const double a = fabs(u);
  const double m = (lim < a) ? lim : a;
  return (u < 0.0 ? -1.0 : (u > 0.0 ? 1.0 : 0.0)) * m;
}

// std::min(a, b)
__device__ __forceinline__ double servo_min(double a, double b) { return (b < a) ? b : a; }

// servo, SOLE_GAIN (:1252-1316): consol_ss, then each active wall moves by its own gain, at
// most max_vel dt (the z walls only under z_servo)
__device__ __forceinline__ void servo_move(ServoDev& s, BodyArrays& B, double dt) {
  const int64_t N = B.n;
  servo_consol_ss(s, B);
  const double lim = s.max_vel * dt;
  if (s.activated[0]) B.pos[0] += -servo_clamp(s.gain[0] * (s.ws[0] - s.goal[0]), lim);
  if (s.activated[1]) B.pos[1] += servo_clamp(s.gain[1] * (s.ws[1] - s.goal[0]), lim);
  if (s.activated[2]) B.pos[N + 2] += -servo_clamp(s.gain[2] * (s.ws[2] - s.goal[1]), lim);
  if (s.activated[3]) B.pos[N + 3] += servo_clamp(s.gain[3] * (s.ws[3] - s.goal[1]), lim);
  if (s.z_servo) {
    B.pos[2 * N + 4] += -servo_clamp(s.gain[4] * (s.ws[4] - s.goal[2]), lim);
    B.pos[2 * N + 5] += servo_clamp(s.gain[5] * (s.ws[5] - s.goal[2]), lim);
  }
}

// The wall moves of the step's branch and the requests.
__global__ void k_servo_step(ServoDev* S, BodyArrays B, double dt, const unsigned long long* iterDev) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  ServoDev& s = *S;
  const int64_t N = B.n;
  s.needGain = 0;
  if (s.z_servo && !s.hydro) {  // generalConsolidation (NEW_SCHEME)
    s.iterate_num += 1;
    s.solve_num += 1;
    servo_move(s, B, dt);
    if (s.iterate_num >= s.iterate_num_max) {
      s.iterate_num = 0;
      s.needGain = 1;
    }
    if (s.solve_num >= s.solve_num_max) {
      s.solve_num = 0;
      s.needCheck = 1;
    }
  } else if (s.z_servo) {  // hydroConsolidation: walls below their goal close at a bounded rate
    servo_box(s, B);
    const double size[3] = {s.width, s.depth, s.height};
    const double area[3] = {s.x_area, s.y_area, s.z_area};
    for (int w = 0; w < 6; w++) {
      const int ax = w >> 1;
      const double stress = B.force[ax * N + w] / area[ax];
      const double delta = s.goal[ax] - fabs(stress);
      if (delta > 0) {
        const double step = servo_min(0.1 * s.gain[w] * delta, s.hydroRate * size[ax] * dt);
        if ((w & 1) == 0) B.pos[ax * N + w] += step;
        else B.pos[ax * N + w] -= step;
      }
    }
    s.iterate_num += 1;
    if (s.iterate_num >= s.iterate_num_max) {
      s.iterate_num = 1;
      s.needGain = 1;
    }
    s.solve_num += 1;
    if (s.solve_num >= s.solve_num_max) {
      s.solve_num = 1;
      s.needCheck = 1;
    }
  } else {  // shear: axial strain control (the gain update was requested by k_servo_begin)
    const double goalz = s.goal[2];
    if (servo_ramping(s)) {
      s.rampNum++;
      const double ramp_inter = 0.5 * (double)s.rampInterval;
      const double chunk = ramp_inter / (double)s.rampChunks;
      double vel;
      if ((double)s.rampNum < ramp_inter)
        vel = (double)((float)((double)s.rampNum / chunk + 1.0) / (float)s.rampChunks) * goalz * 2.0;
      else
        vel = ((double)((float)((double)(s.rampInterval - s.rampNum) / chunk + 1.0) / (float)s.rampChunks) + 1.0) *
              goalz;
      B.pos[2 * N + 4] += vel * dt * s.height;
      B.pos[2 * N + 5] += -vel * dt * s.height;
    } else {
      B.pos[2 * N + 4] += goalz * dt * s.height;
      B.pos[2 * N + 5] += -goalz * dt * s.height;
    }
    if (s.shearMode & 1) {
      servo_move(s, B, dt);  // drained: lateral stress servo (z_servo is off)
    } else {                 // constant volume
      s.width = B.pos[1] - B.pos[0];
      s.depth = B.pos[N + 3] - B.pos[N + 2];
      s.height = B.pos[2 * N + 5] - B.pos[2 * N + 4];
      const double a = s.initVol / s.height;
      const double dd = 0.5 * (a / s.width - s.depth);
      const double dw = a / (s.depth + dd) - s.width;
      B.pos[0] += -0.5 * dw;
      B.pos[1] += 0.5 * dw;
      B.pos[N + 2] += -0.5 * dd;
      B.pos[N + 3] += 0.5 * dd;
    }
    if (log(s.height0 / s.height) > s.targetStrain) {  // stopAtIter = iter + 1
      s.completedIter = (int64_t)*iterDev;
      s.completedNow = 1;
    }
  }
}

// Stability check, per block: the stress tensor sum of (fn + fs) (pos2 - pos1)^T and the sum
// of |fn + fs| over real interactions with a non-zero normal force (9 + 1 partials), and the
// sum of |F + m g| over bodies that are not fully blocked (1 partial).
// The launch covers an upper bound of the container; its size is read from npDev.
__global__ __launch_bounds__(256) void k_servo_check_partial(const ServoDev* S, PairArrays P,
                                                             const unsigned long long* npDev, BodyArrays B,
                                                             double gx, double gy, double gz, double* partial) {
  if (!S->needCheck && !S->needEcho) return;
  const int64_t np = (int64_t)*npDev;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[11];
#pragma unroll
  for (int k = 0; k < 11; k++) v[k] = 0.0;
  const int64_t N = B.n;
  if (i < np) {
    const PairRec& R = P.r[i];
    if (R.hasGeom[0] && R.hasPhys[0]) {
      const double fx = R.fn[0], fy = R.fn[1], fz = R.fn[2];
      if (sqrt(fx * fx + (fy * fy + fz * fz)) != 0.0) {
        const int a = R.id1[0], b = R.id2[0];
        double f[3], l[3];
        for (int k = 0; k < 3; k++) {
          f[k] = R.fn[k] + R.fs[k];
          l[k] = B.pos[k * N + b] - B.pos[k * N + a];
        }
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++) v[3 * r + c] = f[r] * l[c];
        v[9] = sqrt(f[0] * f[0] + (f[1] * f[1] + f[2] * f[2]));
      }
    }
  }
  if (i < N && (unsigned)B.blocked[i] != 63u) {
    const double m = B.mass[i];
    const double x = B.force[i] + m * gx, y = B.force[N + i] + m * gy, z = B.force[2 * N + i] + m * gz;
    v[10] = sqrt(x * x + (y * y + z * z));
  }
  __shared__ double red[11][256];
#pragma unroll
  for (int k = 0; k < 11; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 11; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 11) partial[(int64_t)blockIdx.x * 11 + threadIdx.x] = red[threadIdx.x][0];
}

// getStress and the completion test of the branch (completion in generalConsolidation quiets
// the integrator of this step; hydroConsolidation tests the UnbalancedForce of the last echo),
// then the echo: UnbalancedForce, and consol_ss after the moves in consolidation and in
// constant-volume shear.
__global__ void k_servo_check_final(ServoDev* S, BodyArrays B, const double* partial, int64_t nb,
                                    const unsigned long long* iterDev) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  ServoDev& s = *S;
  s.needGain = 0;
  if (!s.needCheck && !s.needEcho) return;
  double t[11];
  for (int k = 0; k < 11; k++) t[k] = 0.0;
  for (int64_t b = 0; b < nb; b++)
    for (int k = 0; k < 11; k++) t[k] += partial[b * 11 + k];
  if (s.needCheck) {
    s.needCheck = 0;
    servo_box(s, B);  // getStressTensor calls updateBoxSize after the walls moved
    const double vol = s.height * s.z_area;
    double g[9];
    for (int k = 0; k < 9; k++) g[k] = t[k] / vol;
    const double mean = (g[0] + (g[4] + g[8])) / 3.0;
    g[0] -= mean;
    g[4] -= mean;
    g[8] -= mean;
    double q[9];  // squares in Eigen's column-major order
    for (int col = 0; col < 3; col++)
      for (int row = 0; row < 3; row++) q[3 * col + row] = g[3 * row + col] * g[3 * row + col];
    const double sq = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + (q[7] + q[8])));
    const double dev = sqrt(sq) * sqrt(1.5);
    s.mean = mean;
    s.dev = dev;
    if (!s.hydro) {
      if ((fabs(mean - s.goal[0]) < s.goal[0] * s.f_threshold) && (dev / mean < s.fmin_threshold)) {
        s.unbf = t[10] / (t[9] * 2.0);
        if (s.unbf < s.unbf_tol) {  // stopAtIter = iter + 1; newton->quiet_system_flag = true
          s.completedIter = (int64_t)*iterDev;
          s.completedNow = 1;
          s.quiet = 1;
        }
      }
    } else if ((fabs(mean - s.goal[0]) / s.goal[0] < s.f_threshold) && (dev / mean < s.fmin_threshold)) {
      if (s.unbf < s.unbf_tol) {  // stopAtIter = iter + 1
        s.completedIter = (int64_t)*iterDev;
        s.completedNow = 1;
      }
    }
  }
  if (s.needEcho) {
    s.needEcho = 0;
    s.unbf = t[10] / (t[9] * 2.0);
    if (s.z_servo || (s.shearMode & 2)) servo_consol_ss(s, B);
  }
}

// NewtonIntegrator resets quiet_system_flag after its pass (:299).
__global__ void k_servo_end(ServoDev* S) {
  if (blockIdx.x == 0 && threadIdx.x == 0) S->quiet = 0;
}
