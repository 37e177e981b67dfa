// servo.h -- wall stress servo of a cubic specimen, inside the step.
//
// CompressionEngine (pkg/dem/Compression.cpp) as the reference compiles it (SOLE_GAIN and
// NEW_SCHEME, Compression.hpp:26-29), consolidation branch: action_cubic (:395-481) ->
// generalConsolidation (:324-345) -> servo (:1252-1316), get_gain (:1125-1220), consol_ss
// (:1330-1345), getStress (:764-841), Shop::unbalancedForce (pkg/dem/Shop_01.cpp:148-189).
// Walls are bodies 0-5: left, right, front, back, bottom, top.  The controller state lives on
// the device and every decision is taken there, so a step with the servo has the same launch
// sequence as any other (graph capture) and no host round trip:
//   k_servo_begin   first run: box size, counters, gain update requested
//   k_servo_gain    (one wave per wall) kn of the wall's real contacts summed in interaction
//                   order (the wall's per-body list is in that order), gain = alpha area / kn
//   k_servo_step    counters, wall stresses from the wall forces, wall moves; requests the
//                   gain update and the stability check when their intervals are due
//   k_servo_gain    again (the update after the move, with the areas before it)
//   k_servo_check_partial / k_servo_check_final
//                   stress tensor over the box volume, mean and deviator, and when they meet
//                   the target the unbalanced force; completion quiets the integrator
// Sums over interactions and bodies in the check are block-tree sums in block order
// (deterministic; equal to the reference's serial sums to rounding).
#pragma once
#include "state.h"

struct ServoDev {
  // parameters (sdm_servo)
  double goal[3], max_vel, gain_alpha, f_threshold, fmin_threshold, unbf_tol;
  int32_t iterate_num_max, solve_num_max, activated[4], z_servo;
  // controller state
  int32_t first, iterate_num, solve_num, needGain, needCheck, quiet, completedNow, pad_;
  int64_t completedIter;
  double width, depth, height, x_area, y_area, z_area;
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

__global__ void k_servo_begin(ServoDev* S, BodyArrays B) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (!S->first) return;
  servo_box(*S, B);
  S->iterate_num = 0;
  S->solve_num = 0;
  S->needGain = 1;
  S->first = 0;
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

// generalConsolidation (NEW_SCHEME): counters, servo (consol_ss + wall moves), the requests.
__global__ void k_servo_step(ServoDev* S, BodyArrays B, double dt) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  ServoDev& s = *S;
  const int64_t N = B.n;
  s.needGain = 0;
  s.completedNow = 0;
  s.iterate_num += 1;
  s.solve_num += 1;
  // consol_ss, SOLE_GAIN
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
  // servo, SOLE_GAIN: each wall moves by its own gain, at most max_vel dt
  const double lim = s.max_vel * dt;
  if (s.activated[0]) B.pos[0] += -servo_clamp(s.gain[0] * (s.ws[0] - s.goal[0]), lim);
  if (s.activated[1]) B.pos[1] += servo_clamp(s.gain[1] * (s.ws[1] - s.goal[0]), lim);
  if (s.activated[2]) B.pos[N + 2] += -servo_clamp(s.gain[2] * (s.ws[2] - s.goal[1]), lim);
  if (s.activated[3]) B.pos[N + 3] += servo_clamp(s.gain[3] * (s.ws[3] - s.goal[1]), lim);
  if (s.z_servo) {
    B.pos[2 * N + 4] += -servo_clamp(s.gain[4] * (s.ws[4] - s.goal[2]), lim);
    B.pos[2 * N + 5] += servo_clamp(s.gain[5] * (s.ws[5] - s.goal[2]), lim);
  }
  if (s.iterate_num >= s.iterate_num_max) {
    s.iterate_num = 0;
    s.needGain = 1;
  }
  if (s.solve_num >= s.solve_num_max) {
    s.solve_num = 0;
    s.needCheck = 1;
  }
}

// Stability check, per block: the stress tensor sum of (fn + fs) (pos2 - pos1)^T and the sum
// of |fn + fs| over real interactions with a non-zero normal force (9 + 1 partials), and the
// sum of |F + m g| over bodies that are not fully blocked (1 partial).
// The launch covers an upper bound of the container; its size is read from npDev.
__global__ __launch_bounds__(256) void k_servo_check_partial(const ServoDev* S, PairArrays P,
                                                             const unsigned long long* npDev, BodyArrays B,
                                                             double gx, double gy, double gz, double* partial) {
  if (!S->needCheck) return;
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

// getStress and the completion test; completion quiets the integrator of this step.
__global__ void k_servo_check_final(ServoDev* S, BodyArrays B, const double* partial, int64_t nb,
                                    const unsigned long long* iterDev) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  ServoDev& s = *S;
  s.needGain = 0;
  if (!s.needCheck) return;
  s.needCheck = 0;
  double t[11];
  for (int k = 0; k < 11; k++) t[k] = 0.0;
  for (int64_t b = 0; b < nb; b++)
    for (int k = 0; k < 11; k++) t[k] += partial[b * 11 + k];
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
  if ((fabs(mean - s.goal[0]) < s.goal[0] * s.f_threshold) && (dev / mean < s.fmin_threshold)) {
    s.unbf = t[10] / (t[9] * 2.0);
    if (s.unbf < s.unbf_tol) {  // stopAtIter = iter + 1; newton->quiet_system_flag = true
      s.completedIter = (int64_t)*iterDev;
      s.completedNow = 1;
      s.quiet = 1;
    }
  }
}

// NewtonIntegrator resets quiet_system_flag after its pass (:299).
__global__ void k_servo_end(ServoDev* S) {
  if (blockIdx.x == 0 && threadIdx.x == 0) S->quiet = 0;
}
