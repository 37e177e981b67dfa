// Error of the mixed form's difference evaluation (delta_eval, sudodem_amd/csrc/delta.h) against
// the double evaluation of Disfun (pkg/dem/Superquadrics_Ig2.cpp:1624-1678) at the same angles,
// over random pairs: two bodies of random semi-axes (0.8-1.2 mm) and exponents, random
// orientations, centres 1.6-2.1 mm apart, an anchor near the line of centres (mode 0), or with
// probability 1/2 one body's local direction within 1e-7..0.1 of a principal plane (mode 2),
// and steps of 1e-7 to 0.5 rad in random directions.  Per step-size band it prints the
// share of evaluations the difference form takes, the share whose touching test it certifies,
// the certified ones whose decision differs from the double evaluation (must be 0) and the
// largest |f - f_double| / (2^-24 (ds + 4 dk + dl) + anchor rounding) (the solver certifies with
// kDeltaErr times that bound).  Test infrastructure, built by __graft_entry__.build().
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <cmath>
#include "delta.h"

__device__ inline double urand(uint64_t& s) {
  s ^= s << 13; s ^= s >> 7; s ^= s << 17;
  return (double)(s >> 11) * (1.0 / 9007199254740992.0);
}
__device__ M3<double> rand_rot(uint64_t& s) {
  double q[4], n = 0;
  for (int k = 0; k < 4; k++) { double u1 = urand(s) + 1e-300, u2 = urand(s); q[k] = sqrt(-2 * log(u1)) * cos(6.283185307179586 * u2); n += q[k] * q[k]; }
  n = sqrt(n);
  Q4 Q; Q.w = q[0] / n; Q.x = q[1] / n; Q.y = q[2] / n; Q.z = q[3] / n;
  return quat_rot(Q);
}

__global__ void probe(int n, uint64_t seed, double epslo, double epshi, int mode, float* ratio, int* flags, float* step,
                      float* lgrel, double* info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t s = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(i + 1));
  for (int k = 0; k < 4; k++) urand(s);
  PairGeo<double> g;
  g.s1.rx = 0.8e-3 + 0.4e-3 * urand(s); g.s1.ry = 0.8e-3 + 0.4e-3 * urand(s); g.s1.rz = 0.8e-3 + 0.4e-3 * urand(s);
  g.s2.rx = 0.8e-3 + 0.4e-3 * urand(s); g.s2.ry = 0.8e-3 + 0.4e-3 * urand(s); g.s2.rz = 0.8e-3 + 0.4e-3 * urand(s);
  g.s1.eps1 = epslo + (epshi - epslo) * urand(s); g.s1.eps2 = epslo + (epshi - epslo) * urand(s);
  g.s2.eps1 = epslo + (epshi - epslo) * urand(s); g.s2.eps2 = epslo + (epshi - epslo) * urand(s);
  g.G1 = rand_rot(s); g.G2 = rand_rot(s);
  M3<double> R = rand_rot(s);
  const V3<double> dir = mul(R, mk3(1.0, 0.0, 0.0));
  const double L = 2.0e-3 * (0.8 + 0.25 * urand(s));
  g.pos1 = mk3(0.0, 0.0, 0.0);
  g.pos2 = dir * L;
  const M3<double> gC = crotation_mat(g.pos1, g.pos2);
  double a0 = -0.5 + urand(s), a1 = -0.5 + urand(s);
  int smallk = -1;
  if (mode == 1 || (mode == 2 && urand(s) < 0.5)) {
    // anchor direction with a small component of one body's local direction
    V3<double> c = mul(gC, mk3(cos(a0) * cos(a1), sin(a0) * cos(a1), sin(a1)));
    const bool second = urand(s) < 0.5;
    V3<double> nl = second ? mulT(g.G2, c * -1.0) : mulT(g.G1, c);
    smallk = (int)(3 * urand(s)); if (smallk > 2) smallk = 2;
    const double mag = pow(10.0, -7.0 + 6.0 * urand(s)) * (urand(s) < 0.5 ? -1.0 : 1.0);
    double* pc = smallk == 0 ? &nl.x : (smallk == 1 ? &nl.y : &nl.z);
    *pc = mag;
    nl = normalized(nl);
    c = second ? mul(g.G2, nl) * -1.0 : mul(g.G1, nl);
    const V3<double> u = mulT(gC, c);
    a1 = asin(fmax(-1.0, fmin(1.0, u.z)));
    a0 = atan2(u.y, u.x);
  }
  const double sm = pow(10.0, -7.0 + 6.7 * urand(s));
  const double th = 6.283185307179586 * urand(s);
  const double d0 = sm * cos(th), d1 = sm * sin(th);
  DeltaPair P;
  V3<double> p1, p2;
  bool touching = true, ok = true;
  double cdd = 0.0;
  const double fa = delta_anchor(g, gC, a0, a1, P, p1, p2, touching, cdd, ok);
  int fl = 0;
  float r = -1.0f, lr = 0.0f;
  if (ok && isfinite(fa)) {
    fl |= 1;
    bool t = true, okd = true;
    float cd = 0, ds = 0, dk = 0, dl = 0;
    const double fd = delta_eval(P, (float)d0, (float)d1, t, cd, ds, dk, dl, okd);
    if (okd) {
      fl |= 2;
      NMGeo<double> nmg;
      nmg.g = g;
      nmg.gC = gC;
      V3<double> q1, q2;
      bool te = true;
      int nev = 0;
      const double fe = nm_eval<double, true>(nmg, a0 + d0, a1 + d1, q1, q2, te, nev);
      const double qa = fabs(P.b1.q[0]) + fabs(P.b1.q[1]) + fabs(P.b1.q[2]) + fabs(P.b2.q[0]) + fabs(P.b2.q[1]) + fabs(P.b2.q[2]);
      const double bound = 5.9604644775390625e-08 * ((double)ds + 4.0 * (double)dk + (double)dl) + 1e-16 * fe + 3.552713678800501e-15 * qa;
      r = (float)(fabs(fd - fe) / bound);
      double* I = info + 16 * (size_t)i;
      I[0] = fd; I[1] = fe; I[2] = ds; I[3] = dk; I[4] = dl; I[5] = g.s1.eps1; I[6] = g.s1.eps2; I[7] = g.s2.eps1; I[8] = g.s2.eps2; I[9] = smallk; I[10] = d0; I[11] = d1; I[12] = cd; I[13] = fa; I[14] = a0; I[15] = a1;
      lr = (float)((double)dl / ((double)ds + 4.0 * (double)dk + 1e-30));
      const double ef = 4.0 * 5.9604644775390625e-08 * ((double)ds + 4.0 * (double)dk + (double)dl) + 4e-16 * fd;
      const bool cert = fabs((double)cd) > ef + 4.0 * 5.9604644775390625e-08 * fd;
      if (cert) fl |= 4;
      if (cert && (t != te)) fl |= 8;
      if (smallk >= 0) fl |= 16;
    }
  }
  ratio[i] = r;
  flags[i] = fl;
  step[i] = (float)sm;
  lgrel[i] = lr;
}

#define CHK(x)                                                              \
  do {                                                                      \
    if ((x) != hipSuccess) {                                                \
      printf("HIP error at %s:%d\n", __FILE__, __LINE__);                  \
      return 1;                                                             \
    }                                                                       \
  } while (0)

int main(int argc, char** argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 1 << 22;
  const double elo = argc > 2 ? atof(argv[2]) : 0.6, ehi = argc > 3 ? atof(argv[3]) : 1.4;
  const int mode = argc > 4 ? atoi(argv[4]) : 2;
  float *r, *st, *lg;
  double* info;
  int* fl;
  CHK(hipMalloc(&r, n * 4)); CHK(hipMalloc(&fl, n * 4)); CHK(hipMalloc(&st, n * 4)); CHK(hipMalloc(&lg, n * 4));
  CHK(hipMalloc(&info, (size_t)n * 16 * 8));
  hipLaunchKernelGGL(probe, dim3((n + 255) / 256), dim3(256), 0, 0, n, 12345ull + mode, elo, ehi, mode, r, fl, st, lg, info);
  if (hipDeviceSynchronize() != hipSuccess) { printf("kernel failed\n"); return 1; }
  std::vector<float> R(n), S(n), LG(n);
  std::vector<int> F(n);
  CHK(hipMemcpy(R.data(), r, n * 4, hipMemcpyDeviceToHost));
  CHK(hipMemcpy(F.data(), fl, n * 4, hipMemcpyDeviceToHost));
  CHK(hipMemcpy(S.data(), st, n * 4, hipMemcpyDeviceToHost));
  CHK(hipMemcpy(LG.data(), lg, n * 4, hipMemcpyDeviceToHost));
  printf("eps [%.2f, %.2f] mode %d, %d cases\n", elo, ehi, mode, n);
  const double edges[] = {1e-7, 1e-5, 1e-3, 1e-2, 3e-2, 1e-1, 0.5};
  for (int b = 0; b < 6; b++) {
    long na = 0, nok = 0, ncert = 0, nbad = 0, nsm = 0, nsmok = 0;
    std::vector<float> rs;
    float lgmax = 0;
    for (int i = 0; i < n; i++) {
      if (S[i] < edges[b] || S[i] >= edges[b + 1]) continue;
      if (!(F[i] & 1)) continue;
      na++;
      if (F[i] & 16) nsm++;
      if (F[i] & 2) { nok++; rs.push_back(R[i]); lgmax = std::max(lgmax, LG[i]); if (F[i] & 16) nsmok++; }
      if (F[i] & 4) ncert++;
      if (F[i] & 8) nbad++;
    }
    std::sort(rs.begin(), rs.end());
    float mx = rs.empty() ? 0 : rs.back();
    float p999 = rs.empty() ? 0 : rs[(size_t)(0.999 * (rs.size() - 1))];
    printf("band %.0e %.0e anchored %ld in_range %.6f small_in_range %.6f certified %ld wrong %ld p999 %.4f max %.4f dlrel %.3g\n",
           edges[b], edges[b + 1], na, na ? (double)nok / na : 0.0, nsm ? (double)nsmok / nsm : 0.0, ncert, nbad, p999, mx, lgmax);
  }
  // the worst cases
  std::vector<int> idx;
  for (int i = 0; i < n; i++) if (F[i] & 2) idx.push_back(i);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return R[a] > R[b]; });
  std::vector<double> I(16);
  for (int k = 0; k < 5 && k < (int)idx.size(); k++) {
    CHK(hipMemcpy(I.data(), info + 16 * (size_t)idx[k], 16 * 8, hipMemcpyDeviceToHost));
    printf("worst %d: ratio %.3g fd %.17g fe %.17g err %.3g ds %.3g dk %.3g dl %.3g eps %.3f %.3f %.3f %.3f smallk %d d %.3g %.3g cd %.3g fa %.3g a %.4f %.4f\n",
           k, R[idx[k]], I[0], I[1], fabs(I[0] - I[1]), I[2], I[3], I[4], I[5], I[6], I[7], I[8], (int)I[9], I[10], I[11], I[12], I[13], I[14], I[15]);
  }
  CHK(hipFree(r)); CHK(hipFree(fl)); CHK(hipFree(st)); CHK(hipFree(lg)); CHK(hipFree(info));
  return 0;
}
