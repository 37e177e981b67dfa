// observe.h -- observables of a packing, reduced on the device.
//
//   mean coordination number   Shop__meanCoordinationNumber (py/_utils.cpp:522-539)
//   Love-Weber stress          Shop::getStress (pkg/dem/Shop_02.cpp:350-377)
// The reference sums serially over the interaction container; here each block reduces its
// interactions in a fixed tree and the block partials are added in block order, so the result
// is deterministic and equals the serial sum to rounding.
#pragma once
#include "state.h"

// Real interactions whose normal force is non-zero, and bodies that are not walls.
__global__ __launch_bounds__(256) void k_cn_count(PairArrays P, int64_t np, BodyArrays B, int kindWall,
                                                  unsigned long long* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long c = 0, b = 0;
  if (i < np) {
    const PairRec& R = P.r[i];
    if (R.hasGeom[0] && R.hasPhys[0]) {
      const double fx = R.fn[0], fy = R.fn[1], fz = R.fn[2];
      c = (sqrt(fx * fx + (fy * fy + fz * fz)) == 0.0) ? 0ull : 1ull;  // normalForce.norm() == 0 (:531)
    }
  }
  if (i < B.n) b = B.kind[i] != kindWall ? 1ull : 0ull;
  for (int off = 32; off > 0; off >>= 1) {
    c += __shfl_down(c, off);
    b += __shfl_down(b, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (c) atomicAdd(&out[0], c);
    if (b) atomicAdd(&out[1], b);
  }
}

// Per block: sum over its real interactions of (normalForce + shearForce) (x) branch, with
// branch = pos1 - pos2 - hSize cellDist (Shop_02.cpp:369-371); 9 partials per block.
__global__ __launch_bounds__(256) void k_stress_partial(PairArrays P, int64_t np, BodyArrays B, int periodic,
                                                        double L0, double L1, double L2, double* partial) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s[9];
#pragma unroll
  for (int k = 0; k < 9; k++) s[k] = 0.0;
  if (i < np) {
    const PairRec& R = P.r[i];
    if (R.hasGeom[0] && R.hasPhys[0]) {
      const int64_t N = B.n;
      const int a = R.id1[0], b = R.id2[0];
      double br[3], f[3];
      for (int k = 0; k < 3; k++) br[k] = B.pos[k * N + a] - B.pos[k * N + b];
      if (periodic) {
        br[0] -= L0 * (double)R.cellDist[0];
        br[1] -= L1 * (double)R.cellDist[1];
        br[2] -= L2 * (double)R.cellDist[2];
      }
      for (int k = 0; k < 3; k++) f[k] = R.fn[k] + R.fs[k];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) s[3 * r + c] = f[r] * br[c];
    }
  }
  __shared__ double red[9][256];
#pragma unroll
  for (int k = 0; k < 9; k++) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 9; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 9) partial[(int64_t)blockIdx.x * 9 + threadIdx.x] = red[threadIdx.x][0];
}

// Block partials added in block order, divided by the volume.
__global__ void k_stress_final(const double* partial, int64_t nb, double volume, double* out) {
  const int k = threadIdx.x;
  if (blockIdx.x != 0 || k >= 9) return;
  double s = 0.0;
  for (int64_t b = 0; b < nb; b++) s += partial[b * 9 + k];
  out[k] = s / volume;
}

__global__ void k_gather_forces(BodyArrays B, const int32_t* ids, int n, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t N = B.n;
  const int id = ids[i];
  for (int k = 0; k < 3; k++) out[3 * i + k] = B.force[k * N + id];
}
