// state.h -- device layout of the scene.
//
// Bodies and interactions are structure-of-arrays in HBM.  A vector attribute with k
// components occupies k consecutive planes of length N (bodies) or capacity (interactions):
// component c of element i is at ptr[c * stride + i], so a wavefront reading one component
// of 64 consecutive elements issues one coalesced 512-byte load.
#pragma once
#include <stdint.h>

// Body attributes: (type, member, components, attribute name).  Names follow the reference
// attributes (core/State.hpp, pkg/dem/Superquadrics.hpp, core/Bound.hpp).
#define SDM_BODY_FIELDS(X)                               \
  X(double, pos, 3, "state.pos")                         \
  X(double, ori, 4, "state.ori")                         \
  X(double, vel, 3, "state.vel")                         \
  X(double, angVel, 3, "state.angVel")                   \
  X(double, angMom, 3, "state.angMom")                   \
  X(double, mass, 1, "state.mass")                       \
  X(double, inertia, 3, "state.inertia")                 \
  X(int32_t, blocked, 1, "state.blockedDOFs")            \
  X(int32_t, damped, 1, "state.isDamped")                \
  X(int32_t, kind, 1, "shape.kind")                      \
  X(double, rxyz, 3, "shape.rxyz")                       \
  X(double, eps, 2, "shape.eps")                         \
  X(double, rmax, 1, "shape.r_max")                      \
  X(int32_t, isSphere, 1, "shape.isSphere")              \
  X(int32_t, wallAxis, 1, "shape.wallAxis")              \
  X(int32_t, wallSense, 1, "shape.wallSense")            \
  X(double, rotL, 9, "shape.rot_mat2local")              \
  X(double, rotG, 9, "shape.rot_mat2global")             \
  X(int32_t, material, 1, "body.material")               \
  X(int32_t, specimen, 1, "body.specimen")               \
  X(int32_t, hasBound, 1, "bound.exists")                \
  X(double, bmin, 3, "bound.min")                        \
  X(double, bmax, 3, "bound.max")                        \
  X(double, refPos, 3, "bound.refPos")                   \
  X(double, sweep, 1, "bound.sweepLength")               \
  X(int32_t, lastUpd, 1, "bound.lastUpdateIter")         \
  X(double, force, 3, "forces.force")                    \
  X(double, torque, 3, "forces.torque")

// Interaction attributes (core/Interaction.hpp, SuperquadricsGeom(2), SuperquadricsPhys,
// SuperquadricsHertzMindlinPhys), plus the per-contact scatter and solver diagnostics.
#define SDM_PAIR_FIELDS(X)                                   \
  X(int32_t, id1, 1, "intr.id1")                             \
  X(int32_t, id2, 1, "intr.id2")                             \
  SDM_PAIR_FIELDS_NOIDS(X)

// Field order = record layout: the first 64 bytes (ids, flags, iterMadeReal, depth) and the
// angles are what every interaction reads each step; a contact also touches the next lines.
#define SDM_PAIR_FIELDS_NOIDS(X)                             \
  X(int32_t, hasGeom, 1, "intr.hasGeom")                     \
  X(int32_t, hasPhys, 1, "intr.hasPhys")                     \
  X(int32_t, functorSet, 1, "intr.functorSet")               \
  X(int32_t, geomExists, 1, "intr.geomExists")               \
  X(int32_t, isShearNew, 1, "geom.isShearNew")               \
  X(int32_t, gSphere, 1, "geom.isSphere")                    \
  X(int32_t, active, 1, "contact.active")                    \
  X(int32_t, nEval, 1, "diag.nEval")                         \
  X(int32_t, stop, 1, "diag.stop")                           \
  X(int32_t, cellDist, 3, "intr.cellDist")                   \
  X(int64_t, iterMadeReal, 1, "intr.iterMadeReal")           \
  X(double, depth, 1, "geom.PenetrationDepth")               \
  X(double, ang, 2, "geom.contactAngle")                     \
  X(double, normal, 3, "geom.normal")                        \
  X(double, fs, 3, "phys.shearForce")                        \
  X(double, fn, 3, "phys.normalForce")                       \
  X(double, kn, 1, "phys.kn")                                \
  X(double, ks, 1, "phys.ks")                                \
  X(double, tanPhi, 1, "phys.tangensOfFrictionAngle")        \
  X(double, betan, 1, "phys.betan")                          \
  X(double, betas, 1, "phys.betas")                          \
  X(double, nuab, 1, "phys.nuab")                            \
  X(double, Gab, 1, "phys.Gab")                              \
  X(double, mu1, 1, "phys.mu1")                              \
  X(double, mu2, 1, "phys.mu2")                              \
  X(double, cF, 3, "contact.force")                          \
  X(double, cT1, 3, "contact.torque1")                       \
  X(double, cT2, 3, "contact.torque2")                       \
  X(double, cpt, 3, "geom.contactPoint")                     \
  X(double, shearInc, 3, "geom.shearInc")                    \
  X(double, relVn, 3, "geom.relativeVn")                     \
  X(double, relVs, 3, "geom.relativeVs")                     \
  X(double, twist, 3, "geom.twist_axis")                     \
  X(double, orth, 3, "geom.orthonormal_axis")                \
  X(double, point1, 3, "geom.point1")                        \
  X(double, point2, 3, "geom.point2")                        \
  X(double, curv1, 2, "geom.curvatures1")                    \
  X(double, curv2, 2, "geom.curvatures2")                    \
  X(double, curvAngle, 1, "geom.curvatureAngle")             \
  X(double, alpha, 1, "geom.alpha")                          \
  X(double, hK, 1, "geom.K")                                 \
  X(double, hE, 1, "geom.E")                                 \
  X(double, curvSum, 1, "geom.curvatures_sum")

// One interaction = one record (array of structures).  The kernels reach an interaction
// through queues and per-body lists in no particular order, and a contact reads and
// writes ~60 values: as one record these are 5 whole 128-byte lines, as separate arrays
// they were ~60 partly used lines.  640 bytes keeps every record line-aligned.
struct PairRec {
#define SDM_PREC(T, m, k, name) T m[k];
  SDM_PAIR_FIELDS(SDM_PREC)
#undef SDM_PREC
  int32_t pad_[18];
};
static_assert(sizeof(PairRec) == 640, "interaction record size");

struct BodyArrays {
  int32_t n;
#define SDM_DECL(T, m, k, name) T* m;
  SDM_BODY_FIELDS(SDM_DECL)
#undef SDM_DECL
  // Packed copy of what the pair kernels gather per body, one 256-byte record per body:
  // pos[3], r_max, bound min[3], max[3], rot_mat2global[9], semi-axes[3], exponents[2],
  // vel[3], angVel[3], mass, and the packed flags (geo_flags).  Interactions reference
  // bodies in no spatial order, so a record (two 128-byte lines) replaces ~30 scattered
  // reads.  The first line alone (pos, r_max, bounds) settles the bounding-sphere test and
  // the law's erase test, which is all most interactions need.
  double* geo;
  // index of the body within its specimen (body.specimen; = the body id with one specimen)
  int32_t* lid;
  // |pos - refPos| rounded up: how far the body moved since the last collider run
  float* disp;
};

enum {
  GEO_POS = 0, GEO_RMAX = 3, GEO_BMIN = 4, GEO_BMAX = 7, GEO_G = 10, GEO_R = 19, GEO_E = 22,
  GEO_VEL = 24, GEO_ANGVEL = 27, GEO_MASS = 30, GEO_FLAGS = 31, GEO_STRIDE = 32
};
// Static per-body integers in one 64-bit word: material (low 32 bits), then blockedDOFs (8),
// kind (2), isSphere (1), wallAxis (2), wallSense + 1 (2), "among the first six bodies of
// its specimen" (1: the wall ids 0-5 the reference's Ip2/Law2 functors test for).
__host__ __device__ inline uint64_t geo_flags(int material, int blocked, int kind, int isSphere, int wallAxis,
                                              int wallSense, int lid) {
  return (uint64_t)(uint32_t)material | ((uint64_t)(blocked & 0xFF) << 32) | ((uint64_t)(kind & 3) << 40) |
         ((uint64_t)(isSphere & 1) << 42) | ((uint64_t)(wallAxis & 3) << 43) |
         ((uint64_t)((wallSense + 1) & 3) << 45) | ((uint64_t)(lid <= 5 ? 1 : 0) << 47);
}

struct PairArrays {
  PairRec* r;   // [cap] records; the first np are the interaction container in linIntrs order
  int64_t cap;
  // Per interaction, coalesced: the two ids (kept through every change of the container),
  // the positions of its two entries in the per-body lists (rebuilt with them), and the
  // separation margin of a settled apart pair (see k_pairs_prep), NaN when not settled.
  int32_t* cid;
  int32_t* cpos;
  float* gap;
};

struct MaterialDev {
  double density, Kn, Ks, frictionAngle, betan, betas, G, nu;
  double tanF;  // tan(frictionAngle), evaluated on the host by the C library the reference calls
};
