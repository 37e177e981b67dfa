/*
 * sdm_oracle.h -- CPU oracle of the superellipsoid DEM step (test infrastructure only).
 *
 * This is the checker, not the product: a scalar FP64 restatement of the SudoDEM3D
 * per-step loop that only tests/, the smoke check and the CPU-baseline leg of bench.py
 * use.  It exposes the same ABI as include/sdm.h with the prefix sdmcpu_ instead of sdm_,
 * plus two oracle-only switches.
 */
#ifndef SDM_ORACLE_H
#define SDM_ORACLE_H

#include "../include/sdm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdm_ctx sdmcpu_ctx;

const char* sdmcpu_version(void);
sdm_ctx* sdmcpu_create(const sdm_params* params, int32_t n_bodies, int32_t n_materials);
void sdmcpu_destroy(sdm_ctx* ctx);
const char* sdmcpu_last_error(const sdm_ctx* ctx);
int sdmcpu_set_materials(sdm_ctx* ctx, const sdm_material* mats, int32_t n);
int sdmcpu_superquadric_props(double rx, double ry, double rz, double eps1, double eps2,
                              double* r_max, double* volume, double inertia[3]);
int sdmcpu_field_info(sdm_ctx* ctx, const char* name, int32_t* dtype, int64_t* count);
int sdmcpu_get(sdm_ctx* ctx, const char* name, void* dst, int64_t count);
int sdmcpu_set(sdm_ctx* ctx, const char* name, const void* src, int64_t count);
const char* sdmcpu_field_list(void);
int sdmcpu_step(sdm_ctx* ctx, int32_t nsteps);
int sdmcpu_sync(sdm_ctx* ctx);
int sdmcpu_get_counters(sdm_ctx* ctx, sdm_counters* out);
int sdmcpu_reset_counters(sdm_ctx* ctx);
int sdmcpu_set_timing(sdm_ctx* ctx, int32_t on);
int sdmcpu_set_graph(sdm_ctx* ctx, int32_t on);
int sdmcpu_set_tuning(sdm_ctx* ctx, const sdm_tuning* t);
int sdmcpu_mean_coordination(sdm_ctx* ctx, double* cn);
int sdmcpu_stress(sdm_ctx* ctx, double volume, double sigma[9]);
int sdmcpu_body_forces(sdm_ctx* ctx, const int32_t* ids, int32_t n, double* forces);
int sdmcpu_set_servo(sdm_ctx* ctx, const sdm_servo* servo);
int sdmcpu_get_servo(sdm_ctx* ctx, sdm_servo_state* out);
int sdmcpu_get_tuning(sdm_ctx* ctx, sdm_tuning* t);
int sdmcpu_get_timers(sdm_ctx* ctx, double out[8]);

/* Number of OpenMP threads used by the engine loops (1 = the reference's default -j1,
 * serial semantics).  With more than one thread the interaction loop and the body loops
 * run in parallel with per-thread force buffers summed in thread order, as
 * core/ForceContainer.hpp:87,134-162 does. */
int sdmcpu_set_threads(sdm_ctx* ctx, int32_t nthreads);

/* Initial collider sort: 0 = std::sort with Bounds::operator< (what the reference runs,
 * InsertionSortCollider.cpp:339), 1 = std::stable_sort with the same comparator.  The two
 * differ only in the placement of bounds that compare equal (the +-inf bounds of walls). */
int sdmcpu_set_init_sort(sdm_ctx* ctx, int32_t stable);
/* The collider's bound ring of one axis: coordinates, ids, min flags, periods (2N each), loIdx. */
int sdmcpu_dump_bounds(sdm_ctx* ctx, int32_t axis, double* coord, int32_t* id, int32_t* isMin, int32_t* period,
                       int64_t* loIdx);

/* Distance evaluation used by the narrowphase tests: the support-map point of one
 * superellipsoid for a local unit direction (Normal2Phi + getSurface). */
int sdmcpu_support_point(double rx, double ry, double rz, double eps1, double eps2,
                         const double n_local[3], double phi_out[2], double p_out[3]);
/* Principal curvatures at surface parameters phi (Superquadrics::Curvatures). */
int sdmcpu_curvatures(double rx, double ry, double rz, double eps1, double eps2,
                      const double rot_global[9], const double phi[2], double k_out[2],
                      double dir_max[3], double dir_min[3], int32_t* max_flag);
/* SuperquadricsGeom2::HertzMindlin on given curvatures: out = {curvatures_sum, alpha, K, E}. */
int sdmcpu_hertz_mindlin(const double curv1[2], const double curv2[2], double theta, double out[4]);

#ifdef __cplusplus
}
#endif

#endif
