/*
 * sdm.h -- C ABI of the MI355X superellipsoid DEM step library (libsdm_hip.so).
 *
 * One context owns one scene (bodies, interactions, collider state) resident in
 * HBM and advances it with the engine list
 *
 *   ForceResetter()
 *   InsertionSortCollider([Bo1_Superquadrics_Aabb(), Bo1_Wall_Aabb()], verletDist=...)
 *   InteractionLoop([Ig2_Wall_Superquadrics_SuperquadricsGeom(2)(),
 *                    Ig2_Superquadrics_Superquadrics_SuperquadricsGeom(2)()],
 *                   [Ip2_SuperquadricsMat(2)_...()], [SuperquadricsLaw(2)()])
 *   NewtonIntegrator(damping=..., gravity=..., isSuperquadrics=1)
 *
 * which is the canonical SudoDEM3D engine list (examples/SudoDEM3D/example1.py:105-116).
 * A reference engine would call this ABI once per step, never per pair:
 *   - GlobalEngine::action() of each engine (core/Engine.hpp:34, core/GlobalEngine.hpp:13-17)
 *     maps onto sdm_step(), which runs all four engines of one Scene::moveToNextTimeStep
 *     (core/Scene.cpp:70-138);
 *   - the Python attribute system (SUDODEM_CLASS_BASE_DOC_ATTRS_*,
 *     lib/serialization/Serializable.hpp:207-223) maps onto sdm_set()/sdm_get() with the
 *     same attribute names ("state.pos", "geom.PenetrationDepth", "phys.shearForce", ...).
 *
 * Conventions: every int return is 0 on success and negative on failure; the message of
 * the last failure is returned by sdm_last_error(). No C++ exception crosses the ABI.
 * Host buffers belong to the caller; device buffers belong to the context. A context is
 * used by one host thread at a time; calls are stream ordered and sdm_get() synchronises.
 */
#ifndef SDM_H
#define SDM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdm_ctx sdm_ctx;

enum {
  SDM_LAW_LINEAR = 0, /* Ig2_*_SuperquadricsGeom (Levenberg-Marquardt) + Ip2_SuperquadricsMat_* + SuperquadricsLaw */
  SDM_LAW_HERTZ = 1   /* Ig2_*_SuperquadricsGeom2 (Nelder-Mead) + Ip2_SuperquadricsMat2_* + SuperquadricsLaw2 */
};

enum {
  SDM_FP64 = 0,  /* all arithmetic in double, reference semantics (pow, atan2, atan, sin and cos
                    correctly rounded, the values glibc gives the reference) */
  SDM_MIXED = 1  /* Nelder-Mead iterations in float on relative frames, final contact refinement in double;
                    Levenberg-Marquardt in double with the device's own functions */
};

enum { SDM_BODY_SUPERQUADRIC = 0, SDM_BODY_WALL = 1 };

/* Engine attributes.  Each field names the reference attribute it carries. */
typedef struct sdm_params {
  double dt;                    /* Scene::dt */
  double gravity[3];            /* NewtonIntegrator.gravity (pkg/dem/NewtonIntegrator.hpp:73) */
  double damping;               /* NewtonIntegrator.damping (NewtonIntegrator.hpp:70) */
  int32_t exact_aspherical_rot; /* NewtonIntegrator.exactAsphericalRot (NewtonIntegrator.hpp:75) */
  double verlet_dist;           /* InsertionSortCollider.verletDist, >= 0 (InsertionSortCollider.hpp:217) */
  int32_t target_interv;        /* InsertionSortCollider.targetInterv (InsertionSortCollider.hpp:214) */
  double min_sweep_dist_factor; /* InsertionSortCollider.minSweepDistFactor (InsertionSortCollider.hpp:218) */
  int32_t sort_axis;            /* InsertionSortCollider.sortAxis (InsertionSortCollider.hpp:211) */
  int32_t law;                  /* SDM_LAW_* : which Ig2/Ip2/Law2 functor family the InteractionLoop holds */
  int32_t precision;            /* SDM_FP64 or SDM_MIXED */
  int32_t periodic;             /* Scene::isPeriodic: an axis-aligned periodic cell [0, cell) (core/Scene.hpp) */
  double cell[3];               /* Cell size (Cell::getSize, core/Cell.hpp); used when periodic */
} sdm_params;

/* SuperquadricsMat / SuperquadricsMat2 (pkg/dem/Superquadrics.hpp:247-277) + Material.density */
typedef struct sdm_material {
  double density, Kn, Ks, frictionAngle, betan, betas, G, nu;
} sdm_material;

/* Per-run counters (device side, summed over the steps since the last reset). */
typedef struct sdm_counters {
  int64_t steps;          /* engine-loop iterations run */
  int64_t collider_runs;  /* InsertionSortCollider::action executions */
  int64_t contact_evals;  /* Ig2 invocations on superquadric pairs (one per pair-step) */
  int64_t dist_evals;     /* Disfunction / Disfun evaluations inside the narrowphase solvers */
  int64_t pairs_inserted; /* interactions appended by the collider */
  int64_t pairs_erased;   /* interactions erased by the collider */
  int64_t graph_steps;    /* steps replayed from the captured step graph (sdm_set_graph) */
} sdm_counters;

/* Version string of the library. */
const char* sdm_version(void);

/* Create a context for n_bodies bodies and n_materials materials on the current device.
 * Replaces the construction of Scene + O.engines (py/wrapper/sudodemWrapper.cpp:929-991).
 * Returns NULL on failure (message through sdm_last_error(NULL)). */
sdm_ctx* sdm_create(const sdm_params* params, int32_t n_bodies, int32_t n_materials);
void sdm_destroy(sdm_ctx* ctx);
const char* sdm_last_error(const sdm_ctx* ctx);

/* Material table (O.materials). */
int sdm_set_materials(sdm_ctx* ctx, const sdm_material* mats, int32_t n);

/* Geometric properties of a superellipsoid, as Superquadrics::Initial computes them
 * (pkg/dem/Superquadrics.cpp:47-93): r_max, volume and principal inertia per unit density. */
int sdm_superquadric_props(double rx, double ry, double rz, double eps1, double eps2,
                           double* r_max, double* volume, double inertia[3]);
/* The same for n bodies (semi-axes rxyz[3n], exponents eps[2n]; out r_max[n], volume[n],
 * inertia[3n]): the body-construction loop of a scene script (NewSuperquadrics per body,
 * py/_superquadrics_utils.cpp:354-379) in one call.  Bit for bit the single-body values. */
int sdm_superquadric_props_n(int64_t n, const double* rxyz, const double* eps, double* r_max, double* volume,
                             double* inertia);

/* Attribute access.  `name` is one of the attribute names listed by sdm_field_list(); the
 * element type (SDM_F64 / SDM_I32 / SDM_I64) and the number of elements (count) are fixed by
 * the attribute and the current number of bodies or interactions.  Vector attributes are
 * stored row-major, one row per body or interaction ("state.pos" holds x,y,z of body 0,
 * then of body 1, ...).  Setting "intr.id1" resizes the interaction container to count.
 * Attribute reads synchronise with the device. */
enum { SDM_F64 = 0, SDM_I32 = 1, SDM_I64 = 2 };
int sdm_field_info(sdm_ctx* ctx, const char* name, int32_t* dtype, int64_t* count);
int sdm_get(sdm_ctx* ctx, const char* name, void* dst, int64_t count);
int sdm_set(sdm_ctx* ctx, const char* name, const void* src, int64_t count);
/* Newline separated list of the attribute names. */
const char* sdm_field_list(void);

/* Run nsteps iterations of the engine list (Scene::moveToNextTimeStep, core/Scene.cpp:70-138). */
int sdm_step(sdm_ctx* ctx, int32_t nsteps);
/* Block until all queued device work of the context is done. */
int sdm_sync(sdm_ctx* ctx);

int sdm_get_counters(sdm_ctx* ctx, sdm_counters* out);
int sdm_reset_counters(sdm_ctx* ctx);

/* Per-phase device timers (HIP events recorded on the context's stream around each phase),
 * the equivalent of the reference's per-engine TimingInfo (core/Timing.hpp:5-23,
 * core/Scene.cpp:90-97).  Off by default; sdm_set_timing(ctx, 1) resets and starts them.
 * sdm_get_timers fills 8 values summed since then:
 *   out[0] collider [ms]            out[1] interaction loop [ms]   out[2] Newton + force sums [ms]
 *   out[3] timed steps              out[4] narrowphase solver kernel [ms]
 *   out[5] distance evaluations of the solver kernel
 *   out[6] distance evaluations of the restart / refinement kernels
 *   out[7] interactions handed to the solver kernel (summed over the timed steps) */
int sdm_set_timing(sdm_ctx* ctx, int32_t on);
int sdm_get_timers(sdm_ctx* ctx, double out[8]);
/* Per-kernel device times under the same timing switch (events around each launch or group
 * of launches): out[2k] milliseconds and out[2k+1] launches summed since sdm_set_timing(ctx, 1),
 * for k = 0 the integrator (NewtonIntegrator), 1 the ordered force sums (ForceContainer), 2 the
 * bound update of a collider run (BoundDispatcher), 3 the interaction loop's dispatch pass,
 * 4 the solver queue (flag scan and fill), 5 the wall and sphere pairs, 6 the Nelder-Mead first
 * vertex, 7 the Nelder-Mead solver list (scan, fill, cost sort), 8 the main solver launch,
 * 9 the mixed form's resume launch (the Levenberg-Marquardt restarts run inside the main
 * solver launch), 10 the mixed form's
 * double refinement, 11 the finish (curvature, stiffness, force law, stores), 12 the periodic
 * collider's ring walk in its sequential phases (with the run's coordinates), 13 its parallel
 * phases (window checks, sorts, crossings); out[28] the Nelder-Mead pairs touching at their
 * first vertex, out[29] the mixed form's solver evaluations done in double (Nelder-Mead or
 * Levenberg-Marquardt, anchors included), out[30] the pairs handed to the resume launch
 * (summed over the timed steps; out[28] and out[30] are 0 outside the mixed Hertz form),
 * out[31] the launches that extended a sequential phase of the periodic ring walk after a
 * parallel phase's check failed.  The reference's TimingInfo has no per-kernel split. */
int sdm_get_kernel_timers(sdm_ctx* ctx, double out[32]);

/* Steady-state steps as a replayed graph.  No counterpart in the reference (its engines are
 * called one by one).  When on (the default), a step without a collider run, host write or
 * timing since the previous one replays a captured hipGraph of the interaction loop, the force
 * sums and the integration instead of launching ~20 kernels; results are identical.  The
 * graph is recaptured after any reallocation or parameter change. */
int sdm_set_graph(sdm_ctx* ctx, int32_t on);

/* Observables of the current state, reduced on the device.
 * sdm_mean_coordination: Shop__meanCoordinationNumber (py/_utils.cpp:522-539): twice the real
 *   interactions with a non-zero normal force over the bodies that are not walls.
 * sdm_stress: Shop::getStress (pkg/dem/Shop_02.cpp:350-377), the Love-Weber stress
 *   sum over real interactions of (normalForce + shearForce) (x) (pos1 - pos2 - hSize cellDist)
 *   over `volume` (row-major 3x3; volume 0 = the cell volume if periodic, else 1).  Summed in a
 *   fixed block tree: deterministic, equal to the reference's serial sum to rounding.
 * sdm_body_forces: ForceContainer::getForce (core/ForceContainer.hpp) of `n` bodies after the
 *   last step (x, y, z each): the wall forces CompressionEngine turns into wall stresses
 *   (pkg/dem/Compression.cpp:270-275) without reading back every body. */
int sdm_mean_coordination(sdm_ctx* ctx, double* cn);
int sdm_stress(sdm_ctx* ctx, double volume, double sigma[9]);
int sdm_body_forces(sdm_ctx* ctx, const int32_t* ids, int32_t n, double* forces);

/* sdm_stress without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and result as sdm_stress (bit for bit); the
 * block partials live in a static device array of 9 MiB per GPU, held until the process ends,
 * which limits it to 33,554,432 interactions (larger containers: an error, use sdm_stress).
 * Every call on one GPU shares that array: calls that overlap on one GPU (several host threads
 * or streams) race. */
int sdm_stress_noalloc(sdm_ctx* ctx, double volume, double sigma[9]);

/* Wall stress servo: CompressionEngine (pkg/dem/Compression.cpp) on a cubic specimen as the
 * reference compiles it (NEW_SCHEME and SOLE_GAIN, Compression.hpp:26-29): action_cubic
 * (:395-604) with generalConsolidation (:324-345) or hydroConsolidation (:264-322) when z_servo,
 * the triaxial shear branch (:498-603) otherwise; servo (:1252-1316), get_gain (:1125-1220),
 * consol_ss (:1330-1345), getStress (:764-841), Shop::unbalancedForce
 * (pkg/dem/Shop_01.cpp:148-189).  Bodies 0-5 are the walls left, right,
 * front, back, bottom, top (x-, x+, y-, y+, z-, z+), as the engine assumes.  It runs inside
 * every step after the force sums and before the integrator (the engine order of
 * examples/SudoDEM3D/example2_consol.py and example2_shear.py), with no host round trip:
 *   consolidation (z_servo, not hydro_strain): wall stresses, each wall moved by its gain,
 *     gains every iterate_num_max steps (stiffness summed in interaction order) and the
 *     stability check every solve_num_max steps.  When the check finds the target reached
 *     and the unbalanced force below unbf_tol, the integrator quiets the system in that step
 *     (NewtonIntegrator::quiet_system_flag: velocities zeroed, no move);
 *   hydroConsolidation (z_servo and hydro_strain): every wall still below its goal moves in
 *     by min(0.1 gain deficit, hydro_strain_rate size dt); the check compares with the
 *     UnbalancedForce of the last echo and does not quiet the system;
 *   shear (z_servo = 0): the z walls close at the ramped rate, then at goalz height per unit
 *     time; the x and y walls keep the lateral stress (shear_mode bit 1) or the volume;
 *     the run ends once log(height0 / height) exceeds target_strain.
 * Completion returns sdm_step after that step (stopAtIter = iter + 1).  Every echo_interval
 * steps (Scene::iter a multiple of it) UnbalancedForce is recomputed and, in consolidation
 * and in constant-volume shear, the wall stresses are re-evaluated after the moves. */
typedef struct sdm_servo {
  double goal[3];           /* goalx, goaly, goalz: target stresses [Pa] */
  double max_vel;           /* max_vel: wall speed limit [m/s] */
  double gain_alpha;        /* gain_alpha */
  double f_threshold;       /* f_threshold: |mean stress - goalx| < goalx f_threshold ... */
  double fmin_threshold;    /* fmin_threshold: ... and deviator / mean < fmin_threshold ... */
  double unbf_tol;          /* unbf_tol: ... and unbalanced force < unbf_tol: completed */
  int32_t iterate_num_max;  /* iterate_num_max: steps between gain updates (100) */
  int32_t solve_num_max;    /* solve_num_max: steps between stability checks (2000) */
  int32_t activated[4];     /* left/right/front/back_wall_activated (the z walls follow z_servo) */
  int32_t z_servo;          /* z_servo: 1 consolidation (stress control in z), 0 triaxial shear */
  int32_t hydro_strain;     /* hydroStrain: consolidation by a bounded strain rate (hydroConsolidation) */
  double hydro_strain_rate; /* hydroStrainRate */
  int32_t shear_mode;       /* shearMode: bit 1 drained (lateral stress servo), else constant volume */
  int32_t ramp_interval;    /* ramp_interval: steps over which the axial rate ramps up */
  int32_t ramp_chunks;      /* ramp_chunks */
  int32_t echo_interval;    /* echo_interval: every so many steps UnbalancedForce and the wall
                               stresses are refreshed (the engine's console report) */
  double target_strain;     /* target_strain: shear ends at log(height0 / height) above it */
  int32_t keep_state;       /* 1: change the attributes of the running engine (as assigning them in
                               Python does), 0: a new engine (firstRun) */
  int32_t continue_flag;    /* continueFlag: the first run keeps size0 below instead of the box
                               it finds, and shear skips the ramp */
  double size0[3];          /* width0, depth0, height0: read when continue_flag */
} sdm_servo;
typedef struct sdm_servo_state {
  double wall_stress[6];    /* wsxx_left, wsxx_right, wsyy_front, wsyy_back, wszz_bottom, wszz_top */
  double stress[3];         /* wsxx, wsyy, wszz (means of the two walls of each axis) */
  double gain[6];           /* gain_x, gain_x1, gain_y, gain_y1, gain_z, gain_z1 */
  double stiffness[6];      /* avg_xstiff, avg_xstiff1, avg_ystiff, avg_ystiff1, avg_zstiff, avg_zstiff1 */
  double size[3];           /* width, depth, height at the last wall-stress evaluation */
  double mean_stress;       /* getStress() of the last stability check: mean ... */
  double deviator_stress;   /* ... and deviatoric stress (0 before the first check) */
  double unbalanced_force;  /* UnbalancedForce: of the last check that reached the target or the
                               last echo (1 before either) */
  int32_t iterate_num, solve_num;
  int64_t completed_iter;   /* Scene::iter of the step that completed consolidation or shear, -1 if none */
  double size0[3];          /* width0, depth0, height0: the sizes of the first run */
  int32_t ramp_num;         /* rampNum */
  int32_t pad_;
} sdm_servo_state;
/* servo = NULL switches the servo off.  Switching it on starts it afresh (firstRun) unless
 * keep_state asks to keep the state of the running one (e.g. z_servo = 0 after consolidation:
 * shear with the sizes the engine recorded on its first run). */
int sdm_set_servo(sdm_ctx* ctx, const sdm_servo* servo);
int sdm_get_servo(sdm_ctx* ctx, sdm_servo_state* out);

/* Launch shapes of the step (no counterpart in the reference).  None changes a result: they
 * only schedule the same per-interaction work differently.  sudodem_amd/autotune.py picks
 * them by timing a workload's own steps. */
typedef struct sdm_tuning {
  int32_t solver_blocks_per_cu; /* persistent solvers: blocks per CU (0 = the occupancy maximum) */
  int32_t sort_by_cost;         /* Nelder-Mead list sorted by expected cost, longest first (1) or not (0);
                                   default 1 for SDM_FP64, 0 for SDM_MIXED */
  int32_t misc_blocks_per_cu;   /* wall / sphere kernel: blocks per CU (>= 1) */
  int32_t graph;                /* steady-state steps replayed from a captured graph (sdm_set_graph) */
} sdm_tuning;
int sdm_set_tuning(sdm_ctx* ctx, const sdm_tuning* t);
int sdm_get_tuning(sdm_ctx* ctx, sdm_tuning* t);

/* ---------------------------------------------------------------------------------------
 * Differentiable rollout over a frozen contact set (linear law + NewtonIntegrator).
 * No counterpart in the reference: it differentiates, over a window of `steps` steps, the
 * force law of SuperquadricsLaw::go (pkg/dem/Superquadrics.cpp:1181-1363) and the leapfrog of
 * NewtonIntegrator (pkg/dem/NewtonIntegrator.cpp:99-360) with the contacts and their geometry
 * held at the window start (the narrowphase is not differentiated).  Loss = mean, over the
 * last `loss_steps` steps, of the summed normal-force magnitudes of the flagged contacts.
 * Every pointer in the problem and every array argument is DEVICE memory; the caller owns
 * all of it (checkpoint and work buffers of the sizes returned below), nothing is
 * allocated.  Calls run on the null stream and return after completion.
 * ------------------------------------------------------------------------------------- */
typedef struct {
  int32_t n_bodies, n_contacts;
  const double* pos0;         /* [n][3] positions at the window start */
  const double* mass;         /* [n] */
  const double* inv_inertia;  /* [n][9] world-frame inverse inertia at the window start (row-major) */
  const int32_t* dynamic;     /* [n] 1: integrated; 0: moves with wall_vel (walls) */
  const double* wall_vel;     /* [n][3] velocity of the non-dynamic bodies */
  const int32_t* id_a;        /* [m] first body (the wall for wall contacts) */
  const int32_t* id_b;        /* [m] */
  const double* normal;       /* [m][3] unit normal from a to b */
  const double* ra;           /* [m][3] contact point - centre of a */
  const double* rb;           /* [m][3] contact point - centre of b */
  const double* depth0;       /* [m] penetration depth at the window start */
  const double* fs0;          /* [m][3] shear force at the window start */
  const int32_t* in_loss;     /* [m] 1: the contact's normal force enters the loss */
  const int32_t* csr_off;     /* [n+1] per-body contact lists */
  const int32_t* csr_ent;     /* [2m] contact << 1 | side (0: body is a, 1: body is b), in contact order */
  double dt, damping, gravity[3];
  int32_t steps, loss_steps, checkpoint_every;
} sdm_frozen_problem;

/* Sizes (in doubles) of the checkpoint and work buffers. */
int64_t sdm_frozen_checkpoint_size(const sdm_frozen_problem* p);
int64_t sdm_frozen_work_size(const sdm_frozen_problem* p);
/* params = {kn, ks, mu} (device, 3 doubles); v0, w0: [n][3] start velocities and angular
 * velocities (entries of non-dynamic bodies are ignored).  Writes the loss (host). */
int sdm_frozen_forward(const sdm_frozen_problem* p, const double* params, const double* v0, const double* w0,
                       double* checkpoints, double* work, double* loss);
/* After sdm_frozen_forward with the same arguments: dLoss/d{kn, ks, mu} (host, 3 doubles)
 * and dLoss/dv0 ([n][3], device; zero for non-dynamic bodies). */
int sdm_frozen_backward(const sdm_frozen_problem* p, const double* params, const double* checkpoints, double* work,
                        double* grad_params, double* grad_v0);

#ifdef __cplusplus
}
#endif

#endif /* SDM_H */
