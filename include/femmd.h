/*
 * femmd.h -- C ABI of the MI355X-native femmd2D step engine.
 *
 * The engine owns device memory; the caller owns every host buffer passed
 * in or out.  Every call returns an int status (FM_OK = 0, negative on
 * error) and never exits the process; fm_last_error() gives the text.
 * One handle per host thread; each handle owns its own HIP stream.
 *
 * Particle indices in this API are the caller's: the order in which
 * particles were added (fm_add_particles, then fm_add_elastomer nodes),
 * as in the reference (femmd2D/src/particles.cpp:201-231, :295-385).
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   engine -> compute ("operator API"):
 *     computeForcesNNGPUWrapper      femmd2D/src/interactions.h:77-86
 *     computeBondForcesGPUWrapper    femmd2D/src/interactions.h:88-94
 *     computeFeaForcesGPUWrapper     femmd2D/src/interactions.h:96-104
 *     integratorStep1Wrapper         femmd2D/src/integrator.h:39
 *     integratorStep2Wrapper         femmd2D/src/integrator.h:41
 *     NeighborList::buildNNGPU       femmd2D/src/nn.h:74
 *   host API (users):
 *     MDSystem / c_* / System.*      femmd2D/src/md3dsystem.h:14-116,
 *                                    femmd2D/src/interface.cpp:24-81,656-696
 */
#ifndef FEMMD_H
#define FEMMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fm_system fm_system; /* one periodic 2D system */
typedef struct fm_batch fm_batch;   /* many independent replicas */

/* ---- status codes ---- */
#define FM_OK 0
#define FM_ERR_ARG (-1)      /* invalid argument                         */
#define FM_ERR_STATE (-2)    /* call not valid in the current state      */
#define FM_ERR_HIP (-3)      /* HIP runtime failure                      */
#define FM_ERR_DEVICE (-4)   /* device-side sticky error, see below      */
#define FM_ERR_GEOMETRY (-5) /* box too small for the cell grid, etc.    */
#define FM_ERR_NOMEM (-6)

/* ---- device error bits (sticky until fm_clear_errors) ---- */
#define FM_DEV_CELL_OVERFLOW 1u  /* more particles in a cell than capacity */
#define FM_DEV_TILE_OVERFLOW 2u  /* a cell tile did not fit in LDS         */
#define FM_DEV_INVERTED 4u       /* an element with det F <= 0             */
#define FM_DEV_NONFINITE 8u      /* a non-finite coordinate after a drift  */
#define FM_DEV_LIST_OVERFLOW 16u /* more neighbours than the list capacity */
#define FM_DEV_SORT_STALL 32u    /* the counting sort's look-back timed out */

/* ---- material models (femmd2D/src/interactions.cpp:411-689) ---- */
#define FM_NEO_HOOKEAN 0 /* calcFeaForcesCPU2, svk = 0 (reference CPU) */
#define FM_SVK 1         /* calcFeaForcesCPU2, svk = 1                 */
#define FM_LEGACY_GPU 2  /* computeFeaForcesGPU energy (= 2 x SVK)     */

typedef struct fm_params {
    /* periodic box (MDSystem::init, md3dsystem.cpp:49-80); held in FP32
     * like the reference's Box (a real=float vector, def.h:86), so every
     * wrap and image shift uses (float)Lx, (float)Ly */
    double Lx, Ly;
    double dt;      /* time step; reference default 0.005                  */
    int32_t device; /* HIP device ordinal                                  */
    int32_t reserved[7];
} fm_params;

/* Energies and observables, FP64 sums (femmd2D/src/properties.cpp:15-36,
 * femmd2D/src/md3dsystem.cpp:151-179). */
typedef struct fm_energies {
    double ke;          /* 1/2 sum m v^2                                   */
    double u_lj, w_lj;  /* pair energy and virial sum f_ij . r_ij          */
    double u_bond, w_bond;
    double u_fea, w_fea;
    double vv_max;      /* max |v|^2                                       */
    double temperature; /* KE / N (2D, kB = 1)                             */
    double pressure;    /* 0.5 sum m v^2 / V + 0.5 (w_lj + w_fea) / V      */
    int64_t steps;
    int64_t n;
} fm_energies;

/* Running averages of the per-step observables since the last reset
 * (Properties::accumulate / Prop::avg, femmd2D/src/properties.cpp:39-86,
 * femmd2D/src/properties.h:6-13).  Quantity k of mean[] / stdv[]:
 *   0 uKinetic = KE / N          1 uPotential = (U_lj + U_fea + U_bond) / N
 *   2 uTotal                     3 pressure (as fm_energies.pressure)
 *   4 virial = W_lj              5 feaVirial = W_fea
 *   6 uBonded = U_bond / N       7 uFea = U_fea / N
 * (temperature = uKinetic in 2D with kB = 1, properties.cpp:21). */
#define FM_AVG_UKINETIC 0
#define FM_AVG_UPOTENTIAL 1
#define FM_AVG_UTOTAL 2
#define FM_AVG_PRESSURE 3
#define FM_AVG_VIRIAL 4
#define FM_AVG_FEAVIRIAL 5
#define FM_AVG_UBONDED 6
#define FM_AVG_UFEA 7
typedef struct fm_averages {
    int64_t count;  /* steps accumulated                                   */
    double mean[8]; /* Prop::avgval                                        */
    double stdv[8]; /* Prop::stdv = sqrt(max(sum2 / n - avg^2, 0))         */
} fm_averages;

/* Number of HIP devices visible (md_device_init, femmd2D/src/CudaUtil.cu:45-102). */
int fm_device_count(int32_t *n);

/* ---- system construction (replaces MDSystem::init and friends) ---- */
int fm_create(const fm_params *p, fm_system **out);
void fm_destroy(fm_system *h);
const char *fm_last_error(const fm_system *h); /* never NULL */

/* Particles::addParticle (femmd2D/src/particles.cpp:201-231), vectorised.
 * vx, vy may be NULL (zero).  type may be NULL (type 0). */
int fm_add_particles(fm_system *h, int32_t n, const float *x, const float *y,
                     const float *vx, const float *vy, const float *mass,
                     const float *sigma, const int32_t *type,
                     int32_t *first_id);

/* Particles::addElastomer (femmd2D/src/particles.cpp:295-385).
 * verts: 2*nv coordinates relative to (cx, cy); tris: 3*nt indices with
 * index_base (1 for the reference mesh files).  Node masses: node_mass if
 * > 0, else total_mass spread as rho*A/3 per corner.  exclude != 0 puts
 * the nodes in a fresh exclusion group (no LJ among themselves).
 * E_per_tri / nu_per_tri may be NULL (global values of fm_set_material).
 * Unlike the reference, no edge bonds are created implicitly. */
int fm_add_elastomer(fm_system *h, int32_t nv, const float *verts, int32_t nt,
                     const int32_t *tris, int32_t index_base, float cx,
                     float cy, int32_t type, float sigma, float node_mass,
                     float total_mass, int32_t exclude, const float *E_per_tri,
                     const float *nu_per_tri, int32_t *first_node);

/* Bonds::createBond (femmd2D/src/bonds.cpp:14-29), per-bond k and r0;
 * r0 < 0 means the current (minimum-image) distance. */
int fm_add_bonds(fm_system *h, int32_t nb, const int32_t *i, const int32_t *j,
                 const float *k, const float *r0);

/* The edge bonds of Particles::addElastomer (femmd2D/src/particles.cpp:
 * 335-337: Bonds::createBond on the three edges of every triangle, r0 =
 * the current distance, an edge shared by two triangles bonded twice).
 * Opt-in here (fm_add_elastomer creates none): tris / index_base as given
 * to fm_add_elastomer, first_node its result; every bond gets stiffness k;
 * dedup != 0 bonds each edge once.  Edges are taken in the reference's
 * order (i0-i1, i0-i2, i1-i2 per triangle). */
int fm_add_edge_bonds(fm_system *h, int32_t first_node, int32_t nt, const int32_t *tris,
                      int32_t index_base, float k, int32_t dedup);

/* Interactions::setPairForce (femmd2D/src/interactions.cpp:65-79): row-major
 * ntypes x ntypes tables (ntypes in [1, 16]); rcut in sigma units; ushift
 * additive. */
int fm_set_pair_table(fm_system *h, int32_t ntypes, const float *eps,
                      const float *rcut, const float *ushift);

/* setElasticConstants / set_youngs_modulus
 * (femmd2D/src/interactions.cpp:43-50, femmd2D/src/interface.cpp:479-490). */
int fm_set_material(fm_system *h, int32_t model, float E, float nu);
int fm_set_dt(fm_system *h, double dt);
int fm_set_lj(fm_system *h, int32_t on);
/* Exclusion group of one particle (Particles::insertPartInExclusionGroup,
 * femmd2D/src/particles.cpp:289-293); -1 = none.  Before fm_commit. */
int fm_set_exclusion(fm_system *h, int32_t i, int32_t group);

/* Neighbour-list skin (default 0.6; the reference's CPU list uses 0.4,
 * femmd2D/src/nn.h:63), before fm_commit.
 * skin > 0: pair forces run over Verlet lists of radius rc*sigma + skin,
 * rebuilt (binning, counting sort, list build) only when the running sum of
 * per-step maximum displacements exceeds skin / 2 (femmd2D/src/nn.cpp:
 * 114-134), so most steps skip the sort.  skin = 0: the cells are walked
 * every step.  Typed pair tables (ntypes > 1) always walk the cells. */
int fm_set_skin(fm_system *h, float skin);
/* Internal order (no reference counterpart: the reference keeps particles
 * in their insertion order, femmd2D/src/particles.cpp:201-231).  The engine
 * stores particles in cell order; as they diffuse, that order decorrelates
 * from the cells and the step's memory accesses scatter (a fully shuffled
 * order measured 36 % slower on the 1.1 M-particle system).  Every
 * `steps` steps of fm_step (default 1000; 0 = never) the particles that no
 * element, bond or cluster refers to are given new internal ids in the
 * current cell order, on the same internal slots.  Results are unchanged up
 * to FP32 summation order within a cell; caller ids never change.
 * fm_set_reorder_interval: any time.  fm_reorder: re-sort now (after
 * fm_commit).  fm_reorder_info: the interval and the re-sorts so far. */
int fm_set_reorder_interval(fm_system *h, int32_t steps);
int fm_reorder(fm_system *h);
int fm_reorder_info(fm_system *h, int32_t *interval, int64_t *count);

/* Masses of the given particles (Particles::mass, femmd2D/src/particles.h:56).
 * A zero, negative or infinite mass pins the particle (it never moves).
 * Before fm_commit. */
int fm_set_mass(fm_system *h, int32_t n, const int32_t *ids, const float *mass);

/* Upload, topology build, validation (allocate_cuda + init_neighbor_list,
 * femmd2D/src/md3dsystem.h:86-94, femmd2D/src/nn.cpp:192-250). */
int fm_commit(fm_system *h);

/* ---- stepping (Integrator::run / doStep / calcForces) ---- */
int fm_step(fm_system *h, int32_t nsteps);     /* c_run, System.run       */
int fm_compute_forces(fm_system *h);           /* c_calcForces            */
/* Energies and observables of the current state: after fm_step, those of
 * its last step (formed on the device by the step itself). */
int fm_get_energies(fm_system *h, fm_energies *e);
/* Per-step observables (Integrator::doStep -> Properties::compute,
 * femmd2D/src/integrator.cpp:18, femmd2D/src/properties.cpp:88-98): with
 * every = 1 (the default, the reference's cadence) each step of fm_step
 * forms its energies, virials, kinetic energy and pressure on the device and
 * adds them to running sums; with every = 0 only the last step of each
 * fm_step call is formed and accumulated (the others skip the energy and
 * virial sums; forces and trajectories are the same bits either way). */
int fm_set_observe_every(fm_system *h, int32_t every);
int fm_get_averages(fm_system *h, fm_averages *a);
/* Properties::accumulate(0): zero the running sums. */
int fm_reset_averages(fm_system *h);
int fm_get_state(fm_system *h, float *x, float *y, float *vx, float *vy);
int fm_set_state(fm_system *h, const float *x, const float *y, const float *vx,
                 const float *vy);
int fm_get_forces(fm_system *h, float *fx, float *fy);
int fm_num_particles(const fm_system *h, int32_t *n); /* c_getNumberOfParticles */
/* c_getNumberOfBonds / c_getBondsPtr (femmd2D/src/interface.cpp:42,45): the
 * bonds as (i, j) pairs of caller ids, 2 * nb ints. */
int fm_num_bonds(const fm_system *h, int32_t *nb);
int fm_get_bonds(const fm_system *h, int32_t *pairs);
/* Number of CST elements (triangles added by fm_add_elastomer). */
int fm_num_elements(const fm_system *h, int32_t *nt);
int fm_sync(fm_system *h);

/* Launch tuning: cells per tile row and threads per workgroup of the force
 * pass (0 = keep), graph capture of the step loop on/off. */
int fm_set_tuning(fm_system *h, int32_t tile_cells, int32_t block,
                  int32_t use_graph);

/* ---- diagnostics used by the parity tests ---- */
uint32_t fm_device_errors(fm_system *h);
int fm_clear_errors(fm_system *h);
/* Cell grid of the last binning.  key: n, caller order.  start: ncells+1
 * and perm: n (caller ids in sorted order), either may be NULL. */
int fm_get_cells(fm_system *h, int32_t *ncx, int32_t *ncy, int32_t *key,
                 int32_t *start, int32_t *perm);
/* Interacting sets at the current positions: per particle the count, the
 * sum of (j+1) and the sum of (j+1)^2 over its interacting partners j. */
int fm_pair_sets(fm_system *h, int32_t *cnt, int64_t *s1, int64_t *s2);
/* Time one force pass on the device (ms per pass, averaged over reps). */
int fm_time_forces(fm_system *h, int32_t reps, double *ms);
/* Advance nsteps + 1 steps, timing nsteps of them kernel by kernel (no
 * graph): each kernel's dispatch start and end timestamps, taken by HIP
 * events armed on its launch (hipExtLaunchKernel).  ms[6] = per step: kick
 * (with the binning when cells are walked every step), binning (list mode),
 * counting sort, list build, force pass (0 for a kernel the mode does not
 * launch), whole step from the kick's start to the force pass's end. */
int fm_profile_step(fm_system *h, int32_t nsteps, double *ms);
/* Elastomer clusters (one per fm_add_elastomer call, in call order). */
int fm_cluster_count(fm_system *h, int32_t *nc);
/* Cluster observables at the current state (FeaElements::calcClusterProps,
 * femmd2D/src/feaElements.cpp:86-196; MDSystem::evalProps cluster part,
 * femmd2D/src/md3dsystem.cpp:168-174).  props[7 * c + k], k = mass, centre
 * of mass x, y (of the unfolded nodes: a cluster wider than half the box has
 * its nodes beyond L/2 moved down by L, all of it moved up by L if their mean
 * is then negative), centre-of-mass velocity x, y, centroid x, y (circular
 * mean of the wrapped positions).  ux, uy (caller order, may be null): the
 * unfolded positions of cluster nodes, the wrapped position of any other
 * particle.  The cluster kinetic energy sum is sum_c mass_c |v_cm,c|^2. */
int fm_cluster_props(fm_system *h, double *props, float *ux, float *uy);
/* Cluster virial (femmd2D/src/interactions.cpp:370-382): over the
 * interacting pairs (i, j) of particles that both belong to clusters,
 * sum of 1/2 f_ij . (R_c(i) - R_c(j)), minimum image.  Pairs with a particle
 * outside every cluster add nothing (the reference indexes cluster -1 for
 * them).  The molecular pressure is 0.5 (cluster KE sum - virial) / V. */
int fm_cluster_virial(fm_system *h, double *virial);
/* Velocity rescale to temperature temp (MDSystem::setTemperature,
 * femmd2D/src/md3dsystem.cpp:92-136): n = 0, all particles, with
 * T = sum(m v^2 / 2) / N; n > 0, the particles ids[0..n), with the
 * reference's group formula T = 2 sum(m v^2 / 2) / (3 N), N the whole
 * system's count.  v *= sqrt(temp / T). */
int fm_set_temperature(fm_system *h, double temp, int32_t n, const int32_t *ids);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (a thermostat inside a tight serving loop, harnesses that forbid
 * hipMalloc): fm_set_temperature with the same arguments and the same result
 * bits, calling no hipMalloc / hipFree.  Its scratch (selected ids, kinetic
 * energy partials) lives in one file-scope device array of 4.5 MiB that each
 * GPU the library runs on holds until the process ends.  Limits: n <= 2^20
 * ids; with n = 0, N <= 2^24 particles; larger requests fail with FM_ERR_ARG.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from several threads or handles) race. */
int fm_set_temperature_noalloc(fm_system *h, double temp, int32_t n, const int32_t *ids);
/* Neighbour-list statistics: mean and max list length at the last build,
 * rebuilds so far (all zero when the cells are walked every step). */
int fm_list_stats(fm_system *h, double *mean_len, int32_t *max_len,
                  int64_t *rebuilds);

/* ---- adjoint (no reference counterpart; single-type systems) ----
 * Directional derivative of the forces at the current positions: for a
 * per-particle vector w (caller order), j = (df/dx) w, and the parameter
 * terms dE[e] = w . df/dE_e (caller element order), deps = w . df/deps,
 * dsig = w . df/dsigma for the common pair eps and sigma.  Any output may
 * be NULL.  The state is left unchanged. */
int fm_force_jvp(fm_system *h, const float *wx, const float *wy, float *jx,
                 float *jy, double *dE, double *deps, double *dsig);
/* Discrete adjoint of nsteps KDK steps from the current state: runs them
 * forward (the state advances as with fm_step), evaluates
 * loss = (1/K) sum_k |x_{ids[k]} - (tx[k], ty[k])|^2 at the end (minimum
 * image, K = nloss) and sweeps backward, recomputing positions from
 * checkpoints taken every ckpt steps (ckpt <= 0: as many stored steps as a
 * quarter of the free device memory holds, the whole trajectory when it
 * fits).  Returns loss, dL/dE per element (caller order) and dL/deps,
 * dL/dsigma; gradient outputs may be NULL. */
int fm_adjoint(fm_system *h, int32_t nsteps, int32_t ckpt, int32_t nloss,
               const int32_t *ids, const float *tx, const float *ty,
               double *loss, double *dE, double *deps, double *dsig);

/* The adjoint for an arbitrary loss of the final positions, in two calls:
 * fm_adjoint_forward runs nsteps steps from the current state (the state
 * advances as with fm_step), keeping what the backward sweep needs (the
 * stored trajectory, or checkpoints every ckpt steps as fm_adjoint chooses
 * them); fm_adjoint_backward then takes the terminal cotangent dL/dx_N,
 * dL/dy_N (caller order, n floats each; NULL = zero) and returns dL/dE per
 * element (caller order), dL/deps and dL/dsigma, FP64.  The state must not
 * change in between (FM_ERR_STATE otherwise); the backward call consumes
 * the recording.  fm_adjoint_run is both in one call, for a cotangent known
 * beforehand. */
int fm_adjoint_forward(fm_system *h, int32_t nsteps, int32_t ckpt);
int fm_adjoint_backward(fm_system *h, const float *dL_dx, const float *dL_dy, double *dE,
                        double *deps, double *dsig);
int fm_adjoint_run(fm_system *h, int32_t nsteps, int32_t ckpt, const float *dL_dx,
                   const float *dL_dy, double *dE, double *deps, double *dsig);

/* ---- batched replicas (no reference counterpart) ---- */
int fm_batch_create(int32_t n, fm_system *const *systems, fm_batch **out);
void fm_batch_destroy(fm_batch *b);
const char *fm_batch_last_error(const fm_batch *b);
int fm_batch_step(fm_batch *b, int32_t nsteps);
int fm_batch_get_state(fm_batch *b, int32_t r, float *x, float *y, float *vx,
                       float *vy);
int fm_batch_get_energies(fm_batch *b, int32_t r, fm_energies *e);
int fm_batch_get_averages(fm_batch *b, int32_t r, fm_averages *a);
int fm_batch_set_tuning(fm_batch *b, int32_t tile_cells, int32_t block,
                        int32_t use_graph);
/* Autotune tile width and workgroup size per replica-size bucket; results
 * are cached in cache_path (a text file, one line per batch signature) when
 * it is not NULL. */
int fm_batch_tune(fm_batch *b, const char *cache_path);
/* Re-sort interval of the batch's internal order (as fm_set_reorder_interval). */
int fm_batch_set_reorder_interval(fm_batch *b, int32_t steps);
uint32_t fm_batch_device_errors(fm_batch *b);

#ifdef __cplusplus
}
#endif
#endif
