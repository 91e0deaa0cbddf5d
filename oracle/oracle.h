/*
 * oracle.h -- CPU restatement of femmd2D's per-step force-and-integrate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the reference the HIP engine's
 * results are tested against, and the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  Nothing in femmd_old_amd/
 * links, loads or falls back to it.
 *
 * It restates, in plain C++, the reference CPU path (real = float build,
 * DEV = CPU):
 *   - periodic wrap               femmd2D/src/def.h:112-129
 *   - minimum image               femmd2D/src/def.h:91-110
 *   - cell id / cell decode       femmd2D/src/nnCuda.cu:9-37
 *   - cell fill order             femmd2D/src/nn.cpp:8-20
 *   - 3x3 neighbour walk          femmd2D/src/nn.cpp:22-62 (all-pairs: :66-92)
 *   - LJ pair force               femmd2D/src/interactionsCuda.cu:43-100
 *   - pair loop, energy, virial   femmd2D/src/interactions.cpp:303-389
 *   - harmonic bonds              femmd2D/src/interactions.cpp:143-173
 *   - CST Neo-Hookean / SVK       femmd2D/src/interactions.cpp:532-689
 *   - legacy 2xSVK GPU model      femmd2D/src/interactionsCuda.cu:312-451
 *   - Lame constants              femmd2D/src/interactions.cpp:43-50
 *   - KDK "leapfrog"              femmd2D/src/integrator.cpp:8-29,100-135
 *   - KE, max v^2, observables    femmd2D/src/md3dsystem.cpp:151-179,
 *                                 femmd2D/src/properties.cpp:15-36
 *   - Verlet skin rebuild rule    femmd2D/src/nn.cpp:114-134,200-225
 *
 * Precision: the FP64 entry points evaluate forces and energies in double.
 * With pred_f32 = 1 the discrete decisions (minimum-image branch, cutoff
 * predicate) are taken in FP32 exactly as the reference's real=float build
 * takes them, so the FP64 forces are those of the same interacting set.
 *
 * The C++ reference cannot be compiled here (it needs CUDA's runtime, vector
 * types and Thrust device vectors in its host files), so no oracle/_ref
 * build exists; this restatement is pinned by closed forms, invariants,
 * brute force and finite differences in tests/.
 */
#ifndef FEMMD_ORACLE_H
#define FEMMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One periodic 2D system, all arrays in the user's particle order. */
typedef struct or_system {
    int32_t n;
    double Lx, Ly;
    const double *x, *y;        /* positions, n                          */
    const double *sigma;        /* per particle sigma, n                 */
    const int32_t *type;        /* per particle type, n                  */
    const int32_t *excl;        /* exclusion group, -1 = none, n         */
    int32_t ntypes;
    const double *eps, *rc, *ushift;   /* ntypes*ntypes, row-major      */
    int32_t lj_on;
    int32_t nbond;
    const int32_t *bi, *bj;     /* bond ends, nbond                      */
    const double *bk, *br0;     /* stiffness, rest length, nbond         */
    int32_t ntri;
    const int32_t *tri;         /* 3*ntri node indices                   */
    const double *rx, *ry;      /* reference (unwrapped) node positions  */
    const double *tE, *tnu;     /* per element Young modulus and Poisson */
    int32_t model;              /* 0 Neo-Hookean, 1 SVK, 2 legacy 2xSVK  */
} or_system;

/* energies[OR_NEN] layout shared by every entry point.  OR_WS_BOND and
 * OR_WS_FEA are the sums of the magnitudes of the virial terms (per bond
 * |f . r|, per element corner |f_k . (r_k - r_2)|): the scale of the
 * cancellation in W_bond and W_fea, against which an FP32 sum is judged. */
enum {
    OR_U_LJ = 0, OR_W_LJ = 1, OR_U_BOND = 2, OR_W_BOND = 3,
    OR_U_FEA = 4, OR_W_FEA = 5, OR_NPAIRS = 6, OR_NINVERTED = 7,
    OR_WS_BOND = 8, OR_WS_FEA = 9, OR_NEN = 10
};

/* def.h:112 wrap, FP32, in place. */
void or_wrap_f32(int32_t n, float *x, float *y, float Lx, float Ly);

/* nnCuda.cu:9-17 cell id, FP32 arithmetic, clamp of ix == ncx. */
void or_cell_keys_f32(int32_t n, const float *x, const float *y, float Lx,
                      float Ly, int32_t ncx, int32_t ncy, int32_t *key);

/* nn.cpp:212-213: number of cells along a side, floor(L / cmin). */
int32_t or_cell_count(double L, double cmin);

/* Stable counting sort by key (cell fill order of nn.cpp:8-20). */
void or_counting_sort(int32_t n, const int32_t *key, int32_t ncells,
                      int32_t *start, int32_t *perm);

/* Interacting sets under the FP32 predicate of interactionsCuda.cu:84-90.
 * For each i: cnt = |{j}|, s1 = sum (j+1), s2 = sum (j+1)^2.
 * mode 0 = all pairs (nn.cpp:66-92 shape), 1 = 3x3 cell walk (nn.cpp:22-62). */
int32_t or_pair_sets_f32(const or_system *s, int32_t mode, int32_t *cnt,
                         int64_t *s1, int64_t *s2);

/* FP64 forces and energies at the given positions.
 * pred_f32: take cutoff / image decisions in FP32 (positions must then be
 * FP32 values).  all_pairs: 1 = O(N^2), 0 = cell walk.  fscale (may be
 * NULL, 2n doubles) receives per particle, in [0, n), the sum of the
 * magnitudes of its pair, bond and element force terms (the scale of the
 * cancellation in its net force) and, in [n, 2n), the sum over those terms
 * of r |df/dr| (the force change per relative change of each separation:
 * what an FP32 evaluation of the same inputs cannot resolve below).
 * Returns 0 or a negative error (bad input). */
int32_t or_forces(const or_system *s, int32_t pred_f32, int32_t all_pairs,
                  double *fx, double *fy, double *energies, double *fscale);

/* FP64 KDK rollout (integrator.cpp:8-29,100-135) from (x, y, vx, vy).
 * inv_m = 0 marks a clamped node.  The cutoff predicate is FP64 unless
 * pred_f32.  State arrays are updated in place.  If en_hist is not NULL it
 * receives OR_NEN + 2 doubles per step: the energies of or_forces plus
 * KE = 1/2 sum m v^2 and max v^2 at the end of the step. */
int32_t or_run(const or_system *s, double *x, double *y, double *vx,
               double *vy, const double *mass, const double *inv_m,
               int32_t nsteps, double dt, int32_t pred_f32, double *en_hist);

/* Forward-mode derivative of an nsteps or_run rollout along the recorded
 * trajectory tx, ty ((nsteps + 1) * n, step-major) for the parameter
 * direction (dE per element or NULL, deps on every eps entry, dsig on every
 * sigma); force derivatives by FP64 central differences of step hfd.
 * Writes (dx, dy) at step nsteps. */
int32_t or_tangent(const or_system *s, const double *tx, const double *ty, int32_t nsteps,
                   double dt, const double *inv_m, const double *dE, double deps, double dsig,
                   double hfd, double *dx, double *dy);

/* ---- CPU baseline: the reference DEV=CPU path in FP32, Verlet list ---- */
typedef struct or_baseline or_baseline;
or_baseline *or_bl_create(const or_system *s, const float *vx,
                          const float *vy, const float *mass, float dt,
                          float skin, int32_t nthreads);
/* Runs nsteps of doStep(); returns wall seconds (steady clock). */
double or_bl_run(or_baseline *b, int32_t nsteps);
void or_bl_get(or_baseline *b, float *x, float *y, float *vx, float *vy,
               double *props /* KE_sum, U, W, W_fea, nn_updates */);
void or_bl_destroy(or_baseline *b);
int32_t or_bl_threads(void);

#ifdef __cplusplus
}
#endif
#endif
