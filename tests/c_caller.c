/*
 * c_caller.c -- a plain C99 caller of include/femmd.h, as a host program
 * outside Python would bind the engine (create -> add particles, an
 * elastomer, bonds -> pair table, material -> commit -> step -> energies,
 * averages, state).
 *
 * The system: 30 x 30 periodic box; solvent on a 20 x 20 lattice (spacing
 * 1.5) with the sites near the elastomer removed; one 7-node hexagonal
 * elastomer (6 CCW triangles, 1-based indices like the reference's mesh
 * files) at (15.3, 15.2); two harmonic bonds from solvent particles to rim
 * nodes (r0 = current length); LJ eps = 1, rc = 2.5 sigma, no shift;
 * Neo-Hookean E = 50, nu = 0.3; velocities from a fixed LCG.
 *
 * Usage: c_caller NSTEPS
 * Prints the counts ("n") and the inputs ("s" per solvent particle,
 * "verts", "tris", "bonds"), then after the run "energy", "avg" and one
 * "p i x y vx vy" line per particle (%.9g round-trips a float exactly,
 * %.17g a double), so a test can compare it with the same system built
 * through another binding.  Exit status 0 on success, 1 on an engine error
 * (the message on stderr).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "femmd.h"

static int check(fm_system *h, int rc, const char *what) {
    if (rc != FM_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, h ? fm_last_error(h) : "");
        return 1;
    }
    return 0;
}

static unsigned int lcg_state = 12345u;
static float lcg_uniform(void) { /* in [-0.5, 0.5) */
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return (float)(lcg_state >> 8) / 16777216.0f - 0.5f;
}

int main(int argc, char **argv) {
    const int nsteps = argc > 1 ? atoi(argv[1]) : 50;
    fm_params p;
    memset(&p, 0, sizeof p);
    p.Lx = 30.0;
    p.Ly = 30.0;
    p.dt = 0.005;
    p.device = 0;
    fm_system *h = NULL;
    if (check(NULL, fm_create(&p, &h), "fm_create")) return 1;

    /* solvent */
    static float x[400], y[400], vx[400], vy[400], m[400], sig[400];
    const float cx = 15.3f, cy = 15.2f;
    int n = 0;
    for (int i = 0; i < 20; ++i)
        for (int j = 0; j < 20; ++j) {
            const float px = 0.75f + 1.5f * (float)i, py = 0.75f + 1.5f * (float)j;
            const float dx = px - cx, dy = py - cy;
            if (dx * dx + dy * dy < 2.3f * 2.3f) continue;
            x[n] = px;
            y[n] = py;
            vx[n] = lcg_uniform();
            vy[n] = lcg_uniform();
            m[n] = 1.0f;
            sig[n] = 1.0f;
            ++n;
        }
    int32_t first = -1;
    if (check(h, fm_add_particles(h, n, x, y, vx, vy, m, sig, NULL, &first), "fm_add_particles"))
        return 1;

    /* a hexagonal elastomer: centre + 6 rim nodes, 6 triangles (1-based) */
    float verts[14];
    verts[0] = 0.0f;
    verts[1] = 0.0f;
    for (int k = 0; k < 6; ++k) {
        const double a = k * (3.14159265358979323846 / 3.0);
        verts[2 + 2 * k] = (float)(1.1 * cos(a));
        verts[3 + 2 * k] = (float)(1.1 * sin(a));
    }
    int32_t tris[18];
    for (int k = 0; k < 6; ++k) {
        tris[3 * k] = 1;
        tris[3 * k + 1] = 2 + k;
        tris[3 * k + 2] = 2 + (k + 1) % 6;
    }
    int32_t node0 = -1;
    if (check(h, fm_set_material(h, FM_NEO_HOOKEAN, 50.0f, 0.3f), "fm_set_material")) return 1;
    if (check(h, fm_add_elastomer(h, 7, verts, 6, tris, 1, cx, cy, 0, 1.0f, 1.0f, 0.0f, 1, NULL,
                                  NULL, &node0), "fm_add_elastomer"))
        return 1;

    /* two bonds from the nearest solvent particles to rim nodes 1 and 4 */
    int32_t bi[2], bj[2];
    float bk[2] = {100.0f, 100.0f}, br0[2] = {-1.0f, -1.0f};
    for (int b = 0; b < 2; ++b) {
        const int rim = node0 + (b == 0 ? 1 : 4);
        const float rx = cx + verts[2 * (rim - node0)], ry = cy + verts[2 * (rim - node0) + 1];
        int best = 0;
        float bd = 1e30f;
        for (int i = 0; i < n; ++i) {
            const float dx = x[i] - rx, dy = y[i] - ry, d = dx * dx + dy * dy;
            if (d < bd) { bd = d; best = i; }
        }
        bi[b] = best;
        bj[b] = rim;
    }
    if (check(h, fm_add_bonds(h, 2, bi, bj, bk, br0), "fm_add_bonds")) return 1;
    const float eps = 1.0f, rc = 2.5f, ush = 0.0f;
    if (check(h, fm_set_pair_table(h, 1, &eps, &rc, &ush), "fm_set_pair_table")) return 1;
    if (check(h, fm_commit(h), "fm_commit")) return 1;

    int32_t ntot = 0, nb = 0, nt = 0;
    if (check(h, fm_num_particles(h, &ntot), "fm_num_particles") ||
        check(h, fm_num_bonds(h, &nb), "fm_num_bonds") ||
        check(h, fm_num_elements(h, &nt), "fm_num_elements"))
        return 1;
    printf("n %d %d %d %d %d\n", (int)ntot, (int)first, (int)node0, (int)nb, (int)nt);
    /* the inputs, so another binding can build the same system */
    for (int i = 0; i < n; ++i) printf("s %.9g %.9g %.9g %.9g\n", x[i], y[i], vx[i], vy[i]);
    printf("verts");
    for (int k = 0; k < 14; ++k) printf(" %.9g", verts[k]);
    printf("\ntris");
    for (int k = 0; k < 18; ++k) printf(" %d", (int)tris[k]);
    printf("\nbonds %d %d %d %d\n", (int)bi[0], (int)bj[0], (int)bi[1], (int)bj[1]);

    if (check(h, fm_step(h, nsteps), "fm_step")) return 1;
    fm_energies e;
    if (check(h, fm_get_energies(h, &e), "fm_get_energies")) return 1;
    printf("energy %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %lld\n", e.ke, e.u_lj,
           e.w_lj, e.u_fea, e.w_fea, e.u_bond, e.w_bond, e.pressure, e.temperature,
           (long long)e.steps);
    fm_averages a;
    if (check(h, fm_get_averages(h, &a), "fm_get_averages")) return 1;
    printf("avg %lld", (long long)a.count);
    for (int k = 0; k < 8; ++k) printf(" %.17g %.17g", a.mean[k], a.stdv[k]);
    printf("\n");
    float *sx = (float *)malloc(sizeof(float) * 4 * (size_t)ntot);
    if (!sx) return 1;
    float *sy = sx + ntot, *svx = sy + ntot, *svy = svx + ntot;
    if (check(h, fm_get_state(h, sx, sy, svx, svy), "fm_get_state")) return 1;
    for (int i = 0; i < ntot; ++i)
        printf("p %d %.9g %.9g %.9g %.9g\n", i, sx[i], sy[i], svx[i], svy[i]);
    free(sx);
    fm_destroy(h);
    return 0;
}
