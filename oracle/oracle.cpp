// oracle.cpp -- CPU restatement of femmd2D's per-step force-and-integrate
// path.  TEST INFRASTRUCTURE ONLY (what the engine is tested against, and
// the CPU baseline); the product library never loads it.  Every function
// cites the reference lines it restates (paths relative to the reference
// root).
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off, as the reference's
// own g++ -O2 build of femmd2D/src/Makefile:9; the -fopenmp variant enables
// the pair-loop pragma of femmd2D/src/interactions.cpp:354).

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// femmd2D/src/def.h:91-93 (minImage).  Returns the multiple of L removed so
// that an FP64 evaluation can apply the very image an FP32 decision chose.
// The reference compares against the double 0.5*box.x; 0.5*L is exact in
// both precisions, so the float comparison here decides identically.
template <class R>
inline int min_image(R &d, R L) {
    if (d >= R(0.5) * L) { d -= L; return 1; }
    if (d < -R(0.5) * L) { d += L; return -1; }
    return 0;
}

// femmd2D/src/def.h:112 (pbcCalc): r = r - floor(r / L) * L, in R.
template <class R>
inline R wrap1(R r, R L) {
    R q = r / L;
    R f = std::floor(q);
    R p = f * L;
    return r - p;
}

// femmd2D/src/nnCuda.cu:9-17 (cellId): ix = (int)(x * ncx / Lx) in FP32,
// with the ix == ncx clamp; id = ix + iy * ncx.
inline int32_t cell_key_f32(float x, float y, float Lx, float Ly, int32_t ncx,
                            int32_t ncy) {
    float tx = x * (float)ncx;
    tx = tx / Lx;
    int32_t ix = (int32_t)tx;
    if (ix == ncx) ix -= 1;
    float ty = y * (float)ncy;
    ty = ty / Ly;
    int32_t iy = (int32_t)ty;
    if (iy == ncy) iy -= 1;
    return ix + iy * ncx;
}

// femmd2D/src/nnCuda.cu:19-37: decode, then unconditional periodic wrap of
// a neighbour cell coordinate.
inline int32_t wrap_cell(int32_t c, int32_t nc) {
    if (c < 0) return nc - 1;
    if (c == nc) return 0;
    return c;
}

// Lame constants, femmd2D/src/interactions.cpp:43-50.
inline void lame(double E, double nu, double &lam, double &mu) {
    lam = E * nu / ((1 - 2 * nu) * (1 + nu));
    mu = E / (2 * (1 + nu));
}

// ---------------------------------------------------------------------------
// Pair enumeration shared by the set and force oracles.
// ---------------------------------------------------------------------------
struct PairCtx {
    const or_system *s;
    std::vector<float> xf, yf, sf, rcf;
    float Lxf, Lyf;
    double rc_eff_max;  // max over pairs of rc_ij * sigma_ij, FP64
};

void make_ctx(const or_system *s, PairCtx &c) {
    c.s = s;
    int n = s->n;
    c.xf.resize(n); c.yf.resize(n); c.sf.resize(n);
    for (int i = 0; i < n; ++i) {
        c.xf[i] = (float)s->x[i];
        c.yf[i] = (float)s->y[i];
        c.sf[i] = (float)s->sigma[i];
    }
    int nt = s->ntypes;
    c.rcf.resize((size_t)nt * nt);
    double rcmax = 0;
    for (int k = 0; k < nt * nt; ++k) {
        c.rcf[k] = (float)s->rc[k];
        rcmax = std::max(rcmax, s->rc[k]);
    }
    double smax = 0;
    for (int i = 0; i < n; ++i) smax = std::max(smax, s->sigma[i]);
    c.Lxf = (float)s->Lx;
    c.Lyf = (float)s->Ly;
    c.rc_eff_max = rcmax * smax;
}

// femmd2D/src/nn.cpp:44-46 exclusion rule: interact iff i != j and
// (gi == -1 || gj == -1 || gi != gj).
inline bool excluded(const or_system *s, int i, int j) {
    if (i == j) return true;
    int gi = s->excl[i], gj = s->excl[j];
    return !(gi == -1 || gj == -1 || gi != gj);
}

// FP32 predicate of femmd2D/src/interactionsCuda.cu:79-90: sigma_ij =
// (si + sj) / 2, rCut2 = rc*rc*sij*sij, interacting iff dr2 < rCut2.
inline bool pred32(const PairCtx &c, int i, int j, int &sx, int &sy) {
    const or_system *s = c.s;
    float dx = c.xf[i] - c.xf[j];
    sx = min_image(dx, c.Lxf);
    float dy = c.yf[i] - c.yf[j];
    sy = min_image(dy, c.Lyf);
    float dr2 = dx * dx + dy * dy;
    float sij = (c.sf[i] + c.sf[j]) / 2;
    float rc = c.rcf[(size_t)s->type[i] * s->ntypes + s->type[j]];
    float rc2 = rc * rc * sij * sij;
    return dr2 < rc2;
}

// FP64 predicate (same expression in double).
inline bool pred64(const or_system *s, int i, int j, int &sx, int &sy) {
    double dx = s->x[i] - s->x[j];
    sx = min_image(dx, s->Lx);
    double dy = s->y[i] - s->y[j];
    sy = min_image(dy, s->Ly);
    double dr2 = dx * dx + dy * dy;
    double sij = (s->sigma[i] + s->sigma[j]) / 2;
    double rc = s->rc[(size_t)s->type[i] * s->ntypes + s->type[j]];
    double rc2 = rc * rc * sij * sij;
    return dr2 < rc2;
}

// Calls fn(i, j) for every ordered candidate pair (i != j, not excluded).
// all_pairs mirrors femmd2D/src/nn.cpp:66-92; otherwise the 3x3 cell walk of
// femmd2D/src/nn.cpp:22-62 with cx outer, cy inner and in-cell order equal
// to ascending particle index (femmd2D/src/nn.cpp:8-20).
template <class Fn>
void for_candidates(const PairCtx &c, int all_pairs, Fn fn) {
    const or_system *s = c.s;
    int n = s->n;
    // cell side must cover the largest cutoff; a tiny margin absorbs the
    // FP32 rounding of FP64 positions used for binning.
    double cmin = c.rc_eff_max * (1.0 + 1e-6);
    int32_t ncx = (int32_t)std::floor(s->Lx / cmin);
    int32_t ncy = (int32_t)std::floor(s->Ly / cmin);
    if (all_pairs || ncx < 3 || ncy < 3) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (!excluded(s, i, j)) fn(i, j);
        return;
    }
    int32_t ncells = ncx * ncy;
    std::vector<int32_t> key(n), start(ncells + 1), perm(n);
    std::vector<float> xw(c.xf), yw(c.yf);
    for (int i = 0; i < n; ++i) {
        // FP64 positions may sit a rounding away from [0, L): wrap the
        // binning copy (the reference stores wrapped positions).
        xw[i] = wrap1(xw[i], c.Lxf);
        yw[i] = wrap1(yw[i], c.Lyf);
        key[i] = cell_key_f32(xw[i], yw[i], c.Lxf, c.Lyf, ncx, ncy);
        if (key[i] < 0) key[i] = 0;
        if (key[i] >= ncells) key[i] = ncells - 1;
    }
    or_counting_sort(n, key.data(), ncells, start.data(), perm.data());
    for (int i = 0; i < n; ++i) {
        int32_t cx0 = key[i] % ncx, cy0 = key[i] / ncx;
        for (int dx = -1; dx <= 1; ++dx) {
            for (int dy = -1; dy <= 1; ++dy) {
                int32_t cx = wrap_cell(cx0 + dx, ncx);
                int32_t cy = wrap_cell(cy0 + dy, ncy);
                int32_t cc = cx + cy * ncx;
                for (int32_t k = start[cc]; k < start[cc + 1]; ++k) {
                    int j = perm[k];
                    if (!excluded(s, i, j)) fn(i, j);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 force kernels.
// ---------------------------------------------------------------------------

// femmd2D/src/interactionsCuda.cu:43-55 (ljforce) and :57-100 (pairForce),
// accumulated as the full-list loop femmd2D/src/interactions.cpp:347-389:
// only f_i is written; U and W take 1/2 per ordered pair.
void lj_pair64(const or_system *s, int i, int j, int sx, int sy, double *fx,
               double *fy, double *en, double *fs) {
    int ti = s->type[i], tj = s->type[j];
    size_t k = (size_t)ti * s->ntypes + tj;
    double e = s->eps[k];
    double sij = (s->sigma[i] + s->sigma[j]) / 2;
    double dx = (s->x[i] - s->x[j]) - sx * s->Lx;
    double dy = (s->y[i] - s->y[j]) - sy * s->Ly;
    double dr2 = dx * dx + dy * dy;
    double c = 48.0 * e / (sij * sij);
    double dr2i = sij * sij / dr2;
    double dr6i = dr2i * dr2i * dr2i;
    double fc = c * dr6i * (dr6i - 0.5) * dr2i;
    fx[i] += fc * dx;
    fy[i] += fc * dy;
    if (fs) {
        const double r = std::sqrt(dr2);
        fs[i] += std::fabs(fc) * r;
        // r |dF/dr| of F(r) = 48 eps (s6^2 - s6/2) / r, s6 = (sigma/r)^6
        fs[s->n + i] += e * std::fabs(624.0 * dr6i * dr6i - 168.0 * dr6i) / r;
    }
    double u = 4.0 * e * dr6i * (dr6i - 1.0) + s->ushift[k];
    en[OR_U_LJ] += 0.5 * u;
    en[OR_W_LJ] += 0.5 * fc * dr2;
    en[OR_NPAIRS] += 1.0;
}

// femmd2D/src/interactions.cpp:143-173 with per-bond k (the CPU path uses
// one global kBond; Appendix choice: per-bond k and r0).  The reference CPU
// path leaves the bond virial out of sys->virial; it is reported apart.
void bonds64(const or_system *s, const double *x, const double *y, double *fx,
             double *fy, double *en, double *fs) {
    for (int b = 0; b < s->nbond; ++b) {
        int i = s->bi[b], j = s->bj[b];
        double dx = x[i] - x[j];
        double dy = y[i] - y[j];
        min_image(dx, s->Lx);
        min_image(dy, s->Ly);
        double r = std::sqrt(dx * dx + dy * dy);
        double dr = r - s->br0[b];
        double kdr = -s->bk[b] * dr / r;
        fx[i] += kdr * dx; fy[i] += kdr * dy;
        fx[j] -= kdr * dx; fy[j] -= kdr * dy;
        if (fs) {
            fs[i] += std::fabs(kdr) * r;
            fs[j] += std::fabs(kdr) * r;
            fs[s->n + i] += s->bk[b] * r;  // r dF/dr of F = k (r - r0)
            fs[s->n + j] += s->bk[b] * r;
        }
        en[OR_U_BOND] += 0.5 * s->bk[b] * dr * dr;
        en[OR_W_BOND] += kdr * (dx * dx + dy * dy);
        en[OR_WS_BOND] += std::fabs(kdr * (dx * dx + dy * dy));
    }
}

// femmd2D/src/interactions.cpp:532-689 (calcFeaForcesCPU2).  Nodes 1 and 2
// are unwrapped against node 0 by one half-box test (:554-568); Dm from the
// reference positions, Bm = Dm^-1, vol0 = det(Dm)/2 (signed, :594-602);
// F = Ds Bm; Neo-Hookean P = mu (F - F^-T) + lambda lnJ F^-T and
// psi = mu/2 (I1 - 2) - mu lnJ + lambda/2 (lnJ)^2 (:640-657) or SVK
// (:631-639); H = -vol0 P Bm^T, f0 = H[:,0], f1 = H[:,1], f2 = -f0 - f1.
// The legacy GPU model (interactionsCuda.cu:397-433) is twice the SVK energy.
// W_fea = sum_k f_k . r_k over unwrapped corners (the reference sums
// f . unfoldedPos; equal because sum_k f_k = 0 per element).
void cst64(const or_system *s, const double *x, const double *y, double *fx,
           double *fy, double *en, double *fs) {
    double Lx = s->Lx, Ly = s->Ly;
    for (int e = 0; e < s->ntri; ++e) {
        int n0 = s->tri[3 * e], n1 = s->tri[3 * e + 1], n2 = s->tri[3 * e + 2];
        double r0x = x[n0], r0y = y[n0];
        double r1x = x[n1], r1y = y[n1];
        double r2x = x[n2], r2y = y[n2];
        double dx01 = r0x - r1x;
        if (dx01 > Lx / 2.0) r1x += Lx;
        if (dx01 < -Lx / 2.0) r1x -= Lx;
        double dx02 = r0x - r2x;
        if (dx02 > Lx / 2.0) r2x += Lx;
        if (dx02 < -Lx / 2.0) r2x -= Lx;
        double dy01 = r0y - r1y;
        if (dy01 > Ly / 2.0) r1y += Ly;
        if (dy01 < -Ly / 2.0) r1y -= Ly;
        double dy02 = r0y - r2y;
        if (dy02 > Ly / 2.0) r2y += Ly;
        if (dy02 < -Ly / 2.0) r2y -= Ly;

        double Dm00 = s->rx[n0] - s->rx[n2], Dm10 = s->ry[n0] - s->ry[n2];
        double Dm01 = s->rx[n1] - s->rx[n2], Dm11 = s->ry[n1] - s->ry[n2];
        double det = Dm00 * Dm11 - Dm01 * Dm10;
        double vol0 = det / 2;
        double B00 = Dm11 / det, B10 = -Dm10 / det;
        double B01 = -Dm01 / det, B11 = Dm00 / det;

        double Ds00 = r0x - r2x, Ds10 = r0y - r2y;
        double Ds01 = r1x - r2x, Ds11 = r1y - r2y;
        double F00 = Ds00 * B00 + Ds01 * B10;
        double F10 = Ds10 * B00 + Ds11 * B10;
        double F01 = Ds00 * B01 + Ds01 * B11;
        double F11 = Ds10 * B01 + Ds11 * B11;

        double lam, mu;
        lame(s->tE[e], s->tnu[e], lam, mu);
        double P00, P01, P10, P11, psi;
        if (s->model == 0) {
            double J = F00 * F11 - F01 * F10;
            if (!(J > 0)) en[OR_NINVERTED] += 1.0;
            double lJ = std::log(J);
            double i00 = F11 / J, i10 = -F10 / J, i01 = -F01 / J, i11 = F00 / J;
            // F^-T[a][b] = invF[b][a]
            P00 = mu * (F00 - i00) + lam * lJ * i00;
            P10 = mu * (F10 - i01) + lam * lJ * i01;
            P01 = mu * (F01 - i10) + lam * lJ * i10;
            P11 = mu * (F11 - i11) + lam * lJ * i11;
            double I1 = F00 * F00 + F10 * F10 + F01 * F01 + F11 * F11;
            psi = 0.5 * mu * (I1 - 2) - mu * lJ + 0.5 * lam * lJ * lJ;
        } else {
            double sc = (s->model == 2) ? 2.0 : 1.0;
            lam *= sc; mu *= sc;
            double E00 = 0.5 * (F00 * F00 + F10 * F10 - 1);
            double E10 = 0.5 * (F00 * F01 + F11 * F10);
            double E01 = 0.5 * (F00 * F01 + F10 * F11);
            double E11 = 0.5 * (F01 * F01 + F11 * F11 - 1);
            double tr = E00 + E11, ltr = lam * tr;
            P00 = ltr * F00 + 2 * mu * (F00 * E00 + F01 * E10);
            P10 = ltr * F10 + 2 * mu * (F10 * E00 + F11 * E10);
            P01 = ltr * F01 + 2 * mu * (F00 * E01 + F01 * E11);
            P11 = ltr * F11 + 2 * mu * (F10 * E01 + F11 * E11);
            psi = mu * (E00 * E00 + E10 * E10 + E01 * E01 + E11 * E11) +
                  0.5 * lam * tr * tr;
            double J = F00 * F11 - F01 * F10;
            if (!(J > 0)) en[OR_NINVERTED] += 1.0;
        }
        double H00 = -vol0 * (P00 * B00 + P01 * B01);
        double H10 = -vol0 * (P10 * B00 + P11 * B01);
        double H01 = -vol0 * (P00 * B10 + P01 * B11);
        double H11 = -vol0 * (P10 * B10 + P11 * B11);
        double f0x = H00, f0y = H10, f1x = H01, f1y = H11;
        double f2x = -f0x - f1x, f2y = -f0y - f1y;
        fx[n0] += f0x; fy[n0] += f0y;
        fx[n1] += f1x; fy[n1] += f1y;
        fx[n2] += f2x; fy[n2] += f2y;
        if (fs) {
            fs[n0] += std::hypot(f0x, f0y);
            fs[n1] += std::hypot(f1x, f1y);
            fs[n2] += std::hypot(f2x, f2y);
            // element stiffness x length: (lambda + 2 mu) |A0| |Bm|_F
            const double bn = std::sqrt(B00 * B00 + B01 * B01 + B10 * B10 + B11 * B11);
            const double kl = (std::fabs(lam) + 2 * std::fabs(mu)) * std::fabs(vol0) * bn;
            fs[s->n + n0] += kl;
            fs[s->n + n1] += kl;
            fs[s->n + n2] += kl;
        }
        en[OR_U_FEA] += psi * vol0;
        en[OR_W_FEA] += f0x * r0x + f0y * r0y + f1x * r1x + f1y * r1y +
                        f2x * r2x + f2y * r2y;
        en[OR_WS_FEA] += std::fabs(f0x * (r0x - r2x)) + std::fabs(f0y * (r0y - r2y)) +
                         std::fabs(f1x * (r1x - r2x)) + std::fabs(f1y * (r1y - r2y));
    }
}

int32_t forces_impl(const or_system *s, int pred_f32, int all_pairs,
                    double *fx, double *fy, double *en, double *fs = nullptr) {
    if (!s || s->n < 0 || !(s->Lx > 0) || !(s->Ly > 0)) return -1;
    std::fill(fx, fx + s->n, 0.0);
    std::fill(fy, fy + s->n, 0.0);
    std::fill(en, en + OR_NEN, 0.0);
    if (fs) std::fill(fs, fs + 2 * (size_t)s->n, 0.0);
    // force order of femmd2D/src/interactions.cpp:96-102: bonds, FEA, pairs.
    if (s->nbond > 0) bonds64(s, s->x, s->y, fx, fy, en, fs);
    if (s->ntri > 0) cst64(s, s->x, s->y, fx, fy, en, fs);
    if (s->lj_on && s->n > 0) {
        PairCtx c;
        make_ctx(s, c);
        for_candidates(c, all_pairs, [&](int i, int j) {
            int sx, sy;
            bool hit = pred_f32 ? pred32(c, i, j, sx, sy) : pred64(s, i, j, sx, sy);
            if (hit) lj_pair64(s, i, j, sx, sy, fx, fy, en, fs);
        });
    }
    return 0;
}

}  // namespace

// ===========================================================================
extern "C" {

void or_wrap_f32(int32_t n, float *x, float *y, float Lx, float Ly) {
    for (int i = 0; i < n; ++i) {
        x[i] = wrap1(x[i], Lx);
        y[i] = wrap1(y[i], Ly);
    }
}

void or_cell_keys_f32(int32_t n, const float *x, const float *y, float Lx,
                      float Ly, int32_t ncx, int32_t ncy, int32_t *key) {
    for (int i = 0; i < n; ++i) key[i] = cell_key_f32(x[i], y[i], Lx, Ly, ncx, ncy);
}

int32_t or_cell_count(double L, double cmin) {
    return (int32_t)std::floor(L / cmin);
}

// Stable counting sort: the order the reference's cell fill produces
// (femmd2D/src/nn.cpp:12-19 visits particles in ascending index).
void or_counting_sort(int32_t n, const int32_t *key, int32_t ncells,
                      int32_t *start, int32_t *perm) {
    std::vector<int32_t> cnt(ncells, 0);
    for (int i = 0; i < n; ++i) cnt[key[i]]++;
    start[0] = 0;
    for (int c = 0; c < ncells; ++c) start[c + 1] = start[c] + cnt[c];
    std::vector<int32_t> fill(start, start + ncells);
    for (int i = 0; i < n; ++i) perm[fill[key[i]]++] = i;
}

int32_t or_pair_sets_f32(const or_system *s, int32_t mode, int32_t *cnt,
                         int64_t *s1, int64_t *s2) {
    if (!s || s->n < 0) return -1;
    PairCtx c;
    make_ctx(s, c);
    for (int i = 0; i < s->n; ++i) { cnt[i] = 0; s1[i] = 0; s2[i] = 0; }
    for_candidates(c, mode == 0, [&](int i, int j) {
        int sx, sy;
        if (pred32(c, i, j, sx, sy)) {
            int64_t v = (int64_t)j + 1;
            cnt[i] += 1;
            s1[i] += v;
            s2[i] += v * v;
        }
    });
    return 0;
}

int32_t or_forces(const or_system *s, int32_t pred_f32, int32_t all_pairs,
                  double *fx, double *fy, double *energies, double *fscale) {
    return forces_impl(s, pred_f32, all_pairs, fx, fy, energies, fscale);
}

// femmd2D/src/integrator.cpp:8-29 (doStep), :100-135 (leapFrogStep1/2,
// applyPBC, updateAcc): v += dt/2 a; x += dt v; wrap; forces; a = f/m;
// v += dt/2 a.  Clamped nodes carry inv_m = 0.
int32_t or_run(const or_system *s0, double *x, double *y, double *vx,
               double *vy, const double *mass, const double *inv_m,
               int32_t nsteps, double dt, int32_t pred_f32, double *en_hist) {
    if (!s0) return -1;
    or_system s = *s0;
    s.x = x;
    s.y = y;
    int n = s.n;
    std::vector<double> fx(n), fy(n);
    double en[OR_NEN];
    int rc = forces_impl(&s, pred_f32, 0, fx.data(), fy.data(), en);
    if (rc) return rc;
    double h = 0.5 * dt;
    for (int st = 0; st < nsteps; ++st) {
        for (int i = 0; i < n; ++i) {
            vx[i] += h * fx[i] * inv_m[i];
            vy[i] += h * fy[i] * inv_m[i];
            x[i] += dt * vx[i];
            y[i] += dt * vy[i];
            x[i] = wrap1(x[i], s.Lx);
            y[i] = wrap1(y[i], s.Ly);
        }
        rc = forces_impl(&s, pred_f32, 0, fx.data(), fy.data(), en);
        if (rc) return rc;
        double ke = 0, vvmax = 0;
        for (int i = 0; i < n; ++i) {
            vx[i] += h * fx[i] * inv_m[i];
            vy[i] += h * fy[i] * inv_m[i];
            double vv = vx[i] * vx[i] + vy[i] * vy[i];
            ke += mass[i] * vv;
            vvmax = std::max(vvmax, vv);
        }
        if (en_hist) {
            double *hs = en_hist + (OR_NEN + 2) * (size_t)st;
            std::memcpy(hs, en, OR_NEN * sizeof(double));
            hs[OR_NEN] = 0.5 * ke;
            hs[OR_NEN + 1] = vvmax;
        }
    }
    return 0;
}

// Forward-mode derivative of an nsteps KDK rollout (the map of or_run)
// along a recorded trajectory: positions (tx, ty)[k * n + i] at steps
// k = 0..nsteps.  The parameters move along (dE per element, deps added to
// every eps entry, dsig added to every sigma); the initial state does not
// depend on them.  Every force derivative is a central difference in FP64,
//   G_k = [F(x_k + h dx_k; p + h d) - F(x_k - h dx_k; p - h d)] / 2h,
// which is J dx_k + (dF/dp) d to O(h^2); the tangent follows the step:
//   dv += dt/2 G_k inv_m; dx += dt dv; dv += dt/2 G_{k+1} inv_m.
// Both evaluations of G_k use the interacting pairs (and their periodic
// images) of the unperturbed x_k: the derivative of the smooth branch the
// step is on.  (A pair within h |dr| of the cutoff would otherwise enter one
// evaluation and not the other; where the perturbed positions move far, as
// a sigma direction drives them over hundreds of steps, such pairs are
// frequent enough to bias a plain difference by parts in 1e3.)
// Returns (dx, dy) at step nsteps.  Reentrant (callers run directions in
// threads).
int32_t or_tangent(const or_system *s0, const double *tx, const double *ty, int32_t nsteps,
                   double dt, const double *inv_m, const double *dE, double deps, double dsig,
                   double hfd, double *dx, double *dy) {
    if (!s0 || nsteps < 0 || !(hfd > 0)) return -1;
    const int n = s0->n;
    const int nt2 = s0->ntypes * s0->ntypes;
    std::vector<double> sig[2], eps[2], tE[2];
    for (int q = 0; q < 2; ++q) {
        const double sg = q ? -hfd : hfd;
        sig[q].resize(n);
        for (int i = 0; i < n; ++i) sig[q][i] = s0->sigma[i] + sg * dsig;
        eps[q].resize(nt2);
        for (int k = 0; k < nt2; ++k) eps[q][k] = s0->eps[k] + sg * deps;
        tE[q].resize(s0->ntri);
        for (int e = 0; e < s0->ntri; ++e) tE[q][e] = s0->tE[e] + (dE ? sg * dE[e] : 0.0);
    }
    struct PairRec { int32_t i, j, sx, sy; };
    std::vector<PairRec> pairs;
    std::vector<double> px(n), py(n), fpx(n), fpy(n), fmx(n), fmy(n), gx(n), gy(n);
    std::vector<double> vx(n, 0.0), vy(n, 0.0);
    double en[OR_NEN];
    auto G = [&](int k) -> int {
        const double *xk = tx + (size_t)k * n, *yk = ty + (size_t)k * n;
        pairs.clear();
        if (s0->lj_on && n > 0) {
            or_system sb = *s0;
            sb.x = xk;
            sb.y = yk;
            PairCtx c;
            make_ctx(&sb, c);
            for_candidates(c, 0, [&](int i, int j) {
                int sx, sy;
                if (pred64(&sb, i, j, sx, sy)) pairs.push_back({i, j, sx, sy});
            });
        }
        for (int q = 0; q < 2; ++q) {
            const double sg = q ? -hfd : hfd;
            // not wrapped: each pair keeps the image shift of x_k
            for (int i = 0; i < n; ++i) {
                px[i] = xk[i] + sg * dx[i];
                py[i] = yk[i] + sg * dy[i];
            }
            or_system s = *s0;
            s.x = px.data();
            s.y = py.data();
            s.sigma = sig[q].data();
            s.eps = eps[q].data();
            if (s.ntri > 0) s.tE = tE[q].data();
            double *fx = q ? fmx.data() : fpx.data(), *fy = q ? fmy.data() : fpy.data();
            std::fill(fx, fx + n, 0.0);
            std::fill(fy, fy + n, 0.0);
            if (s.nbond > 0) bonds64(&s, s.x, s.y, fx, fy, en, nullptr);
            if (s.ntri > 0) cst64(&s, s.x, s.y, fx, fy, en, nullptr);
            for (const PairRec &p : pairs) lj_pair64(&s, p.i, p.j, p.sx, p.sy, fx, fy, en, nullptr);
        }
        for (int i = 0; i < n; ++i) {
            gx[i] = (fpx[i] - fmx[i]) / (2 * hfd);
            gy[i] = (fpy[i] - fmy[i]) / (2 * hfd);
        }
        return 0;
    };
    std::fill(dx, dx + n, 0.0);
    std::fill(dy, dy + n, 0.0);
    const double h = 0.5 * dt;
    int rc = G(0);
    if (rc) return rc;
    for (int k = 0; k < nsteps; ++k) {
        for (int i = 0; i < n; ++i) {
            vx[i] += h * gx[i] * inv_m[i];
            vy[i] += h * gy[i] * inv_m[i];
            dx[i] += dt * vx[i];
            dy[i] += dt * vy[i];
        }
        rc = G(k + 1);
        if (rc) return rc;
        for (int i = 0; i < n; ++i) {
            vx[i] += h * gx[i] * inv_m[i];
            vy[i] += h * gy[i] * inv_m[i];
        }
    }
    return 0;
}

}  // extern "C"

// ===========================================================================
// CPU baseline: the reference DEV=CPU path, FP32 (real = float), AoS state,
// Verlet list with skin (femmd2D/src/nn.h:63, nn.cpp:114-134,200-225),
// full-list pair loop (interactions.cpp:347-389, OpenMP pragma at :354),
// bonds (:143-173), calcClusterProps + Neo-Hookean CST (:393-399, :532-689,
// feaElements.cpp:86-204), KDK (integrator.cpp:8-183), evalProps
// (md3dsystem.cpp:151-179).
// ===========================================================================
namespace {
struct V2 { float x, y; };

struct Cluster {
    std::vector<int> nodes;
    float mass = 0;
    V2 cm{0, 0}, cmv{0, 0}, centroid{0, 0};
};
}  // namespace

struct or_baseline {
    int n = 0, nthreads = 1;
    float Lx = 0, Ly = 0, dt = 0, skin = 0.4f;
    std::vector<V2> pos, vel, acc, force, refPos, unfolded;
    std::vector<float> mass, sigma, vv;
    std::vector<int> type, excl, cluster_of;
    int ntypes = 1;
    std::vector<float> eps, rc, ush;
    bool lj_on = true;
    std::vector<int> bondList;
    std::vector<float> bondLen, bondK;
    std::vector<int> tri;
    std::vector<float> tlam, tmu;
    std::vector<Cluster> clusters;
    // neighbour list
    int ncx = 0, ncy = 0, maxPPC = 0, maxNbr = 0;
    float nnRadiusSq = 0, dispHi = 0;
    std::vector<int> cellCount, cellContent, nbrList, nbrCount;
    int updates = 0;
    // energies (FP32 accumulators, femmd2D/src/md3dsystem.h:45)
    float potEnergy = 0, virial = 0, feaVirial = 0, clusterVirial = 0;
    float kinEneSum = 0, vvMax = 0, clusterKin = 0;

    void buildNN() {
        std::fill(cellCount.begin(), cellCount.end(), 0);
        // femmd2D/src/nn.cpp:8-20 updateCellsCPU
        for (int p = 0; p < n; ++p) {
            int c = cell_key_f32(pos[p].x, pos[p].y, Lx, Ly, ncx, ncy);
            int slot = cellCount[c]++;
            if (slot < maxPPC) cellContent[(size_t)c * maxPPC + slot] = p;
        }
        // femmd2D/src/nn.cpp:22-62 buildNNCPUSerial
        for (int pi = 0; pi < n; ++pi) {
            int cid = cell_key_f32(pos[pi].x, pos[pi].y, Lx, Ly, ncx, ncy);
            int cx0 = cid % ncx, cy0 = cid / ncx, k = 0;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    int cc = wrap_cell(cx0 + dx, ncx) + wrap_cell(cy0 + dy, ncy) * ncx;
                    int m = std::min(cellCount[cc], maxPPC);
                    for (int q = 0; q < m; ++q) {
                        int pj = cellContent[(size_t)cc * maxPPC + q];
                        int gi = excl[pi], gj = excl[pj];
                        if (pi != pj && (gi == -1 || gj == -1 || gi != gj)) {
                            float ddx = pos[pi].x - pos[pj].x;
                            float ddy = pos[pi].y - pos[pj].y;
                            min_image(ddx, Lx);
                            min_image(ddy, Ly);
                            float d2 = ddx * ddx + ddy * ddy;
                            if (d2 <= nnRadiusSq && k < maxNbr)
                                nbrList[(size_t)pi * maxNbr + k++] = pj;
                        }
                    }
                }
            nbrCount[pi] = k;
        }
    }

    // femmd2D/src/nn.cpp:114-134
    void updateNN(bool restart) {
        if (dispHi > 0.5f * skin) { updates++; dispHi = 0; buildNN(); }
        if (restart) { updates = 1; dispHi = 0; buildNN(); }
    }

    // femmd2D/src/feaElements.cpp:86-151 (unfoldPos) and :162-197
    // (updateClustersCentroid), run every step by calcFeaForces.
    void clusterProps() {
        for (auto &cl : clusters) {
            float xmin = 1e5f, xmax = -1e5f, ymin = 1e5f, ymax = -1e5f;
            for (int p : cl.nodes) {
                unfolded[p] = pos[p];
                xmin = std::min(xmin, pos[p].x); xmax = std::max(xmax, pos[p].x);
                ymin = std::min(ymin, pos[p].y); ymax = std::max(ymax, pos[p].y);
            }
            int es = (int)cl.nodes.size();
            if (xmax - xmin > Lx / 2) {
                float xc = 0;
                for (int p : cl.nodes) {
                    if (unfolded[p].x > Lx / 2) unfolded[p].x -= Lx;
                    xc += unfolded[p].x;
                }
                if (xc / es < 0)
                    for (int p : cl.nodes) unfolded[p].x += Lx;
            }
            if (ymax - ymin > Ly / 2) {
                float yc = 0;
                for (int p : cl.nodes) {
                    if (unfolded[p].y > Ly / 2) unfolded[p].y -= Ly;
                    yc += unfolded[p].y;
                }
                if (yc / es < 0)
                    for (int p : cl.nodes) unfolded[p].y += Ly;
            }
            float xc = 0, yc = 0, vxc = 0, vyc = 0;
            for (int p : cl.nodes) {
                xc += mass[p] * unfolded[p].x; yc += mass[p] * unfolded[p].y;
                vxc += mass[p] * vel[p].x; vyc += mass[p] * vel[p].y;
            }
            cl.cm = {xc / cl.mass, yc / cl.mass};
            cl.cmv = {vxc / cl.mass, vyc / cl.mass};
            const float twopi = 6.28318530717958647692f;
            float cxa = 0, sxa = 0, cya = 0, sya = 0;
            for (int p : cl.nodes) {
                float tx = twopi * pos[p].x / Lx;
                cxa += std::cos(tx); sxa += std::sin(tx);
                float ty = twopi * pos[p].y / Ly;
                cya += std::cos(ty); sya += std::sin(ty);
            }
            cxa /= es; sxa /= es; cya /= es; sya /= es;
            cl.centroid.x = Lx * (std::atan2(-sxa, -cxa) + 3.14159265f) / twopi;
            cl.centroid.y = Ly * (std::atan2(-sya, -cya) + 3.14159265f) / twopi;
        }
    }

    void bondForces() {
        for (size_t l = 0; l < bondLen.size(); ++l) {
            int i = bondList[2 * l], j = bondList[2 * l + 1];
            float dx = pos[i].x - pos[j].x, dy = pos[i].y - pos[j].y;
            min_image(dx, Lx); min_image(dy, Ly);
            float r = std::sqrt(dx * dx + dy * dy);
            float d = r - bondLen[l];
            float kdr = -bondK[l] * d / r;
            force[i].x += kdr * dx; force[i].y += kdr * dy;
            force[j].x -= kdr * dx; force[j].y -= kdr * dy;
            potEnergy += 0.5f * bondK[l] * d * d;
        }
    }

    void feaForces() {
        clusterProps();
        int ntri = (int)tlam.size();
        for (int e = 0; e < ntri; ++e) {
            int p[3] = {tri[3 * e], tri[3 * e + 1], tri[3 * e + 2]};
            V2 r[3] = {pos[p[0]], pos[p[1]], pos[p[2]]};
            V2 r0[3] = {refPos[p[0]], refPos[p[1]], refPos[p[2]]};
            float dx01 = r[0].x - r[1].x;
            if (dx01 > Lx / 2.0f) r[1].x += Lx;
            if (dx01 < -Lx / 2.0f) r[1].x -= Lx;
            float dx02 = r[0].x - r[2].x;
            if (dx02 > Lx / 2.0f) r[2].x += Lx;
            if (dx02 < -Lx / 2.0f) r[2].x -= Lx;
            float dy01 = r[0].y - r[1].y;
            if (dy01 > Ly / 2.0f) r[1].y += Ly;
            if (dy01 < -Ly / 2.0f) r[1].y -= Ly;
            float dy02 = r[0].y - r[2].y;
            if (dy02 > Ly / 2.0f) r[2].y += Ly;
            if (dy02 < -Ly / 2.0f) r[2].y -= Ly;
            float D00 = r0[0].x - r0[2].x, D10 = r0[0].y - r0[2].y;
            float D01 = r0[1].x - r0[2].x, D11 = r0[1].y - r0[2].y;
            float det = D00 * D11 - D01 * D10, vol0 = det / 2;
            float B00 = D11 / det, B10 = -D10 / det, B01 = -D01 / det, B11 = D00 / det;
            float S00 = r[0].x - r[2].x, S10 = r[0].y - r[2].y;
            float S01 = r[1].x - r[2].x, S11 = r[1].y - r[2].y;
            float F00 = S00 * B00 + S01 * B10, F10 = S10 * B00 + S11 * B10;
            float F01 = S00 * B01 + S01 * B11, F11 = S10 * B01 + S11 * B11;
            float lam = tlam[e], mu = tmu[e];
            float J = F00 * F11 - F01 * F10, lJ = std::log(J);
            float i00 = F11 / J, i10 = -F10 / J, i01 = -F01 / J, i11 = F00 / J;
            float P00 = mu * (F00 - i00) + lam * lJ * i00;
            float P10 = mu * (F10 - i01) + lam * lJ * i01;
            float P01 = mu * (F01 - i10) + lam * lJ * i10;
            float P11 = mu * (F11 - i11) + lam * lJ * i11;
            float I1 = F00 * F00 + F10 * F10 + F01 * F01 + F11 * F11;
            float pe = 0.5f * mu * (I1 - 2) - mu * lJ + 0.5f * lam * lJ * lJ;
            V2 f[3];
            f[0] = {-vol0 * (P00 * B00 + P01 * B01), -vol0 * (P10 * B00 + P11 * B01)};
            f[1] = {-vol0 * (P00 * B10 + P01 * B11), -vol0 * (P10 * B10 + P11 * B11)};
            f[2] = {-f[0].x - f[1].x, -f[0].y - f[1].y};
            for (int k = 0; k < 3; ++k) {
                force[p[k]].x += f[k].x; force[p[k]].y += f[k].y;
                feaVirial += f[k].x * unfolded[p[k]].x + f[k].y * unfolded[p[k]].y;
            }
            potEnergy += pe * vol0;
        }
    }

    // femmd2D/src/interactions.cpp:347-389 with the omp pragma of :354;
    // energy sums are made race-free with per-thread partials.
    void pairForces() {
        float U = 0, W = 0, CW = 0;
        int nt = ntypes;
#pragma omp parallel num_threads(nthreads) reduction(+ : U, W, CW)
        {
#pragma omp for schedule(static)
            for (int pi = 0; pi < n; ++pi) {
                for (int q = 0; q < nbrCount[pi]; ++q) {
                    int pj = nbrList[(size_t)pi * maxNbr + q];
                    int ti = type[pi], tj = type[pj];
                    float e = eps[ti * nt + tj], rcij = rc[ti * nt + tj];
                    float us = ush[ti * nt + tj];
                    float sij = (sigma[pi] + sigma[pj]) / 2;
                    float dx = pos[pi].x - pos[pj].x, dy = pos[pi].y - pos[pj].y;
                    min_image(dx, Lx); min_image(dy, Ly);
                    float dr2 = dx * dx + dy * dy;
                    float rc2 = rcij * rcij * sij * sij;
                    if (dr2 < rc2) {
                        float c = 48.0f * e / (sij * sij);
                        float dr2i = sij * sij / dr2;
                        float dr6i = dr2i * dr2i * dr2i;
                        float fc = c * dr6i * (dr6i - 0.5f) * dr2i;
                        float fx = fc * dx, fy = fc * dy;
                        force[pi].x += fx; force[pi].y += fy;
                        U += 0.5f * (4.0f * e * dr6i * (dr6i - 1.0f) + us);
                        W += 0.5f * fc * dr2;
                        int ci = cluster_of[pi], cj = cluster_of[pj];
                        if (!clusters.empty() && ci >= 0 && cj >= 0) {
                            float cx = clusters[ci].cm.x - clusters[cj].cm.x;
                            float cy = clusters[ci].cm.y - clusters[cj].cm.y;
                            min_image(cx, Lx); min_image(cy, Ly);
                            CW += 0.5f * (fx * cx + fy * cy);
                        }
                    }
                }
            }
        }
        potEnergy += U; virial += W; clusterVirial += CW;
    }

    void calcForces() {
        if (!bondLen.empty()) bondForces();
        if (!tlam.empty()) feaForces();
        if (lj_on) pairForces();
    }

    void zeroStuff() {
        for (int i = 0; i < n; ++i) { acc[i] = {0, 0}; force[i] = {0, 0}; }
    }
    void updateAcc() {
        for (int i = 0; i < n; ++i) {
            acc[i].x = force[i].x / mass[i];
            acc[i].y = force[i].y / mass[i];
        }
    }

    // femmd2D/src/md3dsystem.cpp:151-179
    void evalProps() {
        for (int p = 0; p < n; ++p) {
            vv[p] = vel[p].x * vel[p].x + vel[p].y * vel[p].y;
            vvMax = std::max(vvMax, vv[p]);
            kinEneSum += mass[p] * vv[p];
        }
        for (auto &cl : clusters)
            clusterKin += cl.mass * (cl.cmv.x * cl.cmv.x + cl.cmv.y * cl.cmv.y);
        dispHi += std::sqrt(vvMax) * dt;
    }

    // femmd2D/src/integrator.cpp:8-29
    void doStep() {
        potEnergy = virial = feaVirial = clusterVirial = 0;
        kinEneSum = vvMax = clusterKin = 0;
        float hdt = 0.5f * dt;
        for (int i = 0; i < n; ++i) {  // leapFrogStep1
            vel[i].x += hdt * acc[i].x; vel[i].y += hdt * acc[i].y;
            pos[i].x += dt * vel[i].x; pos[i].y += dt * vel[i].y;
        }
        for (int i = 0; i < n; ++i) {  // applyPBC
            pos[i].x = wrap1(pos[i].x, Lx);
            pos[i].y = wrap1(pos[i].y, Ly);
        }
        zeroStuff();
        updateNN(false);
        calcForces();
        updateAcc();  // integratorStep2
        for (int i = 0; i < n; ++i) {
            vel[i].x += hdt * acc[i].x; vel[i].y += hdt * acc[i].y;
        }
        evalProps();
    }
};

extern "C" {

or_baseline *or_bl_create(const or_system *s, const float *vx, const float *vy,
                          const float *mass, float dt, float skin,
                          int32_t nthreads) {
    if (!s || s->n <= 0) return nullptr;
    auto *b = new or_baseline();
    int n = s->n;
    b->n = n;
    b->nthreads = nthreads > 0 ? nthreads : 1;
    b->Lx = (float)s->Lx; b->Ly = (float)s->Ly;
    b->dt = dt; b->skin = skin;
    b->pos.resize(n); b->vel.resize(n); b->acc.resize(n); b->force.resize(n);
    b->refPos.resize(n); b->unfolded.resize(n);
    b->mass.resize(n); b->sigma.resize(n); b->vv.resize(n);
    b->type.resize(n); b->excl.resize(n); b->cluster_of.assign(n, -1);
    for (int i = 0; i < n; ++i) {
        b->pos[i] = {(float)s->x[i], (float)s->y[i]};
        b->vel[i] = {vx[i], vy[i]};
        b->mass[i] = mass[i];
        b->sigma[i] = (float)s->sigma[i];
        b->type[i] = s->type[i];
        b->excl[i] = s->excl[i];
        b->refPos[i] = {(float)(s->rx ? s->rx[i] : 0), (float)(s->ry ? s->ry[i] : 0)};
    }
    b->ntypes = s->ntypes;
    int nt2 = s->ntypes * s->ntypes;
    float rcmax = 0;
    for (int k = 0; k < nt2; ++k) {
        b->eps.push_back((float)s->eps[k]);
        b->rc.push_back((float)s->rc[k]);
        b->ush.push_back((float)s->ushift[k]);
        rcmax = std::max(rcmax, (float)s->rc[k]);
    }
    b->lj_on = s->lj_on != 0;
    for (int l = 0; l < s->nbond; ++l) {
        b->bondList.push_back(s->bi[l]);
        b->bondList.push_back(s->bj[l]);
        b->bondLen.push_back((float)s->br0[l]);
        b->bondK.push_back((float)s->bk[l]);
    }
    int gmax = -1;
    for (int i = 0; i < n; ++i) gmax = std::max(gmax, b->excl[i]);
    for (int e = 0; e < s->ntri; ++e) {
        for (int k = 0; k < 3; ++k) b->tri.push_back(s->tri[3 * e + k]);
        double lam, mu;
        lame(s->tE[e], s->tnu[e], lam, mu);
        b->tlam.push_back((float)lam);
        b->tmu.push_back((float)mu);
    }
    // Clusters: one per exclusion group of FEA nodes (each elastomer gets
    // its own group in femmd2D/src/particles.cpp:299,312).
    if (s->ntri > 0) {
        std::vector<char> isnode(n, 0);
        for (int v : b->tri) isnode[v] = 1;
        b->clusters.resize(gmax + 1);
        for (int i = 0; i < n; ++i)
            if (isnode[i] && b->excl[i] >= 0) {
                b->clusters[b->excl[i]].nodes.push_back(i);
                b->clusters[b->excl[i]].mass += b->mass[i];
                b->cluster_of[i] = b->excl[i];
            }
        b->clusters.erase(std::remove_if(b->clusters.begin(), b->clusters.end(),
                                         [](const Cluster &c) { return c.nodes.empty(); }),
                          b->clusters.end());
        for (int i = 0; i < n; ++i) b->cluster_of[i] = -1;
        for (size_t c = 0; c < b->clusters.size(); ++c)
            for (int p : b->clusters[c].nodes) b->cluster_of[p] = (int)c;
    }
    // femmd2D/src/nn.cpp:200-225: cells of side >= rc + skin (sigma
    // ignored, :196), capacities from the reference heuristics; the cell
    // capacity is raised to the actual maximum so that no write overflows.
    float nnR = rcmax + skin;
    b->nnRadiusSq = nnR * nnR;
    b->ncx = (int)std::floor(b->Lx / nnR);
    b->ncy = (int)std::floor(b->Ly / nnR);
    if (b->ncx < 3 || b->ncy < 3) { delete b; return nullptr; }
    float vcell = (b->Lx / b->ncx) * (b->Ly / b->ncy);
    float dens = std::max(1.0f, n / (b->Lx * b->Ly));
    b->maxPPC = std::max(32, (int)(2 * dens * vcell));
    b->maxNbr = std::max(64, (int)(2 * 5.0f * nnR * nnR * nnR * dens));
    b->cellCount.assign((size_t)b->ncx * b->ncy, 0);
    b->cellContent.assign((size_t)b->ncx * b->ncy * b->maxPPC, 0);
    b->nbrList.assign((size_t)n * b->maxNbr, 0);
    b->nbrCount.assign(n, 0);
    // Integrator::run prologue, femmd2D/src/integrator.cpp:39-43
    b->updateNN(true);
    b->zeroStuff();
    b->calcForces();
    b->updateAcc();
    return b;
}

double or_bl_run(or_baseline *b, int32_t nsteps) {
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < nsteps; ++i) b->doStep();
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

void or_bl_get(or_baseline *b, float *x, float *y, float *vx, float *vy,
               double *props) {
    for (int i = 0; i < b->n; ++i) {
        if (x) x[i] = b->pos[i].x;
        if (y) y[i] = b->pos[i].y;
        if (vx) vx[i] = b->vel[i].x;
        if (vy) vy[i] = b->vel[i].y;
    }
    if (props) {
        props[0] = b->kinEneSum;
        props[1] = b->potEnergy;
        props[2] = b->virial;
        props[3] = b->feaVirial;
        props[4] = b->updates;
    }
}

void or_bl_destroy(or_baseline *b) { delete b; }

int32_t or_bl_threads(void) {
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}

}  // extern "C"
