// capi.hip -- the extern "C" boundary declared in include/femmd.h.
//
// Argument checking, caller-order bookkeeping before fm_commit and status
// codes; all device work is delegated to fm::Engine.  Errors never exit
// the process (the reference's HANDLE_ERROR exits, femmd2D/src/CudaUtil.h:4-14).
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <set>
#include <vector>

#include "handles.hpp"

namespace {

int fail(fm_system *h, int code, const std::string &msg) {
    if (h) h->err = msg;
    return code;
}
int failb(fm_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

#define FM_GUARD(h)                                              \
    if (!(h)) return FM_ERR_ARG;                                 \
    try {
#define FM_END(h)                                                \
    }                                                            \
    catch (const std::bad_alloc &) {                             \
        return fail(h, FM_ERR_NOMEM, "host allocation failed");  \
    }                                                            \
    catch (const std::exception &ex) {                           \
        return fail(h, FM_ERR_ARG, ex.what());                   \
    }

bool frozen(fm_system *h) { return h->eng != nullptr; }

// femmd2D/src/def.h:91-93 minimum image in double (bond rest lengths)
double min_image(double d, double L) {
    if (d >= 0.5 * L) return d - L;
    if (d < -0.5 * L) return d + L;
    return d;
}

}  // namespace

extern "C" {

int fm_device_count(int32_t *n) {
    if (!n) return FM_ERR_ARG;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *n = (e == hipSuccess) ? c : 0;
    return FM_OK;
}

int fm_create(const fm_params *p, fm_system **out) {
    if (!p || !out) return FM_ERR_ARG;
    *out = nullptr;
    if (!(p->Lx > 0) || !(p->Ly > 0) || !(p->dt > 0)) return FM_ERR_ARG;
    fm_system *h = new (std::nothrow) fm_system();
    if (!h) return FM_ERR_NOMEM;
    h->hs.Lx = p->Lx;
    h->hs.Ly = p->Ly;
    h->hs.dt = p->dt;
    h->hs.device = p->device;
    *out = h;
    return FM_OK;
}

void fm_destroy(fm_system *h) { delete h; }

const char *fm_last_error(const fm_system *h) { return h ? h->err.c_str() : "null handle"; }

int fm_add_particles(fm_system *h, int32_t n, const float *x, const float *y, const float *vx,
                     const float *vy, const float *mass, const float *sigma, const int32_t *type,
                     int32_t *first_id) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (n < 0 || (n > 0 && (!x || !y || !mass || !sigma))) return fail(h, FM_ERR_ARG, "bad particle arrays");
    fm::HostSystem &S = h->hs;
    // validate every input before anything is appended (all or nothing)
    for (int32_t k = 0; k < n; ++k) {
        if (!std::isfinite(x[k]) || !std::isfinite(y[k])) return fail(h, FM_ERR_ARG, "non-finite position");
        if ((vx && !std::isfinite(vx[k])) || (vy && !std::isfinite(vy[k])))
            return fail(h, FM_ERR_ARG, "non-finite velocity");
        if (!(sigma[k] > 0) || !std::isfinite(sigma[k])) return fail(h, FM_ERR_ARG, "sigma must be positive");
        if (std::isnan(mass[k])) return fail(h, FM_ERR_ARG, "mass is NaN");
        if (type && type[k] < 0) return fail(h, FM_ERR_ARG, "negative particle type");
    }
    if ((int64_t)S.x.size() + n > INT32_MAX) return fail(h, FM_ERR_ARG, "too many particles");
    if (first_id) *first_id = (int32_t)S.x.size();
    for (int32_t k = 0; k < n; ++k) {
        S.x.push_back(x[k]);
        S.y.push_back(y[k]);
        S.vx.push_back(vx ? vx[k] : 0.0f);
        S.vy.push_back(vy ? vy[k] : 0.0f);
        S.mass.push_back(mass[k]);
        S.sigma.push_back(sigma[k]);
        S.type.push_back(type ? type[k] : 0);
        S.excl.push_back(-1);
        S.rx.push_back(x[k]);
        S.ry.push_back(y[k]);
    }
    return FM_OK;
    FM_END(h)
}

// femmd2D/src/particles.cpp:295-385 (without the implicit edge bonds and
// with the per-element area in the mass split, see femmd.h).
int fm_add_elastomer(fm_system *h, int32_t nv, const float *verts, int32_t nt, const int32_t *tris,
                     int32_t index_base, float cx, float cy, int32_t type, float sigma,
                     float node_mass, float total_mass, int32_t exclude, const float *E_per_tri,
                     const float *nu_per_tri, int32_t *first_node) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (nv <= 0 || nt < 0 || !verts || (nt > 0 && !tris)) return fail(h, FM_ERR_ARG, "bad mesh arrays");
    if (!(sigma > 0)) return fail(h, FM_ERR_ARG, "sigma must be positive");
    fm::HostSystem &S = h->hs;
    const int32_t off = (int32_t)S.x.size();
    // validate every input before anything is appended or a group is taken
    // (all or nothing)
    for (int32_t k = 0; k < 3 * nt; ++k) {
        const int32_t v = tris[k] - index_base;
        if (v < 0 || v >= nv) return fail(h, FM_ERR_ARG, "triangle index out of range");
    }
    for (int32_t k = 0; k < nv; ++k) {
        const float px = verts[2 * k] + cx, py = verts[2 * k + 1] + cy;
        if (!std::isfinite(px) || !std::isfinite(py)) return fail(h, FM_ERR_ARG, "non-finite vertex");
    }
    if (type < 0) return fail(h, FM_ERR_ARG, "negative particle type");
    for (int32_t e = 0; e < nt; ++e) {
        // setElasticConstants (femmd2D/src/interactions.cpp:43-50): lambda and
        // mu are finite only for nu in (-1, 1/2)
        if (E_per_tri && !(E_per_tri[e] >= 0 && std::isfinite(E_per_tri[e])))
            return fail(h, FM_ERR_ARG, "per-element E must be finite and >= 0");
        if (nu_per_tri && !(nu_per_tri[e] > -1.0f && nu_per_tri[e] < 0.5f))
            return fail(h, FM_ERR_ARG, "per-element nu must lie in (-1, 0.5)");
    }
    if ((int64_t)off + nv > INT32_MAX) return fail(h, FM_ERR_ARG, "too many particles");
    std::vector<double> m(nv, 0.0);
    if (node_mass > 0) {
        std::fill(m.begin(), m.end(), (double)node_mass);
    } else if (total_mass > 0 && nt > 0) {
        double area = 0.0;
        std::vector<double> a(nt);
        for (int32_t e = 0; e < nt; ++e) {
            const int i0 = tris[3 * e] - index_base, i1 = tris[3 * e + 1] - index_base,
                      i2 = tris[3 * e + 2] - index_base;
            const double x0 = verts[2 * i0], y0 = verts[2 * i0 + 1];
            const double x1 = verts[2 * i1], y1 = verts[2 * i1 + 1];
            const double x2 = verts[2 * i2], y2 = verts[2 * i2 + 1];
            a[e] = 0.5 * ((y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0));
            area += a[e];
        }
        const double rho = total_mass / area;
        for (int32_t e = 0; e < nt; ++e)
            for (int k = 0; k < 3; ++k) m[tris[3 * e + k] - index_base] += rho * a[e] / 3.0;
    } else {
        return fail(h, FM_ERR_ARG, "node_mass or total_mass must be positive");
    }
    for (int32_t k = 0; k < nv; ++k)
        if (!std::isfinite(m[k])) return fail(h, FM_ERR_ARG, "non-finite node mass");
    const int32_t group = exclude ? S.next_group++ : -1;
    if (first_node) *first_node = off;
    for (int32_t k = 0; k < nv; ++k) {
        const float px = verts[2 * k] + cx, py = verts[2 * k + 1] + cy;
        S.x.push_back(px);
        S.y.push_back(py);
        S.vx.push_back(0.0f);
        S.vy.push_back(0.0f);
        S.mass.push_back((float)m[k]);
        S.sigma.push_back(sigma);
        S.type.push_back(type);
        S.excl.push_back(group);
        S.rx.push_back(px);
        S.ry.push_back(py);
    }
    S.cl_first.push_back(off);
    S.cl_n.push_back(nv);
    for (int32_t e = 0; e < nt; ++e) {
        for (int k = 0; k < 3; ++k) S.tri.push_back(off + tris[3 * e + k] - index_base);
        S.tE.push_back(E_per_tri ? E_per_tri[e] : -1.0f);
        S.tnu.push_back(nu_per_tri ? nu_per_tri[e] : -1.0f);
    }
    return FM_OK;
    FM_END(h)
}

int fm_add_bonds(fm_system *h, int32_t nb, const int32_t *i, const int32_t *j, const float *k,
                 const float *r0) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (nb < 0 || (nb > 0 && (!i || !j || !k))) return fail(h, FM_ERR_ARG, "bad bond arrays");
    fm::HostSystem &S = h->hs;
    const int32_t n = (int32_t)S.x.size();
    // validate every bond before any is appended (all or nothing)
    for (int32_t l = 0; l < nb; ++l) {
        if (i[l] < 0 || i[l] >= n || j[l] < 0 || j[l] >= n || i[l] == j[l])
            return fail(h, FM_ERR_ARG, "bond end out of range");
        if (!std::isfinite(k[l])) return fail(h, FM_ERR_ARG, "bond stiffness must be finite");
        if (r0 && !(std::isfinite(r0[l]) || (std::isinf(r0[l]) && r0[l] < 0)))
            return fail(h, FM_ERR_ARG, "bond rest length is NaN or +inf");
    }
    for (int32_t l = 0; l < nb; ++l) {
        float len = r0 ? r0[l] : -1.0f;
        if (len < 0) {
            // Bonds::createBond (femmd2D/src/bonds.cpp:17-20): current length
            const double dx = min_image((double)S.x[i[l]] - S.x[j[l]], S.Lx);
            const double dy = min_image((double)S.y[i[l]] - S.y[j[l]], S.Ly);
            len = (float)std::sqrt(dx * dx + dy * dy);
        }
        S.bi.push_back(i[l]);
        S.bj.push_back(j[l]);
        S.bk.push_back(k[l]);
        S.br0.push_back(len);
    }
    return FM_OK;
    FM_END(h)
}

// Particles::addElastomer's edge bonds (femmd2D/src/particles.cpp:335-337:
// createBond on the three edges of every triangle, shared edges twice),
// explicit and opt-in; dedup keeps the first bond of each edge.
int fm_add_edge_bonds(fm_system *h, int32_t first_node, int32_t nt, const int32_t *tris,
                      int32_t index_base, float k, int32_t dedup) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (nt < 0 || (nt > 0 && !tris)) return fail(h, FM_ERR_ARG, "bad triangle array");
    if (!std::isfinite(k)) return fail(h, FM_ERR_ARG, "bond stiffness must be finite");
    const int64_t n = (int64_t)h->hs.x.size();
    std::vector<int32_t> bi, bj;
    bi.reserve(3 * (size_t)nt);
    bj.reserve(3 * (size_t)nt);
    std::set<std::pair<int32_t, int32_t>> seen;
    for (int32_t t = 0; t < nt; ++t) {
        int32_t v[3];
        for (int q = 0; q < 3; ++q) {
            const int64_t u = (int64_t)first_node + tris[3 * t + q] - index_base;
            if (u < 0 || u >= n) return fail(h, FM_ERR_ARG, "triangle node out of range");
            v[q] = (int32_t)u;
        }
        const int32_t e[3][2] = {{v[0], v[1]}, {v[0], v[2]}, {v[1], v[2]}};
        for (const auto &ed : e) {
            if (ed[0] == ed[1]) return fail(h, FM_ERR_ARG, "degenerate triangle");
            if (dedup && !seen.insert({std::min(ed[0], ed[1]), std::max(ed[0], ed[1])}).second)
                continue;
            bi.push_back(ed[0]);
            bj.push_back(ed[1]);
        }
    }
    const std::vector<float> kk(bi.size(), k);
    return fm_add_bonds(h, (int32_t)bi.size(), bi.data(), bj.data(), kk.data(), nullptr);
    FM_END(h)
}

int fm_set_pair_table(fm_system *h, int32_t ntypes, const float *eps, const float *rcut,
                      const float *ushift) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (ntypes < 1 || ntypes > fm::kMaxPairTypes || !eps || !rcut || !ushift)
        return fail(h, FM_ERR_ARG, "ntypes must be in [1, 16] with three tables");
    const int n2 = ntypes * ntypes;
    for (int q = 0; q < n2; ++q)
        if (!(rcut[q] > 0)) return fail(h, FM_ERR_ARG, "cutoffs must be positive");
    h->hs.ntypes = ntypes;
    h->hs.eps.assign(eps, eps + n2);
    h->hs.rc.assign(rcut, rcut + n2);
    h->hs.ush.assign(ushift, ushift + n2);
    return FM_OK;
    FM_END(h)
}

int fm_set_material(fm_system *h, int32_t model, float E, float nu) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (model < 0 || model > 2) return fail(h, FM_ERR_ARG, "unknown material model");
    if (!(E >= 0) || !(nu > -1.0f) || !(nu < 0.5f)) return fail(h, FM_ERR_ARG, "bad E or nu");
    h->hs.model = model;
    h->hs.E = E;
    h->hs.nu = nu;
    return FM_OK;
    FM_END(h)
}

int fm_set_dt(fm_system *h, double dt) {
    FM_GUARD(h)
    if (!(dt > 0)) return fail(h, FM_ERR_ARG, "dt must be positive");
    h->hs.dt = dt;
    if (h->eng) return h->eng->set_dt(0, dt, h->err);
    return FM_OK;
    FM_END(h)
}

int fm_set_lj(fm_system *h, int32_t on) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    h->hs.lj_on = on != 0;
    return FM_OK;
    FM_END(h)
}

int fm_set_skin(fm_system *h, float skin) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (!(skin >= 0.0f) || !std::isfinite(skin)) return fail(h, FM_ERR_ARG, "skin must be >= 0");
    h->hs.skin = skin;
    return FM_OK;
    FM_END(h)
}

int fm_set_reorder_interval(fm_system *h, int32_t steps) {
    FM_GUARD(h)
    if (steps < 0) return fail(h, FM_ERR_ARG, "steps must be >= 0");
    h->hs.reorder_every = steps;
    if (h->eng) h->eng->set_reorder_interval(steps);
    return FM_OK;
    FM_END(h)
}

int fm_reorder_info(fm_system *h, int32_t *interval, int64_t *count) {
    FM_GUARD(h)
    if (interval) *interval = h->eng ? h->eng->reorder_interval() : h->hs.reorder_every;
    if (count) *count = h->eng ? h->eng->reorders() : 0;
    return FM_OK;
    FM_END(h)
}

int fm_set_exclusion(fm_system *h, int32_t i, int32_t group) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (i < 0 || i >= (int32_t)h->hs.x.size() || group < -1) return fail(h, FM_ERR_ARG, "bad index");
    h->hs.excl[i] = group;
    if (group >= h->hs.next_group) h->hs.next_group = group + 1;
    return FM_OK;
    FM_END(h)
}

int fm_set_mass(fm_system *h, int32_t n, const int32_t *ids, const float *mass) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    if (n < 0 || (n > 0 && (!ids || !mass))) return fail(h, FM_ERR_ARG, "bad mass arrays");
    for (int32_t k = 0; k < n; ++k)
        if (ids[k] < 0 || ids[k] >= (int32_t)h->hs.x.size()) return fail(h, FM_ERR_ARG, "bad index");
    for (int32_t k = 0; k < n; ++k) {
        h->hs.mass[ids[k]] = mass[k];
        if (!(mass[k] > 0) || !std::isfinite(mass[k])) {
            h->hs.vx[ids[k]] = 0.0f;
            h->hs.vy[ids[k]] = 0.0f;
        }
    }
    return FM_OK;
    FM_END(h)
}

int fm_commit(fm_system *h) {
    FM_GUARD(h)
    if (frozen(h)) return fail(h, FM_ERR_STATE, "system already committed");
    std::vector<const fm::HostSystem *> v{&h->hs};
    return fm::Engine::build(v, h->eng, h->err);
    FM_END(h)
}

#define FM_NEED_ENGINE(h) \
    if (!(h)->eng) return fail(h, FM_ERR_STATE, "call fm_commit first")

int fm_reorder(fm_system *h) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->reorder(h->err);
    FM_END(h)
}

int fm_step(fm_system *h, int32_t nsteps) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->step(nsteps, h->err);
    FM_END(h)
}

int fm_compute_forces(fm_system *h) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->compute_forces(h->err);
    FM_END(h)
}

int fm_get_energies(fm_system *h, fm_energies *e) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (!e) return fail(h, FM_ERR_ARG, "null output");
    return h->eng->energies(0, e, h->err);
    FM_END(h)
}

int fm_set_observe_every(fm_system *h, int32_t every) {
    FM_GUARD(h)
    if (every < 0 || every > 1) return fail(h, FM_ERR_ARG, "every must be 0 or 1");
    h->hs.observe_every = every;
    if (h->eng) return h->eng->set_observe_every(every, h->err);
    return FM_OK;
    FM_END(h)
}

int fm_get_averages(fm_system *h, fm_averages *a) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (!a) return fail(h, FM_ERR_ARG, "null output");
    return h->eng->averages(0, a, h->err);
    FM_END(h)
}

int fm_reset_averages(fm_system *h) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->reset_averages(h->err);
    FM_END(h)
}

int fm_get_state(fm_system *h, float *x, float *y, float *vx, float *vy) {
    FM_GUARD(h)
    if (!h->eng) {
        const fm::HostSystem &S = h->hs;
        const size_t n = S.x.size();
        if (x) std::memcpy(x, S.x.data(), n * sizeof(float));
        if (y) std::memcpy(y, S.y.data(), n * sizeof(float));
        if (vx) std::memcpy(vx, S.vx.data(), n * sizeof(float));
        if (vy) std::memcpy(vy, S.vy.data(), n * sizeof(float));
        return FM_OK;
    }
    return h->eng->get_state(0, x, y, vx, vy, h->err);
    FM_END(h)
}

int fm_set_state(fm_system *h, const float *x, const float *y, const float *vx, const float *vy) {
    FM_GUARD(h)
    const size_t n = h->hs.x.size();
    for (size_t u = 0; u < n; ++u) {
        if ((x && !std::isfinite(x[u])) || (y && !std::isfinite(y[u])) ||
            (vx && !std::isfinite(vx[u])) || (vy && !std::isfinite(vy[u])))
            return fail(h, FM_ERR_ARG, "non-finite state");
    }
    if (!h->eng) {
        fm::HostSystem &S = h->hs;
        if (x) std::memcpy(S.x.data(), x, n * sizeof(float));
        if (y) std::memcpy(S.y.data(), y, n * sizeof(float));
        if (vx) std::memcpy(S.vx.data(), vx, n * sizeof(float));
        if (vy) std::memcpy(S.vy.data(), vy, n * sizeof(float));
        return FM_OK;
    }
    return h->eng->set_state(0, x, y, vx, vy, h->err);
    FM_END(h)
}

int fm_get_forces(fm_system *h, float *fx, float *fy) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->get_forces(0, fx, fy, h->err);
    FM_END(h)
}

int fm_num_particles(const fm_system *h, int32_t *n) {
    if (!h || !n) return FM_ERR_ARG;
    *n = (int32_t)h->hs.x.size();
    return FM_OK;
}

int fm_num_bonds(const fm_system *h, int32_t *nb) {
    if (!h || !nb) return FM_ERR_ARG;
    *nb = (int32_t)h->hs.bi.size();
    return FM_OK;
}

int fm_get_bonds(const fm_system *h, int32_t *pairs) {
    if (!h || !pairs) return FM_ERR_ARG;
    for (size_t l = 0; l < h->hs.bi.size(); ++l) {
        pairs[2 * l] = h->hs.bi[l];
        pairs[2 * l + 1] = h->hs.bj[l];
    }
    return FM_OK;
}

int fm_num_elements(const fm_system *h, int32_t *nt) {
    if (!h || !nt) return FM_ERR_ARG;
    *nt = (int32_t)(h->hs.tri.size() / 3);
    return FM_OK;
}

int fm_sync(fm_system *h) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->sync(h->err);
    FM_END(h)
}

int fm_set_tuning(fm_system *h, int32_t tile_cells, int32_t block, int32_t use_graph) {
    FM_GUARD(h)
    if (tile_cells < 0 || tile_cells > fm::kMaxTileW) return fail(h, FM_ERR_ARG, "tile_cells in [0, 64]");
    if (block != 0 && block != 64 && block != 128 && block != 192 && block != 256)
        return fail(h, FM_ERR_ARG, "block must be 0, 64, 128, 192 or 256");
    if (tile_cells > 0) h->hs.tile_w = tile_cells;
    if (block > 0) h->hs.block = block;
    if (use_graph >= 0) h->hs.use_graph = use_graph;
    if (h->eng) return h->eng->set_tuning(tile_cells, block, use_graph, h->err);
    return FM_OK;
    FM_END(h)
}

uint32_t fm_device_errors(fm_system *h) {
    if (!h || !h->eng) return 0;
    return h->eng->errors();
}

int fm_clear_errors(fm_system *h) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->clear_errors();
    FM_END(h)
}

int fm_get_cells(fm_system *h, int32_t *ncx, int32_t *ncy, int32_t *key, int32_t *start,
                 int32_t *perm) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->get_cells(0, ncx, ncy, key, start, perm, h->err);
    FM_END(h)
}

int fm_pair_sets(fm_system *h, int32_t *cnt, int64_t *s1, int64_t *s2) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (!cnt || !s1 || !s2) return fail(h, FM_ERR_ARG, "null output");
    return h->eng->pair_sets(0, cnt, s1, s2, h->err);
    FM_END(h)
}

int fm_time_forces(fm_system *h, int32_t reps, double *ms) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (reps <= 0 || !ms) return fail(h, FM_ERR_ARG, "reps > 0 and ms required");
    return h->eng->time_forces(reps, ms, h->err);
    FM_END(h)
}

int fm_profile_step(fm_system *h, int32_t nsteps, double *ms) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (nsteps <= 0 || !ms) return fail(h, FM_ERR_ARG, "nsteps > 0 and ms[6] required");
    return h->eng->profile(nsteps, ms, h->err);
    FM_END(h)
}

int fm_list_stats(fm_system *h, double *mean_len, int32_t *max_len, int64_t *rebuilds) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->list_stats(0, mean_len, max_len, rebuilds, h->err);
    FM_END(h)
}

int fm_cluster_count(fm_system *h, int32_t *nc) {
    FM_GUARD(h)
    if (!nc) return fail(h, FM_ERR_ARG, "null output");
    *nc = h->eng ? h->eng->ncluster(0) : (int32_t)h->hs.cl_first.size();
    return FM_OK;
    FM_END(h)
}

int fm_cluster_props(fm_system *h, double *props, float *ux, float *uy) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->cluster_props(0, props, ux, uy, h->err);
    FM_END(h)
}

int fm_cluster_virial(fm_system *h, double *virial) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (!virial) return fail(h, FM_ERR_ARG, "null output");
    return h->eng->cluster_virial(0, virial, h->err);
    FM_END(h)
}

int fm_set_temperature(fm_system *h, double temp, int32_t n, const int32_t *ids) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (n < 0 || (n > 0 && !ids)) return fail(h, FM_ERR_ARG, "n >= 0 and ids for n > 0");
    return h->eng->set_temperature(0, temp, n, ids, h->err);
    FM_END(h)
}

int fm_set_temperature_noalloc(fm_system *h, double temp, int32_t n, const int32_t *ids) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (n < 0 || (n > 0 && !ids)) return fail(h, FM_ERR_ARG, "n >= 0 and ids for n > 0");
    return h->eng->set_temperature_noalloc(0, temp, n, ids, h->err);
    FM_END(h)
}

int fm_force_jvp(fm_system *h, const float *wx, const float *wy, float *jx, float *jy,
                 double *dE, double *deps, double *dsig) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    if (!wx || !wy) return fail(h, FM_ERR_ARG, "wx and wy are required");
    return h->eng->adjoint_op(wx, wy, jx, jy, dE, deps, dsig, h->err);
    FM_END(h)
}

int fm_adjoint(fm_system *h, int32_t nsteps, int32_t ckpt, int32_t nloss, const int32_t *ids,
               const float *tx, const float *ty, double *loss, double *dE, double *deps,
               double *dsig) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->adjoint_run(nsteps, ckpt, nloss, ids, tx, ty, loss, dE, deps, dsig, h->err);
    FM_END(h)
}

int fm_adjoint_forward(fm_system *h, int32_t nsteps, int32_t ckpt) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    return h->eng->adjoint_forward(nsteps, ckpt, h->err);
    FM_END(h)
}

int fm_adjoint_backward(fm_system *h, const float *dL_dx, const float *dL_dy, double *dE,
                        double *deps, double *dsig) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    const int32_t n = (int32_t)h->hs.x.size();
    for (int32_t u = 0; u < n; ++u)
        if ((dL_dx && !std::isfinite(dL_dx[u])) || (dL_dy && !std::isfinite(dL_dy[u])))
            return fail(h, FM_ERR_ARG, "non-finite cotangent");
    return h->eng->adjoint_backward(dL_dx, dL_dy, dE, deps, dsig, h->err);
    FM_END(h)
}

int fm_adjoint_run(fm_system *h, int32_t nsteps, int32_t ckpt, const float *dL_dx,
                   const float *dL_dy, double *dE, double *deps, double *dsig) {
    FM_GUARD(h)
    FM_NEED_ENGINE(h);
    const int rc = h->eng->adjoint_forward(nsteps, ckpt, h->err);
    if (rc) return rc;
    return fm_adjoint_backward(h, dL_dx, dL_dy, dE, deps, dsig);
    FM_END(h)
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
int fm_batch_create(int32_t n, fm_system *const *systems, fm_batch **out) {
    if (n <= 0 || !systems || !out) return FM_ERR_ARG;
    *out = nullptr;
    fm_batch *b = new (std::nothrow) fm_batch();
    if (!b) return FM_ERR_NOMEM;
    try {
        b->hs.reserve(n);
        for (int32_t r = 0; r < n; ++r) {
            fm_system *s = systems[r];
            if (!s) { delete b; return FM_ERR_ARG; }
            b->hs.push_back(s->hs);
            if (s->eng) {
                // take the live state of a committed system
                fm::HostSystem &H = b->hs.back();
                std::string e;
                int rc = s->eng->get_state(0, H.x.data(), H.y.data(), H.vx.data(), H.vy.data(), e);
                if (rc) { delete b; return rc; }
            }
        }
        std::vector<const fm::HostSystem *> v;
        for (auto &h : b->hs) v.push_back(&h);
        int rc = fm::Engine::build(v, b->eng, b->err);
        if (rc) {
            // keep the handle so the caller can read the error text
            *out = b;
            return rc;
        }
    } catch (const std::bad_alloc &) {
        delete b;
        return FM_ERR_NOMEM;
    }
    *out = b;
    return FM_OK;
}

void fm_batch_destroy(fm_batch *b) { delete b; }

const char *fm_batch_last_error(const fm_batch *b) { return b ? b->err.c_str() : "null handle"; }

#define FM_NEED_BENGINE(b) \
    if (!(b)->eng) return failb(b, FM_ERR_STATE, "batch build failed")

int fm_batch_step(fm_batch *b, int32_t nsteps) {
    if (!b) return FM_ERR_ARG;
    FM_NEED_BENGINE(b);
    return b->eng->step(nsteps, b->err);
}

int fm_batch_get_state(fm_batch *b, int32_t r, float *x, float *y, float *vx, float *vy) {
    if (!b) return FM_ERR_ARG;
    FM_NEED_BENGINE(b);
    return b->eng->get_state(r, x, y, vx, vy, b->err);
}

int fm_batch_get_energies(fm_batch *b, int32_t r, fm_energies *e) {
    if (!b || !e) return FM_ERR_ARG;
    FM_NEED_BENGINE(b);
    return b->eng->energies(r, e, b->err);
}

int fm_batch_get_averages(fm_batch *b, int32_t r, fm_averages *a) {
    if (!b || !a) return FM_ERR_ARG;
    FM_NEED_BENGINE(b);
    return b->eng->averages(r, a, b->err);
}

int fm_batch_set_tuning(fm_batch *b, int32_t tile_cells, int32_t block, int32_t use_graph) {
    if (!b) return FM_ERR_ARG;
    FM_NEED_BENGINE(b);
    if (tile_cells < 0 || tile_cells > fm::kMaxTileW) return failb(b, FM_ERR_ARG, "tile_cells in [0, 64]");
    if (block != 0 && block != 64 && block != 128 && block != 192 && block != 256)
        return failb(b, FM_ERR_ARG, "block must be 0, 64, 128, 192 or 256");
    return b->eng->set_tuning(tile_cells, block, use_graph, b->err);
}

int fm_batch_set_reorder_interval(fm_batch *b, int32_t steps) {
    if (!b) return FM_ERR_ARG;
    FM_NEED_BENGINE(b);
    if (steps < 0) return failb(b, FM_ERR_ARG, "steps must be >= 0");
    b->eng->set_reorder_interval(steps);
    return FM_OK;
}

uint32_t fm_batch_device_errors(fm_batch *b) {
    if (!b || !b->eng) return 0;
    return b->eng->errors();
}

}  // extern "C"
