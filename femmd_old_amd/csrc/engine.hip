// engine.hip -- build of the device image and step orchestration.
//
// Replaces the reference's per-step host loop (femmd2D/src/integrator.cpp:
// 8-57) and its upload / topology setup (femmd2D/src/particles.cpp:105-199,
// femmd2D/src/bonds.cpp:31-83, femmd2D/src/nn.cpp:192-250).  One step is
// five kernels on one stream (kick+bin, scan reduce, scan apply, scatter,
// forces) with no host synchronisation; the step loop is replayed from a
// captured HIP graph.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace fm {

namespace {

constexpr int kGraphSteps = 16;  // step bodies per captured graph of a run
// node kicks of at least this many workgroups gather through LDS (measured:
// 1950 workgroups, 500k nodes: faster; 364 workgroups: slower)
constexpr size_t kGatherLdsSpans = 1024;

float wrap_host(float r, float L) {  // femmd2D/src/def.h:112, FP32
    float q = r / L;
    float f = std::floor(q);
    float p = f * L;
    return r - p;
}

int cell_key_host(float x, float y, const RepDesc &R) {  // nnCuda.cu:9-17
    float tx = x * R.fncx;
    tx = tx / R.Lx;
    int ix = (int)tx;
    if (ix == R.ncx) ix -= 1;
    float ty = y * R.fncy;
    ty = ty / R.Ly;
    int iy = (int)ty;
    if (iy == R.ncy) iy -= 1;
    ix = std::min(std::max(ix, 0), R.ncx - 1);
    iy = std::min(std::max(iy, 0), R.ncy - 1);
    return ix + iy * R.ncx;
}

template <class T>
hipError_t upload(DBuf<T> &b, const std::vector<T> &v, hipStream_t s) {
    hipError_t e = b.alloc(std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

std::string dev_error_text(uint32_t e) {
    std::string s = "device error:";
    if (e & FM_DEV_CELL_OVERFLOW) s += " cell capacity exceeded;";
    if (e & FM_DEV_TILE_OVERFLOW) s += " cell tile exceeded LDS capacity;";
    if (e & FM_DEV_INVERTED) s += " inverted element (det F <= 0);";
    if (e & FM_DEV_NONFINITE) s += " non-finite coordinate;";
    if (e & FM_DEV_LIST_OVERFLOW) s += " neighbour list capacity exceeded;";
    if (e & FM_DEV_SORT_STALL) s += " counting sort look-back timed out;";
    return s;
}

}  // namespace

Engine::~Engine() {
    if (device_ >= 0) (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::drop_graph() {
    for (auto &g : graphs_) (void)hipGraphExecDestroy(g.second);
    graphs_.clear();
    return FM_OK;
}

int Engine::check(hipError_t e, const char *what, std::string &err) {
    if (e == hipSuccess) return FM_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return FM_ERR_HIP;
}

double Engine::density_of(int r) const {
    const RepDesc &R = reps_h_[r];
    return (double)(R.p1 - R.p0) / ((double)R.Lx * R.Ly);
}

int Engine::build(const std::vector<const HostSystem *> &reps, std::unique_ptr<Engine> &out,
                  std::string &err) {
    if (reps.empty()) { err = "no systems"; return FM_ERR_ARG; }
    std::unique_ptr<Engine> E(new Engine());
    E->device_ = reps[0]->device;
    int rc = E->check(hipSetDevice(E->device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = E->check(hipStreamCreateWithFlags(&E->stream_, hipStreamNonBlocking), "stream", err);
    if (rc) return rc;
    hipStream_t st = E->stream_;
    const int R = (int)reps.size();
    E->R_ = R;
    E->model_ = reps[0]->model;
    E->block_ = reps[0]->block;
    E->use_graph_ = reps[0]->use_graph;
    E->reorder_every_ = reps[0]->reorder_every;
    E->obs_every_ = reps[0]->observe_every;

    // neighbour-list mode: one skin for the batch
    const float skin = reps[0]->skin;
    bool any_lj = false;
    for (int r = 0; r < R; ++r) {
        if (reps[r]->skin != skin) { err = "all replicas of a batch must use one skin"; return FM_ERR_ARG; }
        any_lj |= reps[r]->lj_on;
    }
    if (!(skin >= 0.0f) || !std::isfinite(skin)) { err = "skin must be >= 0"; return FM_ERR_ARG; }
    E->list_ = skin > 0.0f && any_lj;
    E->skin_ = E->list_ ? skin : 0.0f;

    // ---- pass 1: validation and cell geometry per replica ----
    std::vector<RepDesc> rd(R);
    double rlmax = 0.0;
    std::vector<float> smin(R), smax(R);
    double amax = 0.0, rho_cp = 0.0;
    for (int r = 0; r < R; ++r) {
        const HostSystem &S = *reps[r];
        const size_t n = S.x.size();
        if (n == 0) { err = "system " + std::to_string(r) + " has no particles"; return FM_ERR_ARG; }
        if (!(S.Lx > 0) || !(S.Ly > 0)) { err = "box must be positive"; return FM_ERR_ARG; }
        if (S.model != E->model_) { err = "all replicas of a batch must use one material model"; return FM_ERR_ARG; }
        if (S.ntypes < 1 || S.ntypes > kMaxPairTypes) {
            err = "ntypes must be in [1, 16]";
            return FM_ERR_ARG;
        }
        for (int32_t t : S.type)
            if (t < 0 || t >= S.ntypes) { err = "particle type out of range of the pair table"; return FM_ERR_ARG; }
        float lo = S.sigma[0], hi = S.sigma[0];
        for (float s : S.sigma) { lo = std::min(lo, s); hi = std::max(hi, s); }
        if (!(lo > 0)) { err = "sigma must be positive"; return FM_ERR_ARG; }
        smin[r] = lo;
        smax[r] = hi;
        if (lo != hi) E->sigma_ = true;
        if (S.ntypes > 1) E->types_ = true;
        double rcmax = 0;
        for (float v : S.rc) rcmax = std::max(rcmax, (double)v);
        const double rc_eff = rcmax * hi + E->skin_;
        rlmax = std::max(rlmax, rc_eff);
        // cell side >= the largest cutoff (plus skin) plus the FP32 slop of
        // the key
        const double cmin = rc_eff * (1.0 + 1e-5) + 1e-6 * std::max(S.Lx, S.Ly);
        RepDesc &D = rd[r];
        std::memset(&D, 0, sizeof(D));
        D.Lx = (float)S.Lx;
        D.Ly = (float)S.Ly;
        D.hLx = 0.5f * D.Lx;
        D.hLy = 0.5f * D.Ly;
        D.ncx = (int32_t)std::floor(S.Lx / cmin);
        D.ncy = (int32_t)std::floor(S.Ly / cmin);
        if (D.ncx < 3 || D.ncy < 3) {
            err = "box too small: fewer than 3 cells of side >= rc*sigma + skin per axis";
            return FM_ERR_GEOMETRY;
        }
        D.fncx = (float)D.ncx;
        D.fncy = (float)D.ncy;
        D.dt = (float)S.dt;
        D.hdt = 0.5f * D.dt;  // femmd2D/src/integrator.cpp:102 hdt = 0.5*dt
        D.ntypes = S.ntypes;
        const float s0 = hi;
        const float rcf = S.rc[0];
        D.rc2_uniform = rcf * rcf * s0 * s0;  // reference order ((rc*rc)*s)*s
        D.s2_uniform = s0 * s0;
        D.s_uniform = s0;
        D.skin = E->skin_;
        {
            // 4 ulp of the box: one rounding of x + dt v and one of the wrap
            // per coordinate, both coordinates
            const float Lm = std::max(D.Lx, D.Ly);
            D.dslack = 4.0f * (std::nextafter(Lm, INFINITY) - Lm);
        }
        // lists cover the largest cutoff of the table (a typed pair with a
        // smaller one is filtered by the pass)
        float rcl = 0.0f;
        for (int k = 0; k < S.ntypes * S.ntypes; ++k) rcl = std::max(rcl, S.rc[k]);
        D.rc_list = rcl;
        const double rl = (double)rcl * s0 + E->skin_;
        D.rl2_uniform = (float)(rl * rl * 1.00002);
        amax = std::max(amax, (double)(D.Lx / D.ncx) * (D.Ly / D.ncy));
        rho_cp = std::max(rho_cp, 1.16 / ((double)lo * lo));
    }
    // per-cell list capacity: twice the close-packed count of a cell
    E->rho_cp_ = rho_cp;
    E->cap_ = std::min(kMaxCellCap, std::max(32, (int)std::ceil((2.0 * rho_cp * amax + 4.0) / 8.0) * 8));
    // neighbour-list capacity: 1.3 x the close-packed count of the list disc
    if (E->list_) {
        const double m = 1.3 * M_PI * rlmax * rlmax * rho_cp;
        E->nbr_cap_ = std::min(512, std::max(32, (int)std::ceil(m / 8.0) * 8));
    }

    // ---- pass 2: layout ----
    std::vector<float> hx, hy, hvx, hvy, hm, him, hs;
    std::vector<int32_t> hg, ht, huid;
    std::vector<PairEntry> hpt;
    std::vector<int32_t> ht0, ht1, ht2, hbi, hbj;
    std::vector<float4> hdmbm;
    std::vector<float2> halm;
    std::vector<float> heE;
    std::vector<float> hbk, hbr0;
    std::vector<Span> ks, es, bs;
    std::vector<int32_t> erange(2 * R), brange(2 * R);
    E->iid_h_.resize(R);
    E->lj_on_.resize(R);
    for (int r = 0; r < R; ++r) E->lj_on_[r] = reps[r]->lj_on ? 1 : 0;
    int p0 = 0, cell0 = 0;
    for (int r = 0; r < R; ++r) {
        const HostSystem &S = *reps[r];
        RepDesc &D = rd[r];
        const int n = (int)S.x.size();
        D.p0 = p0;
        D.p1 = p0 + n;
        D.cell0 = cell0;
        D.pt = (int32_t)hpt.size();
        for (int k = 0; k < S.ntypes * S.ntypes; ++k) {
            PairEntry pe;
            pe.e48 = 48.0f * S.eps[k];
            pe.e4 = 4.0f * S.eps[k];
            pe.shift = S.ush[k];
            pe.rc = S.rc[k];
            hpt.push_back(pe);
        }
        // wrap, key, internal order of this replica: the particles without
        // element or bond corners in cell order, then those with corners (and
        // every particle of a replica without pair forces) in cell order.
        // The force pass kicks the first group itself; the node kick takes
        // the second, [pg, p1), as one contiguous range.
        std::vector<float> xw(n), yw(n);
        std::vector<int> key(n), order(n);
        std::vector<char> gat(n, S.lj_on ? 0 : 1);
        for (int32_t u : S.tri) gat[u] = 1;
        for (int32_t u : S.bi) gat[u] = 1;
        for (int32_t u : S.bj) gat[u] = 1;
        for (int u = 0; u < n; ++u) {
            xw[u] = wrap_host(S.x[u], D.Lx);
            yw[u] = wrap_host(S.y[u], D.Ly);
            key[u] = cell_key_host(xw[u], yw[u], D);
        }
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return gat[a] != gat[b] ? gat[a] < gat[b] : key[a] < key[b];
        });
        D.pg = p0 + (int32_t)std::count(gat.begin(), gat.end(), (char)0);
        std::vector<int32_t> &iid = E->iid_h_[r];
        iid.assign(n, 0);
        for (int k = 0; k < n; ++k) iid[order[k]] = p0 + k;
        for (int k = 0; k < n; ++k) {
            const int u = order[k];
            hx.push_back(xw[u]);
            hy.push_back(yw[u]);
            hvx.push_back(S.vx[u]);
            hvy.push_back(S.vy[u]);
            const float m = S.mass[u];
            const bool pinned = !(m > 0.0f) || !std::isfinite(m);
            hm.push_back(std::isfinite(m) ? m : 0.0f);
            him.push_back(pinned ? 0.0f : 1.0f / m);
            hs.push_back(S.sigma[u]);
            hg.push_back(S.excl[u]);
            ht.push_back(S.type[u]);
            huid.push_back(u);
        }
        {  // one inverse mass for the whole replica (none pinned)?
            const float f = n > 0 ? him[p0] : 0.0f;
            bool uni = f > 0.0f;
            for (int k = 1; uni && k < n; ++k) uni = him[p0 + k] == f;
            D.im_uniform = uni ? f : 0.0f;
        }
        D.ks0 = (int32_t)ks.size();
        for (int b = p0; b < p0 + n; b += kKickThreads)
            ks.push_back(Span{r, b, std::min(b + kKickThreads, p0 + n), 0});
        D.nks = (int32_t)ks.size() - D.ks0;
        // elements, ordered by their lowest internal node for locality
        const int nt = (int)S.tri.size() / 3;
        std::vector<int> eorder(nt);
        std::iota(eorder.begin(), eorder.end(), 0);
        auto lowest = [&](int e) {
            return std::min(iid[S.tri[3 * e]], std::min(iid[S.tri[3 * e + 1]], iid[S.tri[3 * e + 2]]));
        };
        std::stable_sort(eorder.begin(), eorder.end(), [&](int a, int b) { return lowest(a) < lowest(b); });
        const int e_begin = (int)ht0.size();
        for (int q = 0; q < nt; ++q) {
            const int e = eorder[q];
            const int u0 = S.tri[3 * e], u1 = S.tri[3 * e + 1], u2 = S.tri[3 * e + 2];
            ht0.push_back(iid[u0]);
            ht1.push_back(iid[u1]);
            ht2.push_back(iid[u2]);
            // femmd2D/src/interactions.cpp:589-602: Dm = [X0-X2, X1-X2],
            // vol0 = det(Dm)/2 (signed), Bm = Dm^-1; evaluated in FP64
            const double D00 = (double)S.rx[u0] - S.rx[u2], D10 = (double)S.ry[u0] - S.ry[u2];
            const double D01 = (double)S.rx[u1] - S.rx[u2], D11 = (double)S.ry[u1] - S.ry[u2];
            const double det = D00 * D11 - D01 * D10;
            if (!(std::fabs(det) > 0)) { err = "degenerate element (zero reference area)"; return FM_ERR_ARG; }
            const double A0 = det / 2;
            const double B00 = D11 / det, B10 = -D10 / det, B01 = -D01 / det, B11 = D00 / det;
            hdmbm.push_back(make_float4((float)D00, (float)D01, (float)D10, (float)D11));
            hdmbm.push_back(make_float4((float)B00, (float)B01, (float)B10, (float)B11));
            const double Ee = S.tE[e] >= 0 ? S.tE[e] : S.E;
            const double nue = S.tnu[e] >= 0 ? S.tnu[e] : S.nu;
            E->ecaller_h_.push_back(e);
            heE.push_back((float)Ee);
            // femmd2D/src/interactions.cpp:44-45
            double lam = Ee * nue / ((1 - 2 * nue) * (1 + nue));
            double mu = Ee / (2 * (1 + nue));
            if (S.model == FM_LEGACY_GPU) { lam *= 2; mu *= 2; }
            halm.push_back(make_float2((float)(A0 * lam), (float)(A0 * mu)));
        }
        erange[2 * r] = e_begin;
        erange[2 * r + 1] = (int)ht0.size();
        for (int b = e_begin; b < (int)ht0.size(); b += kElemThreads)
            es.push_back(Span{r, b, std::min(b + kElemThreads, (int)ht0.size()), 0});
        const int b_begin = (int)hbi.size();
        for (size_t l = 0; l < S.bi.size(); ++l) {
            hbi.push_back(iid[S.bi[l]]);
            hbj.push_back(iid[S.bj[l]]);
            hbk.push_back(S.bk[l]);
            hbr0.push_back(S.br0[l]);
        }
        brange[2 * r] = b_begin;
        brange[2 * r + 1] = (int)hbi.size();
        for (int b = b_begin; b < (int)hbi.size(); b += kElemThreads)
            bs.push_back(Span{r, b, std::min(b + kElemThreads, (int)hbi.size()), 0});
        p0 += n;
        // one extra (always empty) cell closes the replica's range of cell
        // starts, so a replica's counts never depend on its neighbour's
        // starts (list-mode rebuilds rescan only the rebuilding replicas)
        cell0 += D.ncx * D.ncy + 1;
    }
    E->N_ = p0;
    E->ncells_ = cell0;
    // the list pass indexes list groups with 32-bit element indices: larger
    // systems walk the cells every step (cells of side >= rc*sigma + skin)
    if (E->list_ && (double)E->nbr_cap_ / 4 * p0 >= 4294967295.0) E->list_ = false;

    E->ntri_ = (int)ht0.size();
    E->nbond_ = (int)hbi.size();
    E->gather_ = (E->ntri_ + E->nbond_) > 0;
    E->reps_h_ = rd;
    E->uid_h_ = huid;
    E->kspans_h_ = ks;
    E->espans_h_ = es;
    E->bspans_h_ = bs;

    // gather CSR: per particle, its corner slots in ascending slot order
    std::vector<int32_t> goff, gent, gpos;
    if (E->gather_) {
        const int N = E->N_;
        std::vector<int32_t> cnt(N + 1, 0);
        for (int e = 0; e < E->ntri_; ++e) { cnt[ht0[e]]++; cnt[ht1[e]]++; cnt[ht2[e]]++; }
        for (int l = 0; l < E->nbond_; ++l) { cnt[hbi[l]]++; cnt[hbj[l]]++; }
        goff.assign(N + 1, 0);
        for (int i = 0; i < N; ++i) goff[i + 1] = goff[i] + cnt[i];
        gent.assign(goff[N], 0);
        std::vector<int32_t> fill(goff.begin(), goff.end() - 1);
        for (int e = 0; e < E->ntri_; ++e) {
            gent[fill[ht0[e]]++] = 3 * e;
            gent[fill[ht1[e]]++] = 3 * e + 1;
            gent[fill[ht2[e]]++] = 3 * e + 2;
        }
        const int b0 = 3 * E->ntri_;
        for (int l = 0; l < E->nbond_; ++l) {
            gent[fill[hbi[l]]++] = b0 + 2 * l;
            gent[fill[hbj[l]]++] = b0 + 2 * l + 1;
        }
        // the force pass writes each corner straight to its gather entry
        gpos.assign(gent.size(), 0);
        for (size_t e = 0; e < gent.size(); ++e) gpos[gent[e]] = (int32_t)e;
    }

    // ---- device allocations and uploads ----
    // The copies are asynchronous: their host sources (the vectors above and
    // these) stay alive until the stream has drained, on every return path
    // (the guard, declared after them, is destroyed first).
    std::vector<int2> hbrep;
    std::vector<Span> hns;
    std::vector<int32_t> hnr;
    struct DrainGuard {
        hipStream_t s;
        ~DrainGuard() { (void)hipStreamSynchronize(s); }
    } drain_guard{st};
    const int N = E->N_;
    hipError_t he = hipSuccess;
#define FM_UP(buf, vec) \
    if ((he = upload(E->buf, vec, st)) != hipSuccess) return E->check(he, "upload " #buf, err)
    FM_UP(reps_, rd);
    FM_UP(ptab_, hpt);
    FM_UP(x_, hx);
    FM_UP(y_, hy);
    FM_UP(vx_, hvx);
    FM_UP(vy_, hvy);
    FM_UP(mass_, hm);
    FM_UP(inv_m_, him);
    FM_UP(sig_, hs);
    FM_UP(grp_, hg);
    FM_UP(typ_, ht);
    FM_UP(uid_, huid);
    FM_UP(kspans_, ks);
    FM_UP(espans_, es);
    FM_UP(bspans_, bs);
    FM_UP(t0_, ht0);
    FM_UP(t1_, ht1);
    FM_UP(t2_, ht2);
    FM_UP(dmbm_, hdmbm);
    FM_UP(alm_, halm);
    FM_UP(eE_, heE);
    FM_UP(bi_, hbi);
    FM_UP(bj_, hbj);
    FM_UP(bk_, hbk);
    FM_UP(br0_, hbr0);
    FM_UP(goff_, goff);
    FM_UP(gent_, gent);
    FM_UP(gpos_, gpos);
    {
        // movable particles: no element, bond or cluster refers to them, so
        // a re-ordering may give them new internal ids
        std::vector<uint8_t> hmov((size_t)N, 1);
        for (const std::vector<int32_t> *v : {&ht0, &ht1, &ht2, &hbi, &hbj})
            for (int32_t q : *v) hmov[q] = 0;
        for (int r = 0; r < R; ++r)
            for (size_t c = 0; c < reps[r]->cl_first.size(); ++c)
                for (int32_t k = 0; k < reps[r]->cl_n[c]; ++k)
                    hmov[E->iid_h_[r][reps[r]->cl_first[c] + k]] = 0;
        std::vector<int32_t> hmovids;
        for (int q = 0; q < N; ++q)
            if (hmov[q]) hmovids.push_back(q);
        E->nmov_ = (int)hmovids.size();
        if (hmovids.empty()) hmovids.push_back(0);
        FM_UP(mov_, hmov);
        FM_UP(movids_, hmovids);
    }
#undef FM_UP
#define FM_AL(buf, cnt) \
    if ((he = E->buf.alloc(cnt)) != hipSuccess) return E->check(he, "alloc " #buf, err)
    FM_AL(fx_, N);
    FM_AL(fy_, N);
    FM_AL(outx_, N);
    FM_AL(outy_, N);
    FM_AL(count_, (size_t)E->ncells_);
    FM_AL(tmp_, (size_t)E->ncells_ * E->cap_);
    E->nsort_ = (E->ncells_ + kSortCells - 1) / kSortCells;
    FM_AL(sort_status_, std::max(E->nsort_, 1));
    hbrep.assign(std::max(E->nsort_, 1), make_int2(0, 0));
    {
        // per sort block: the replicas its cells belong to (replica r owns
        // cells [cell0, cell0 + ncx * ncy + 1))
        int r = 0;
        for (int b = 0; b < E->nsort_; ++b) {
            const int c0 = b * kSortCells, c1 = std::min(c0 + kSortCells, E->ncells_) - 1;
            while (r + 1 < R && rd[r + 1].cell0 <= c0) ++r;
            int r1 = r;
            while (r1 + 1 < R && rd[r1 + 1].cell0 <= c1) ++r1;
            hbrep[b] = make_int2(r, r1);
        }
        if ((he = upload(E->brep_, hbrep, st)) != hipSuccess) return E->check(he, "upload brep", err);
    }
    if (E->list_ && R > kDecideMinReps) {
        FM_AL(fspan_, std::max<size_t>(ks.size(), 1));
        FM_AL(rinfo_, R);
        FM_AL(mark_, std::max(E->nsort_, 1));
        if ((he = hipMemsetAsync(E->mark_.p, 0, E->mark_.n, st)) != hipSuccess)
            return E->check(he, "memset", err);
        FM_AL(nf_, 5);
        if ((he = hipMemsetAsync(E->nf_.p, 0, 5 * sizeof(int32_t), st)) != hipSuccess)
            return E->check(he, "memset", err);
        FM_AL(fblock_, std::max(E->nsort_, 1));
        FM_AL(sort_status2_, std::max(E->nsort_, 1));
        if ((he = hipMemsetAsync(E->sort_status2_.p, 0, E->sort_status2_.n * sizeof(unsigned long long),
                                 st)) != hipSuccess)
            return E->check(he, "memset", err);
    }
    FM_AL(sort_ticket_, 1);
    FM_AL(start_, E->ncells_ + 1);
    FM_AL(perm_, N);
    FM_AL(sorted_, N);
    FM_AL(sorted_typ_, N);
    FM_AL(cf_, std::max(1, 3 * E->ntri_ + 2 * E->nbond_));
    FM_AL(ke_part_, 2 * std::max<size_t>(ks.size(), 1));
    {
        // the node kick: per replica its particles [pg, p1) (those with
        // element or bond corners, and every particle of a replica without
        // pair forces: no tile kicks them), spans of kKickThreads
        hnr.assign(2 * R, 0);
        E->nnode_ = 0;
        for (int r = 0; r < R; ++r) {
            hnr[2 * r] = (int)hns.size();
            for (int b = rd[r].pg; b < rd[r].p1; b += kKickThreads)
                hns.push_back(Span{r, b, std::min(b + kKickThreads, rd[r].p1), 0});
            hnr[2 * r + 1] = (int)hns.size();
            E->nnode_ += rd[r].p1 - rd[r].pg;
        }
        E->nspans_h_ = hns;
        if ((he = upload(E->nspans_, hns, st)) != hipSuccess ||
            (he = upload(E->nrng_, hnr, st)) != hipSuccess)
            return E->check(he, "upload node kick", err);
        FM_AL(ke_node_, 2 * std::max<size_t>(hns.size(), 1));
    }
    FM_AL(fea_part_, 2 * std::max<size_t>(es.size(), 1));
    FM_AL(bond_part_, 2 * std::max<size_t>(bs.size(), 1));
    FM_AL(obs_chunk_, (size_t)R * kObsChunks * kObsRaw);
    FM_AL(obs_, (size_t)R * kObsStride);
    FM_AL(err_, 1);
    if (E->list_) {
        FM_AL(nbr_, (size_t)E->nbr_cap_ * N / 2);  // uint16 entries
        FM_AL(nnum_, N);
        FM_AL(ctl_, R);
        FM_AL(any_, 1);
        // two sets of displacement slots and two sorted-order position
        // buffers: a drift writes the set its force pass does not read
        FM_AL(dmax_, (size_t)2 * R * kDispSlots * kDispStride);
        FM_AL(spos_, N);
        FM_AL(xys_, (size_t)2 * N);
        if (E->sigma_) FM_AL(sigs_, N);
    }
#undef FM_AL
    if ((he = hipMemsetAsync(E->count_.p, 0, sizeof(int32_t) * E->ncells_, st)) != hipSuccess)
        return E->check(he, "memset", err);
    // status words of launch 0 must not read as valid: state 0
    if ((he = hipMemsetAsync(E->sort_status_.p, 0, E->sort_status_.n * sizeof(unsigned long long),
                             st)) != hipSuccess ||
        (he = hipMemsetAsync(E->sort_ticket_.p, 0, sizeof(unsigned long long), st)) != hipSuccess)
        return E->check(he, "memset", err);
    // every sorted slot holds a valid particle index even where a cell
    // overflow left it unwritten (the list pass reads perm over whole ranges)
    if ((he = hipMemsetAsync(E->perm_.p, 0, sizeof(int32_t) * N, st)) != hipSuccess)
        return E->check(he, "memset", err);
    if (E->list_) {
        if ((he = hipMemsetAsync(E->nnum_.p, 0, sizeof(int32_t) * N, st)) != hipSuccess ||
            (he = hipMemsetAsync(E->ctl_.p, 0, sizeof(Ctl) * R, st)) != hipSuccess ||
            (he = hipMemsetAsync(E->any_.p, 0, sizeof(int32_t), st)) != hipSuccess ||
            (he = hipMemsetAsync(E->spos_.p, 0, sizeof(int32_t) * N, st)) != hipSuccess ||
            (he = hipMemsetAsync(E->dmax_.p, 0, E->dmax_.n * sizeof(uint32_t), st)) != hipSuccess)
            return E->check(he, "memset", err);
    }
    if ((he = hipMemsetAsync(E->err_.p, 0, sizeof(uint32_t), st)) != hipSuccess)
        return E->check(he, "memset", err);
    if ((he = hipMemsetAsync(E->fx_.p, 0, sizeof(float) * N, st)) != hipSuccess)
        return E->check(he, "memset", err);
    if ((he = hipMemsetAsync(E->fy_.p, 0, sizeof(float) * N, st)) != hipSuccess)
        return E->check(he, "memset", err);
    if ((he = hipMemsetAsync(E->ke_part_.p, 0, E->ke_part_.n * sizeof(double), st)) != hipSuccess)
        return E->check(he, "memset", err);
    if ((he = hipMemsetAsync(E->fea_part_.p, 0, E->fea_part_.n * sizeof(double), st)) != hipSuccess)
        return E->check(he, "memset", err);
    if ((he = hipMemsetAsync(E->bond_part_.p, 0, E->bond_part_.n * sizeof(double), st)) != hipSuccess)
        return E->check(he, "memset", err);
    if ((he = hipMemsetAsync(E->obs_chunk_.p, 0, E->obs_chunk_.n * sizeof(double), st)) != hipSuccess ||
        (he = hipMemsetAsync(E->obs_.p, 0, E->obs_.n * sizeof(double), st)) != hipSuccess)
        return E->check(he, "memset", err);

    rc = E->build_clusters(reps, err);
    if (rc) return rc;
    // tiles (and LDS capacity) for the requested widths
    std::vector<int> tw(R);
    for (int r = 0; r < R; ++r) tw[r] = reps[r]->tile_w;
    rc = E->build_tiles(tw, err);
    if (rc) return rc;
    rc = E->check(hipStreamSynchronize(st), "commit", err);
    if (rc) return rc;
    out = std::move(E);
    return FM_OK;
}

// Tiles of one cell row each: (replica, row, first column, width).  Width 0
// picks one from the mean cell occupancy so a tile's centre cells hold about
// 0.9 x block particles.
int Engine::build_tiles(const std::vector<int> &tile_w, std::string &err) {
    drop_graph();
    tiles_h_.clear();
    tile_w_.assign(R_, 0);
    rng_h_.assign(8 * R_, 0);
    int kb = 0, eb = 0, bb = 0;
    double worst = 0.0;
    for (int r = 0; r < R_; ++r) {
        const RepDesc &D = reps_h_[r];
        const double ppc = (double)(D.p1 - D.p0) / (D.ncx * D.ncy);
        int W = (r < (int)tile_w.size()) ? tile_w[r] : 0;
        if (W <= 0) W = (int)std::lround(0.9 * block_ / std::max(ppc, 0.5));
        W = std::max(1, std::min(W, std::min(kMaxTileW, (int)D.ncx)));
        tile_w_[r] = W;
        const int t_begin = (int)tiles_h_.size();
        // rows 0 and ncy - 1 first (their halo rows cross the periodic
        // boundary, so all their lists take the minimum image): the force
        // pass hands each XCD a contiguous band of tiles in order, and the
        // costlier tiles then start early instead of closing the last band
        for (int iy = 0; iy < (lj_on_[r] ? D.ncy : 0); ++iy) {
            const int cy = iy == 0 ? 0 : (iy == 1 ? D.ncy - 1 : iy - 1);
            for (int cx0 = 0; cx0 < D.ncx; cx0 += W) {
                const int w = std::min(W, D.ncx - cx0);
                // interior tiles (no staged cell wraps, >= 5 cells per axis)
                // never need the minimum image: |dx| < 2 cells <= 0.4 L
                const bool big = D.ncx >= 5 && D.ncy >= 5;
                const bool in_x = big && cx0 >= 1 && cx0 + w <= D.ncx - 1;
                const bool in_y = big && cy >= 1 && cy <= D.ncy - 2;
                // bit 16: some list may need the minimum image; bits 17, 18:
                // in x, in y (a particle that wraps between list builds is
                // then more than rc from every centre in an interior row or
                // column, so only the edge rows and columns need an axis)
                tiles_h_.push_back(Tile{r, cy, cx0,
                                        w | (in_x && in_y ? 0 : (1 << 16)) | (in_x ? 0 : (1 << 17)) |
                                            (in_y ? 0 : (1 << 18))});
            }
        }
        // staged particles of a tile at close packing (1.16 / sigma_min^2)
        const double acell = (double)(D.Lx / D.ncx) * (D.Ly / D.ncy);
        worst = std::max(worst, 3.0 * (W + 2) * acell * rho_cp_);
        int32_t *g = &rng_h_[8 * r];
        // kick spans
        g[0] = kb;
        while (kb < (int)kspans_h_.size() && kspans_h_[kb].rep == r) ++kb;
        g[1] = kb;
        g[2] = t_begin;
        g[3] = (int)tiles_h_.size();
        g[4] = eb;
        while (eb < (int)espans_h_.size() && espans_h_[eb].rep == r) ++eb;
        g[5] = eb;
        g[6] = bb;
        while (bb < (int)bspans_h_.size() && bspans_h_[bb].rep == r) ++bb;
        g[7] = bb;
    }
    // LDS capacity: 1.3 x the close-packed count of the staged cells (an
    // elastomer under compression can exceed close packing a little), rounded
    // up to a multiple of 128
    int cap = ((int)std::ceil(1.3 * worst) + 127) / 128 * 128;
    const int hard = (types_ ? 2560 : 3072);  // dynamic LDS stays <= 64 KB
    lds_cap_ = std::max(256, std::min(cap, hard));
    hipError_t he;
    if (list_) lb_grid_ = list_build_resident(sigma_, block_, force_args(), device_);
    if ((he = upload(tiles_, tiles_h_, stream_)) != hipSuccess) return check(he, "tiles", err);
    if ((he = upload(rng_, rng_h_, stream_)) != hipSuccess) return check(he, "ranges", err);
    if (list_ && R_ > kDecideMinReps &&
        (he = ftile_.alloc(std::max<size_t>(tiles_h_.size(), 1))) != hipSuccess)
        return check(he, "tile list", err);
    if ((he = lj_part_.alloc(2 * std::max<size_t>(tiles_h_.size(), 1))) != hipSuccess ||
        (he = ke_tile_.alloc(2 * std::max<size_t>(tiles_h_.size(), 1))) != hipSuccess)
        return check(he, "tile partials", err);
    if ((he = hipMemsetAsync(lj_part_.p, 0, lj_part_.n * sizeof(double), stream_)) != hipSuccess ||
        (he = hipMemsetAsync(ke_tile_.p, 0, ke_tile_.n * sizeof(double), stream_)) != hipSuccess)
        return check(he, "memset", err);
    primed_ = false;
    return FM_OK;
}

int Engine::set_dt(int r, double dt, std::string &err) {
    if (r < 0 || r >= R_ || !(dt > 0)) { err = "bad replica or dt"; return FM_ERR_ARG; }
    adj_.valid = false;
    reps_h_[r].dt = (float)dt;
    reps_h_[r].hdt = 0.5f * reps_h_[r].dt;
    int rc = check(hipMemcpyAsync(reps_.p + r, &reps_h_[r], sizeof(RepDesc), hipMemcpyHostToDevice,
                                  stream_), "set_dt", err);
    if (rc) return rc;
    return check(hipStreamSynchronize(stream_), "set_dt", err);
}

int Engine::retile(const std::vector<int> &tile_w, int block, std::string &err) {
    if (block > 0) {
        if (block != 64 && block != 128 && block != 192 && block != 256) {
            err = "block must be 64, 128, 192 or 256";
            return FM_ERR_ARG;
        }
        block_ = block;
    }
    int rc = check(hipStreamSynchronize(stream_), "sync", err);
    if (rc) return rc;
    return build_tiles(tile_w, err);
}

int Engine::set_tuning(int tile_w, int block, int use_graph, std::string &err) {
    if (use_graph >= 0) {
        use_graph_ = use_graph;
        drop_graph();
    }
    if (tile_w > 0 || block > 0) {
        std::vector<int> tw(R_, tile_w > 0 ? tile_w : 0);
        if (tile_w <= 0) tw = tile_w_;
        if (block > 0 && tile_w <= 0) tw.assign(R_, 0);  // re-derive from the new block
        return retile(tw, block, err);
    }
    return FM_OK;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
// Binning of the positions the force pass reads (xcur); gated = list mode,
// only the replicas whose rebuild flag is set.
int Engine::launch_bin(bool gated) {
    if (!gated && bins_pending_) {
        // a drift's binning no sort consumed (a run broken off after a body):
        // start from empty cells
        if (launch_fill_int(count_.p, 0, (size_t)ncells_, stream_) != hipSuccess) return -1;
        bins_pending_ = false;
    }
    KickArgs a{};
    a.reps = reps_.p;
    a.spans = kspans_.p;
    a.x = xcur();
    a.y = ycur();
    a.count = count_.p;
    a.tmp = tmp_.p;
    a.cap = cap_;
    a.err = err_.p;
    a.nspans = (int)kspans_h_.size();
    a.obs = obs_args();
    a.obs_l1 = (gated && l1_) ? 1 : 0;
    if (gated) {  // list mode: decide and bin only the replicas that rebuild
        a.ctl = ctl_.p;
        a.any = any_.p;
        a.dmax = dmax_cur();  // the drift that produced the current positions
        a.one_rep = R_ == 1;
        a.rblog = R_ == 1 ? rblog_ : nullptr;
        if (R_ > kDecideMinReps) {
            DecideArgs d{};
            d.reps = reps_.p;
            d.ctl = ctl_.p;
            d.dmax = dmax_cur();
            d.rng = rng_.p;
            d.nrep = R_;
            d.any = any_.p;
            d.fspan = fspan_.p;
            d.ftile = ftile_.p;
            d.fblock = fblock_.p;
            d.nsort = nsort_;
            d.nf = nf_.p;
            d.rinfo = rinfo_.p;
            d.mark = mark_.p;
            if (launch_decide(d, stream_) != hipSuccess) return -1;
            a.fspan = fspan_.p;
            a.nfs = nf_.p;
        }
    }
    return launch_kick(0, false, false, (int)kspans_h_.size(), a, stream_) == hipSuccess ? 0 : -1;
}

// The kick arguments shared by the first half-kick of a run (every
// particle) and the node kick (the particles with element or bond corners).
KickArgs Engine::kick_args(bool drift) {
    KickArgs k{};
    k.reps = reps_.p;
    k.spans = kspans_.p;
    k.x = x_.p;
    k.y = y_.p;
    k.vx = vx_.p;
    k.vy = vy_.p;
    k.fx = fx_.p;
    k.fy = fy_.p;
    k.cf = cf_.p;
    k.goff = gather_ ? goff_.p : nullptr;
    k.mass = mass_.p;
    k.inv_m = inv_m_.p;
    k.count = count_.p;
    k.tmp = tmp_.p;
    k.cap = cap_;
    k.ke_part = ke_part_.p;
    k.err = err_.p;
    k.rec = drift ? rec_ : nullptr;
    k.n = N_;
    if (list_) {
        // the drift writes the positions of the next buffer of the pair
        k.dmax = dmax_nxt();
        k.spos = spos_.p;
        k.xys = xys_nxt();
    }
    return k;
}

// One piece of a run (femmd2D/src/integrator.cpp:8-29 per step):
//   open  (drift, !kick2): first half-kick, drift and wrap of every particle
//         from the stored forces (k_kick);
//   body  (drift, kick2):  the neighbour update of the current positions
//         (list mode: binning, counting sort and list build gated on the
//         rebuild decision; cell walk: the sort of the last binning), then
//         the force pass, which also gives every particle without element or
//         bond corners its second half-kick of this step and the first
//         half-kick, drift and wrap of the next, then the node kick of the
//         others (their gathered corner forces added);
//   close (!drift, kick2): as a body with the second half-kick only; the
//         forces are stored for the next run's first half-kick.
// In list mode the drift writes the other buffer of the sorted-order
// positions and displacement slots (par_ flips).
int Engine::launch_step_body(bool kick2, bool drift, bool energy, bool l1, bool l2) {
    struct Restore {  // the flags of this body, restored on every return
        Engine *e;
        bool en, a, b;
        ~Restore() { e->energy_ = en; e->l1_ = a; e->l2_ = b; }
    } restore{this, energy_, l1_, l2_};
    energy_ = energy;
    l1_ = l1;
    l2_ = l2;
    if (!kick2) {
        if (!drift) return 0;
        const KickArgs k = kick_args(true);
        const hipError_t e = launch_kick(1, gather_, false, (int)kspans_h_.size(), k, stream_);
        if (e != hipSuccess) return -1;
        if (list_) par_ ^= 1;
        else bins_pending_ = true;
        return 0;
    }
    if (list_) {
        // binning, sort and list build run only for the replicas where some
        // particle moved more than skin / 2 since the last build
        time_next(1);
        const int bad = launch_bin(true);
        arm_kernel_timer(nullptr, nullptr);
        if (bad) return -1;
        if (launch_sort(true)) return -1;
        if (launch_list_build(true)) return -1;
    } else {
        if (launch_sort()) return -1;  // the binning of the last drift
    }
    {
        const ForceArgs f = force_args();
        time_next(4);
        const hipError_t e =
            launch_forces(sigma_, types_, false, list_, energy_, drift ? 2 : 1, block_, f, stream_);
        arm_kernel_timer(nullptr, nullptr);
        if (e != hipSuccess) return -1;
        ene_valid_ = energy_ || !drift;
    }
    if (nnode_ > 0) {
        KickArgs k = kick_args(drift);
        k.spans = nspans_.p;
        k.ke_part = ke_node_.p;
        // (a throughput-bound node kick reads its corner forces coalesced)
        k.gather_lds = nspans_h_.size() >= kGatherLdsSpans ? 1 : 0;
        time_next(0);
        const hipError_t e =
            launch_kick(drift ? 3 : 2, gather_, energy_ || !drift, (int)nspans_h_.size(), k, stream_);
        arm_kernel_timer(nullptr, nullptr);
        if (e != hipSuccess) return -1;
    }
    if (drift && list_) par_ ^= 1;
    if (drift && !list_) bins_pending_ = true;  // the next body's sort consumes it
    return 0;
}

// Counting sort of the last binning (single pass: scan and scatter).
int Engine::launch_sort(bool gated) {
    ScanArgs s{};
    s.reps = reps_.p;
    s.nrep = R_;
    s.ncells = ncells_;
    s.cap = cap_;
    s.nsort = nsort_;
    s.brep = brep_.p;
    s.count = count_.p;
    s.tmp = tmp_.p;
    s.status = sort_status_.p;
    s.ticket = sort_ticket_.p;
    s.start = start_.p;
    s.perm = perm_.p;
    s.uid = uid_.p;
    s.x = xcur();
    s.y = ycur();
    s.grp = grp_.p;
    s.sig = sigma_ ? sig_.p : nullptr;
    s.typ = types_ ? typ_.p : nullptr;
    s.sorted = sorted_.p;
    s.sorted_typ = types_ ? sorted_typ_.p : nullptr;
    s.err = err_.p;
    s.obs = obs_args();
    s.obs_l1 = (!list_ && l1_) ? 1 : 0;  // the cell walk's sort forms level 1
    if (list_) {
        s.spos = spos_.p;
        s.xys = xys_cur();
        s.sigs = sigma_ ? sigs_.p : nullptr;
    }
    if (gated) {
        s.ctl = ctl_.p;
        s.any = any_.p;
        if (R_ > kDecideMinReps) {  // the blocks listed by k_decide
            s.fblock = fblock_.p;
            s.nf = nf_.p;
            s.status2 = sort_status2_.p;
        }
    }
    time_next(2);
    const hipError_t e = ::fm::launch_sort(s, stream_);
    arm_kernel_timer(nullptr, nullptr);
    if (!gated) bins_pending_ = false;
    return e == hipSuccess ? 0 : -1;
}

// Neighbour lists of the replicas flagged for a rebuild (gated), or of all.
int Engine::launch_list_build(bool gated) {
    ForceArgs f = force_args();
    if (gated) f.ctl = ctl_.p;
    if (gated && R_ > kDecideMinReps) {  // the tiles listed by k_decide
        f.ftile = ftile_.p;
        f.nft = nf_.p + 1;
    }
    time_next(3);
    const hipError_t e = ::fm::launch_list_build(sigma_, block_, f, stream_);
    arm_kernel_timer(nullptr, nullptr);
    return e == hipSuccess ? 0 : -1;
}

int Engine::launch_pair_pass() {
    const ForceArgs f = force_args();
    time_next(4);
    const hipError_t e = launch_forces(sigma_, types_, false, true, energy_, 0, block_, f, stream_);
    arm_kernel_timer(nullptr, nullptr);
    ene_valid_ = energy_;
    return e == hipSuccess ? 0 : -1;
}

// Cell-walk force pass over the current binning (sort first).
int Engine::launch_forces_only(bool debug) {
    if (launch_sort()) return -1;  // (the binning is in the kick)
    const ForceArgs f = force_args();
    time_next(4);
    const hipError_t e =
        launch_forces(sigma_, types_, debug, false, energy_ || debug, 0, block_, f, stream_);
    arm_kernel_timer(nullptr, nullptr);
    ene_valid_ = energy_ || debug;
    return e == hipSuccess ? 0 : -1;
}

ForceArgs Engine::force_args() {
    ForceArgs f{};
    f.reps = reps_.p;
    f.ptab = ptab_.p;
    f.tiles = tiles_.p;
    f.ntiles = (int)tiles_h_.size();
    f.lds_cap = lds_cap_;
    f.start = start_.p;
    f.perm = perm_.p;
    f.sorted = sorted_.p;
    f.sorted_typ = sorted_typ_.p;
    f.fx = fx_.p;
    f.fy = fy_.p;
    f.lj_part = lj_part_.p;
    f.espans = espans_.p;
    f.nespans = (int)espans_h_.size();
    f.t0 = t0_.p;
    f.t1 = t1_.p;
    f.t2 = t2_.p;
    f.dmbm = dmbm_.p;
    f.alm = alm_.p;
    f.model = (model_ == FM_NEO_HOOKEAN) ? 0 : 1;
    f.x = xcur();
    f.y = ycur();
    f.cf = cf_.p;
    f.gpos = gpos_.p;
    f.fea_part = fea_part_.p;
    f.bspans = bspans_.p;
    f.nbspans = (int)bspans_h_.size();
    f.bi = bi_.p;
    f.bj = bj_.p;
    f.bk = bk_.p;
    f.br0 = br0_.p;
    f.cf_bond0 = 3 * ntri_;
    f.bond_part = bond_part_.p;
    f.nbr_cap = nbr_cap_;
    f.nbr = nbr_.p;
    f.nnum = nnum_.p;
    // stage list entries in LDS when the tile and the buffer fit in 64 KB
    // (static LDS of the kernel: ~2 KB)
    f.lbuf = list_build_lds_bytes(lds_cap_, cap_, nbr_cap_, block_, true) <= 62 * 1024;
    f.cell_cap = cap_;
    f.gated_grid = lb_grid_;
    f.sigs = sigs_.p;
    f.any = list_ ? any_.p : nullptr;
    f.ctl_w = ctl_.p;
    f.nrep = R_;
    f.n = N_;
    f.uid = uid_.p;
    f.dbg_cnt = dbg_cnt_.p;
    f.dbg_s1 = dbg_s1_.p;
    f.dbg_s2 = dbg_s2_.p;
    f.err = err_.p;
    f.obs = obs_args();
    f.obs_l2 = l2_ ? 1 : 0;
    f.xys = list_ ? xys_cur() : nullptr;
    f.dmax = list_ ? dmax_cur() : nullptr;
    // the kick of the force pass
    f.xw = x_.p;
    f.yw = y_.p;
    f.vx = vx_.p;
    f.vy = vy_.p;
    f.inv_m = inv_m_.p;
    f.mass = mass_.p;
    f.goff = gather_ ? goff_.p : nullptr;
    f.xys_out = list_ ? xys_nxt() : nullptr;
    f.dmax_out = list_ ? dmax_nxt() : nullptr;
    f.rec = rec_;
    f.ke_tile = ke_tile_.p;
    f.count = count_.p;
    f.tmp = tmp_.p;
    return f;
}

ObsArgs Engine::obs_args() {
    ObsArgs o{};
    o.reps = reps_.p;
    o.rng = rng_.p;
    o.nrng = nrng_.p;
    o.ke_kick = 0;
    o.ke_part = ke_part_.p;
    o.ke_tile = ke_tile_.p;
    o.ke_node = ke_node_.p;
    o.lj_part = lj_part_.p;
    o.fea_part = fea_part_.p;
    o.bond_part = bond_part_.p;
    o.chunk = obs_chunk_.p;
    o.obs = obs_.p;
    o.nrep = R_;
    // a single system's partials in 16 chunks (one workgroup each); a batch
    // of many replicas one chunk per replica (each replica's partials are
    // few, and the jobs are many)
    o.nchunk = R_ > 8 ? 1 : kObsChunks;
    return o;
}

// Per-kernel device time of nsteps step bodies (no graph): HIP events on
// the engine's stream around each launch.  out[0..5] = kick (+ bin in cell
// mode), bin, sort, list build, forces, whole body, in ms per step (list
// mode: the rebuild kernels are gated, so their time is the average over
// rebuild and skip steps).
int Engine::profile(int nsteps, double *out, std::string &err) {
    if (nsteps <= 0) { err = "nsteps must be > 0"; return FM_ERR_ARG; }
    int rc = ensure_primed(err);
    if (rc) return rc;
    const int K = 2 * kProfKernels;
    std::vector<hipEvent_t> ev(K * nsteps);
    for (auto &e : ev) {
        rc = check(hipEventCreate(&e), "event", err);
        if (rc) return rc;
    }
    // the first half-kick of this batch of steps (velocities stay a
    // consistent KDK state: a final half-kick closes the loop below); the
    // bodies carry the observables as the timed steps of a run do
    const bool every = obs_every_ > 0;
    obs_valid_ = false;
    (void)every;
    if (launch_bodies_eager(kind_open())) return check(hipGetLastError(), "step", err);
    for (int s = 0; s < nsteps; ++s) {
        prof_ev_ = &ev[K * s];
        const BodyKind k = s == 0 ? kind_first() : kind_body(1);
        int bad = launch_step_body(k.kick2, k.drift, k.energy, k.l1, k.l2);
        prof_ev_ = nullptr;
        if (bad) return check(hipGetLastError(), "step", err);
    }
    if (launch_bodies_eager(kind_close(nsteps >= 1)))
        return check(hipGetLastError(), "final kick", err);
    steps_ += nsteps + 1;
    rc = sync(err);
    if (!rc) obs_valid_ = true;
    // out[k] = dispatch duration of kernel k (0 where the mode launches no
    // such kernel: its events were never recorded), out[5] = whole body,
    // from its first kernel's start (binning in list mode, else the sort) to
    // its last kernel's end (the node kick, else the force pass)
    double acc[kProfKernels + 1] = {0};
    const int k_first = list_ ? 1 : 2, k_last = nnode_ > 0 ? 0 : 4;
    for (int s = 0; s < nsteps && !rc; ++s) {
        float t;
        hipEvent_t *e = &ev[K * s];
        for (int k = 0; k < kProfKernels; ++k)
            if (hipEventElapsedTime(&t, e[2 * k], e[2 * k + 1]) == hipSuccess) acc[k] += t;
        if (hipEventElapsedTime(&t, e[2 * k_first], e[2 * k_last + 1]) == hipSuccess)
            acc[kProfKernels] += t;
    }
    (void)hipGetLastError();  // the elapsed times of unrecorded events failed
    for (auto &e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k <= kProfKernels; ++k) out[k] = acc[k] / nsteps;
    return rc;
}

int Engine::ensure_primed(std::string &err) {
    if (primed_) return FM_OK;
    if (launch_bin()) return check(hipGetLastError(), "bin", err);
    if (list_) {
        // a fresh list epoch at the current positions
        if (launch_sort()) return check(hipGetLastError(), "sort", err);
        if (launch_list_build(false)) return check(hipGetLastError(), "list build", err);
        if (launch_ctl_reset(ctl_.p, R_, dmax_.p, 2 * R_ * kDispSlots, stream_) != hipSuccess)
            return check(hipGetLastError(), "ctl", err);
        if (launch_pair_pass()) return check(hipGetLastError(), "forces", err);
    } else {
        if (launch_forces_only(false)) return check(hipGetLastError(), "forces", err);
    }
    primed_ = true;
    return FM_OK;
}

// Re-sort of the internal order (kernels in kernels.hip): the movable
// particles take the movable internal slots in the order of the last
// counting sort.  Every per-particle array is permuted; perm and spos are
// rewritten, so the sorted order, the neighbour lists and the forces stay
// valid.  Asynchronous; the host id maps are refreshed on first use.
int Engine::launch_reorder_now() {
    if (nmov_ <= 1) return 0;
    hipError_t he;
    const int nblk = (N_ + kReorderChunk - 1) / kReorderChunk;
    ReorderArgs a{};
    a.narr = 0;
    for (DBuf<float> *b : {&x_, &y_, &vx_, &vy_, &fx_, &fy_, &mass_, &inv_m_, &sig_})
        if (b->n >= (size_t)N_) a.arr[a.narr++] = reinterpret_cast<uint32_t *>(b->p);
    for (DBuf<int32_t> *b : {&grp_, &typ_, &uid_})
        if (b->n >= (size_t)N_) a.arr[a.narr++] = reinterpret_cast<uint32_t *>(b->p);
    if (ro_bsum_.n < (size_t)nblk && (he = ro_bsum_.alloc(nblk)) != hipSuccess) return -1;
    if (ro_oldof_.n < (size_t)N_) {
        if ((he = ro_oldof_.alloc(N_)) != hipSuccess) return -1;
        if ((he = hipMemsetAsync(ro_oldof_.p, 0, (size_t)N_ * sizeof(int32_t), stream_)) != hipSuccess)
            return -1;
    }
    if (ro_tmp_.n < (size_t)a.narr * N_ && (he = ro_tmp_.alloc((size_t)a.narr * N_)) != hipSuccess)
        return -1;
    a.n = N_;
    a.nblk = nblk;
    a.mov = mov_.p;
    a.movids = movids_.p;
    a.nmov = nmov_;
    a.perm = perm_.p;
    a.spos = list_ ? spos_.p : nullptr;
    a.bsum = ro_bsum_.p;
    a.oldof = ro_oldof_.p;
    a.tmp = ro_tmp_.p;
    if (launch_reorder(a, stream_) != hipSuccess) return -1;
    order_stale_ = true;
    ++nreorder_;
    if (!list_ && bins_pending_) {
        // the cell walk's pending binning (made by the last drift) holds the
        // old ids: bin the current positions again under the new ones (the
        // cells are cleared first) for the next step's sort
        if (launch_bin()) return -1;
        bins_pending_ = true;
    }
    return 0;
}

int Engine::reorder(std::string &err) {
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_primed(err);  // perm of the current positions
    if (rc) return rc;
    if (launch_reorder_now()) return check(hipGetLastError(), "reorder", err);
    since_reorder_ = 0;
    return sync(err);
}

int Engine::ensure_order(std::string &err) {
    if (!order_stale_) return FM_OK;
    int rc = check(hipMemcpyAsync(uid_h_.data(), uid_.p, (size_t)N_ * sizeof(int32_t),
                                  hipMemcpyDeviceToHost, stream_), "uid", err);
    if (rc) return rc;
    rc = check(hipStreamSynchronize(stream_), "uid", err);
    if (rc) return rc;
    for (int r = 0; r < R_; ++r) {
        std::vector<int32_t> &iid = iid_h_[r];
        for (int q = reps_h_[r].p0; q < reps_h_[r].p1; ++q) iid[uid_h_[q]] = q;
    }
    order_stale_ = false;
    return FM_OK;
}

// A run of identical bodies, or the opening / closing piece of a run: each
// distinct kind (with the buffer parity it starts from) is captured into a
// graph the first time it is launched and replayed afterwards, so that
// steady-state runs of any length launch only graphs.
int Engine::launch_bodies_eager(const BodyKind &k) {
    for (int q = 0; q < k.count; ++q)
        if (launch_step_body(k.kick2, k.drift, k.energy, k.l1, k.l2)) return -1;
    if (k.close) {
        // femmd2D/src/integrator.cpp:18-19: the observables of the last step
        // (Properties::compute)
        return launch_observe(obs_args(), true, stream_) == hipSuccess ? 0 : -1;
    }
    return 0;
}

hipGraphExec_t Engine::graph_of(const BodyKind &k) {
    const int key = k.key() | (par_ << 14);
    for (auto &g : graphs_)
        if (g.first == key) return g.second;
    // capturing records the launches of the kind from the current parity;
    // nothing runs, so the parity is restored afterwards
    const int par0 = par_;
    const bool pend0 = bins_pending_;
    if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    const int bad = launch_bodies_eager(k);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(stream_, &g);
    par_ = par0;
    bins_pending_ = pend0;
    if (bad || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return nullptr;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return nullptr;
    graphs_.emplace_back(key, ex);
    return ex;
}

int Engine::launch_bodies(const BodyKind &k) {
    if (!use_graph_ || rec_ || prof_ev_) return launch_bodies_eager(k);
    hipGraphExec_t ex = graph_of(k);
    if (!ex) return -1;
    if (hipGraphLaunch(ex, stream_) != hipSuccess) return -1;
    if (list_ && k.drift) par_ ^= (k.count & 1);  // each drift flips the parity
    if (!list_) bins_pending_ = k.drift;          // a body or the opening kick bins
    return 0;
}

// The kinds of a run: the first half-kick; the first body (its step's
// observables are formed, nothing earlier is accumulated); further bodies
// (the previous step's observables accumulated); the close.  n >= 2 closes
// with the previous step accumulated.
Engine::BodyKind Engine::kind_open() const { return BodyKind{1, false, true, false, false, false, false}; }
Engine::BodyKind Engine::kind_first() const {
    const bool every = obs_every_ > 0;
    return BodyKind{1, true, true, every, false, false, false};
}
Engine::BodyKind Engine::kind_body(int count) const {
    const bool every = obs_every_ > 0;
    return BodyKind{count, true, true, every, every, every, false};
}
Engine::BodyKind Engine::kind_close(bool after_body) const {
    const bool acc = obs_every_ > 0 && after_body;
    return BodyKind{1, true, false, true, acc, acc, true};
}

int Engine::prepare_graphs() {
    if (!use_graph_) return 0;
    const int par0 = par_;
    const BodyKind kinds[] = {kind_open(), kind_first(), kind_body(kGraphSteps), kind_body(1),
                              kind_close(true), kind_close(false)};
    for (int p = 0; p < (list_ ? 2 : 1); ++p) {
        par_ = p;
        for (const BodyKind &k : kinds)
            if (!graph_of(k)) {
                par_ = par0;
                return -1;
            }
    }
    par_ = par0;
    return 0;
}

int Engine::step(int nsteps, std::string &err) {
    if (nsteps < 0) { err = "nsteps must be >= 0"; return FM_ERR_ARG; }
    if (nsteps == 0) return FM_OK;
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_primed(err);
    if (rc) return rc;
    if (graphs_.empty() && prepare_graphs()) {
        drop_graph();
        return check(hipErrorUnknown, "graph capture", err);
    }
    // femmd2D/src/integrator.cpp:8-29 per step; the second half-kick of
    // step s is fused with the first half-kick of step s+1 in the force pass
    // of step s.  Observed every step (obs_every_): every step's energies,
    // virials and kinetic energy are formed by its force pass and node kick,
    // and accumulated by the next body (or the close); otherwise only the
    // last step's, by the close.
    // The internal order is re-sorted between pieces of a run every
    // reorder_every_ steps (asynchronously, outside the step graphs).
    auto advance = [&](int done) -> int {
        since_reorder_ += done;
        if (reorder_every_ <= 0 || since_reorder_ < reorder_every_) return 0;
        since_reorder_ = 0;
        return launch_reorder_now();
    };
    auto fail = [&](const char *what) {
        obs_valid_ = false;
        const hipError_t e = hipGetLastError();
        return check(e != hipSuccess ? e : hipErrorUnknown, what, err);
    };
    if (launch_bodies(kind_open())) return fail("step");
    if (advance(1)) return fail("reorder");
    int rem = nsteps - 1;  // drifts still to do
    if (rem > 0) {
        if (launch_bodies(kind_first())) return fail("step");
        if (advance(1)) return fail("reorder");
        --rem;
    }
    while (rem >= kGraphSteps) {
        if (launch_bodies(kind_body(kGraphSteps))) return fail("step");
        rem -= kGraphSteps;
        if (advance(kGraphSteps)) return fail("reorder");
    }
    for (; rem > 0; --rem) {
        if (launch_bodies(kind_body(1))) return fail("step");
        if (advance(1)) return fail("reorder");
    }
    if (launch_bodies(kind_close(nsteps >= 2))) return fail("final kick");
    steps_ += nsteps;
    ene_valid_ = true;  // the closing force pass formed the energies
    obs_valid_ = true;
    return sync(err);
}

int Engine::sync(std::string &err) {
    int rc = check(hipStreamSynchronize(stream_), "step", err);
    if (rc) return rc;
    uint32_t e = errors();
    if (e) {
        err = dev_error_text(e);
        return FM_ERR_DEVICE;
    }
    return FM_OK;
}

uint32_t Engine::errors() {
    uint32_t e = 0;
    (void)hipMemcpyAsync(&e, err_.p, sizeof(e), hipMemcpyDeviceToHost, stream_);
    (void)hipStreamSynchronize(stream_);
    return e;
}

int Engine::clear_errors() {
    (void)hipMemsetAsync(err_.p, 0, sizeof(uint32_t), stream_);
    (void)hipStreamSynchronize(stream_);
    return FM_OK;
}

int Engine::compute_forces(std::string &err) {
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    primed_ = false;
    obs_valid_ = false;
    return form_observables(err);
}

// Forces, energies and virials at the current positions (the same bits as
// a step's force pass: same lists, or the same deterministic sort), the
// kinetic-energy partials of the current velocities, and the observables of
// this state (not accumulated).
int Engine::form_observables(std::string &err) {
    if (obs_valid_) return FM_OK;
    int rc = ensure_primed(err);
    if (rc) return rc;
    if (!ene_valid_) {
        rc = energy_pass(err);
        if (rc) return rc;
    }
    KickArgs k{};
    k.spans = kspans_.p;
    k.vx = vx_.p;
    k.vy = vy_.p;
    k.mass = mass_.p;
    k.ke_part = ke_part_.p;
    if (launch_kick(4, false, false, (int)kspans_h_.size(), k, stream_) != hipSuccess)
        return check(hipGetLastError(), "ke", err);
    ObsArgs o = obs_args();
    o.ke_kick = 1;  // the partials k_ke just formed
    if (launch_observe(o, false, stream_) != hipSuccess)
        return check(hipGetLastError(), "observe", err);
    rc = sync(err);
    if (rc) return rc;
    obs_valid_ = true;
    return FM_OK;
}

int Engine::energies(int r, fm_energies *out, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = form_observables(err);
    if (rc) return rc;
    obs_h_.resize((size_t)kObsStride * R_);
    rc = check(hipMemcpyAsync(obs_h_.data(), obs_.p, obs_h_.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream_), "energy copy", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    const double *v = &obs_h_[(size_t)kObsStride * r];
    const RepDesc &D = reps_h_[r];
    const double n = D.p1 - D.p0, V = (double)D.Lx * D.Ly;
    std::memset(out, 0, sizeof(*out));
    out->ke = 0.5 * v[0];
    out->vv_max = v[1];
    out->u_lj = v[2];
    out->w_lj = v[3];
    out->u_fea = v[4];
    out->w_fea = v[5];
    out->u_bond = v[6];
    out->w_bond = v[7];
    // femmd2D/src/properties.cpp:18-34
    out->temperature = out->ke / n;
    out->pressure = 0.5 * v[0] / V + 0.5 * (out->w_lj + out->w_fea) / V;
    out->steps = steps_;
    out->n = (int64_t)n;
    return FM_OK;
}

// Running averages since the last reset (Prop::avg, femmd2D/src/properties.h:11).
int Engine::averages(int r, fm_averages *out, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    obs_h_.resize((size_t)kObsStride * R_);
    rc = check(hipMemcpyAsync(obs_h_.data(), obs_.p, obs_h_.size() * sizeof(double),
                              hipMemcpyDeviceToHost, stream_), "averages copy", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    const double *v = &obs_h_[(size_t)kObsStride * r];
    std::memset(out, 0, sizeof(*out));
    const double cnt = v[kObsRaw];
    out->count = (int64_t)cnt;
    for (int k = 0; k < kObsAvg; ++k) {
        const double s = v[kObsRaw + 1 + k], s2 = v[kObsRaw + 1 + kObsAvg + k];
        const double m = cnt > 0 ? s / cnt : 0.0;
        out->mean[k] = m;
        out->stdv[k] = cnt > 0 ? std::sqrt(std::max(s2 / cnt - m * m, 0.0)) : 0.0;
    }
    return FM_OK;
}

int Engine::reset_averages(std::string &err) {
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    // keep each replica's last totals, zero its count and sums
    for (int r = 0; r < R_; ++r) {
        rc = check(hipMemsetAsync(obs_.p + (size_t)kObsStride * r + kObsRaw, 0,
                                  (kObsStride - kObsRaw) * sizeof(double), stream_),
                   "reset", err);
        if (rc) return rc;
    }
    return sync(err);
}

int Engine::set_observe_every(int every, std::string &err) {
    if (every < 0 || every > 1) { err = "observe_every must be 0 or 1"; return FM_ERR_ARG; }
    if (every != obs_every_) drop_graph();  // the bodies of a run change
    obs_every_ = every;
    return FM_OK;
}

int Engine::get_state(int r, float *x, float *y, float *vx, float *vy, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    {
        const int orc = ensure_order(err);
        if (orc) return orc;
    }
    const RepDesc &D = reps_h_[r];
    const int n = D.p1 - D.p0;
    std::vector<float> buf(n);
    float *outs[4] = {x, y, vx, vy};
    float *srcs[4] = {x_.p, y_.p, vx_.p, vy_.p};
    const std::vector<int32_t> &iid = iid_h_[r];
    for (int q = 0; q < 4; ++q) {
        if (!outs[q]) continue;
        int rc = check(hipMemcpyAsync(buf.data(), srcs[q] + D.p0, n * sizeof(float),
                                      hipMemcpyDeviceToHost, stream_), "get_state", err);
        if (rc) return rc;
        rc = check(hipStreamSynchronize(stream_), "get_state", err);
        if (rc) return rc;
        for (int u = 0; u < n; ++u) outs[q][u] = buf[iid[u] - D.p0];
    }
    return FM_OK;
}

int Engine::set_state(int r, const float *x, const float *y, const float *vx, const float *vy,
                      std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    {
        const int orc = ensure_order(err);
        if (orc) return orc;
    }
    const RepDesc &D = reps_h_[r];
    const int n = D.p1 - D.p0;
    std::vector<float> buf(n);
    const float *ins[4] = {x, y, vx, vy};
    float *dsts[4] = {x_.p, y_.p, vx_.p, vy_.p};
    const std::vector<int32_t> &iid = iid_h_[r];
    for (int q = 0; q < 4; ++q) {
        if (!ins[q]) continue;
        for (int u = 0; u < n; ++u) {
            float v = ins[q][u];
            if (q == 0) v = wrap_host(v, D.Lx);
            if (q == 1) v = wrap_host(v, D.Ly);
            buf[iid[u] - D.p0] = v;
        }
        int rc = check(hipMemcpyAsync(dsts[q] + D.p0, buf.data(), n * sizeof(float),
                                      hipMemcpyHostToDevice, stream_), "set_state", err);
        if (rc) return rc;
        rc = check(hipStreamSynchronize(stream_), "set_state", err);
        if (rc) return rc;
    }
    if (x || y) primed_ = false;
    // the kinetic-energy partials (velocities) and the energies (positions)
    // of the last formed observables no longer describe the state, and a
    // recorded adjoint forward pass no longer ends at it
    if (x || y || vx || vy) {
        obs_valid_ = false;
        adj_.valid = false;
    }
    return FM_OK;
}

int Engine::get_forces(int r, float *fx, float *fy, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    {
        const int orc = ensure_order(err);
        if (orc) return orc;
    }
    int rc = ensure_primed(err);
    if (rc) return rc;
    // corner forces are held in gather order (no entry table)
    if (launch_total_force(N_, fx_.p, fy_.p, cf_.p, gather_ ? goff_.p : nullptr, nullptr,
                           outx_.p, outy_.p, stream_) != hipSuccess)
        return check(hipGetLastError(), "total force", err);
    const RepDesc &D = reps_h_[r];
    const int n = D.p1 - D.p0;
    std::vector<float> bx(n), by(n);
    rc = check(hipMemcpyAsync(bx.data(), outx_.p + D.p0, n * sizeof(float), hipMemcpyDeviceToHost,
                              stream_), "forces", err);
    if (rc) return rc;
    rc = check(hipMemcpyAsync(by.data(), outy_.p + D.p0, n * sizeof(float), hipMemcpyDeviceToHost,
                              stream_), "forces", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    const std::vector<int32_t> &iid = iid_h_[r];
    for (int u = 0; u < n; ++u) {
        if (fx) fx[u] = bx[iid[u] - D.p0];
        if (fy) fy[u] = by[iid[u] - D.p0];
    }
    return FM_OK;
}

int Engine::get_cells(int r, int32_t *ncx, int32_t *ncy, int32_t *key, int32_t *start,
                      int32_t *perm, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    {
        const int orc = ensure_order(err);
        if (orc) return orc;
    }
    int rc = ensure_primed(err);
    if (rc) return rc;
    // bin and sort the current positions (list mode sorts only on rebuilds;
    // this sort is not followed by a list build: start a fresh list epoch
    // on the next use)
    if (launch_bin() || launch_sort()) return check(hipGetLastError(), "cells", err);
    if (list_) primed_ = false;
    const RepDesc &D = reps_h_[r];
    if (ncx) *ncx = D.ncx;
    if (ncy) *ncy = D.ncy;
    const int nc = D.ncx * D.ncy, n = D.p1 - D.p0;
    std::vector<int32_t> st(nc + 1), pm(n);
    rc = check(hipMemcpyAsync(st.data(), start_.p + D.cell0, (nc + 1) * sizeof(int32_t),
                              hipMemcpyDeviceToHost, stream_), "cells", err);
    if (rc) return rc;
    rc = check(hipMemcpyAsync(pm.data(), perm_.p + D.p0, n * sizeof(int32_t),
                              hipMemcpyDeviceToHost, stream_), "cells", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    for (int c = 0; c <= nc; ++c) {
        if (start) start[c] = st[c] - D.p0;
    }
    for (int c = 0; c < nc; ++c) {
        for (int p = st[c]; p < st[c + 1]; ++p) {
            const int internal = pm[p - D.p0];
            const int u = uid_h_[internal];
            if (key) key[u] = c;
            if (perm) perm[p - D.p0] = u;
        }
    }
    return FM_OK;
}

int Engine::pair_sets(int r, int32_t *cnt, int64_t *s1, int64_t *s2, std::string &err) {
    if (R_ != 1 || r != 0) { err = "pair_sets is available for single systems"; return FM_ERR_STATE; }
    {
        const int orc = ensure_order(err);
        if (orc) return orc;
    }
    int rc = ensure_primed(err);
    if (rc) return rc;
    hipError_t he;
    if ((he = dbg_cnt_.alloc(N_)) != hipSuccess) return check(he, "alloc", err);
    if ((he = dbg_s1_.alloc(N_)) != hipSuccess) return check(he, "alloc", err);
    if ((he = dbg_s2_.alloc(N_)) != hipSuccess) return check(he, "alloc", err);
    if (list_) {
        // the interacting sets the neighbour lists give at the current
        // positions (lists as last built, possibly steps ago): the list pass
        // with the dump, which leaves the rebuild state alone
        const ForceArgs f = force_args();
        if (launch_forces(sigma_, types_, true, true, true, 0, block_, f, stream_) != hipSuccess)
            return check(hipGetLastError(), "pair sets", err);
    } else {
        // rebin (the debug pass must see the binning of the current positions)
        if (launch_bin()) return check(hipGetLastError(), "bin", err);
        if (launch_forces_only(true)) return check(hipGetLastError(), "pair sets", err);
    }
    std::vector<long long> a(N_), b(N_);
    rc = check(hipMemcpyAsync(cnt, dbg_cnt_.p, N_ * sizeof(int32_t), hipMemcpyDeviceToHost, stream_),
               "copy", err);
    if (rc) return rc;
    rc = check(hipMemcpyAsync(a.data(), dbg_s1_.p, N_ * sizeof(long long), hipMemcpyDeviceToHost,
                              stream_), "copy", err);
    if (rc) return rc;
    rc = check(hipMemcpyAsync(b.data(), dbg_s2_.p, N_ * sizeof(long long), hipMemcpyDeviceToHost,
                              stream_), "copy", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    for (int i = 0; i < N_; ++i) {
        s1[i] = a[i];
        s2[i] = b[i];
    }
    dbg_cnt_.release();
    dbg_s1_.release();
    dbg_s2_.release();
    // the cell walk's dump re-sorted the particles (list mode: the dump
    // wrote the forces and energies of the current positions again, same bits)
    if (!list_) primed_ = false;
    return FM_OK;
}

int Engine::time_forces(int reps, double *ms, std::string &err) {
    int rc = ensure_primed(err);
    if (rc) return rc;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, stream_);
    energy_ = false;  // the production force pass
    for (int k = 0; k < reps; ++k) {
        int bad = launch_bin();
        if (!bad && list_) {
            // a rebuild (sort + list build) and the list pair pass
            bad = launch_sort() || launch_list_build(false) || launch_pair_pass();
        } else if (!bad) {
            bad = launch_forces_only(false);
        }
        if (bad) {
            energy_ = true;
            return check(hipGetLastError(), "forces", err);
        }
    }
    energy_ = true;
    (void)hipEventRecord(b, stream_);
    rc = check(hipEventSynchronize(b), "time", err);
    float t = 0;
    (void)hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / std::max(reps, 1);
    if (rc) return rc;
    return sync(err);
}

int Engine::time_parts(int reps, double *ms_update, double *ms_pass, std::string &err) {
    int rc = ensure_primed(err);
    if (rc) return rc;
    hipEvent_t ev[3];
    for (auto &e : ev) (void)hipEventCreate(&e);
    auto cleanup = [&]() {
        for (auto &e : ev) (void)hipEventDestroy(e);
    };
    int bad = 0;
    (void)hipEventRecord(ev[0], stream_);
    // (the cell walk's update runs every step: timed with its pass)
    for (int k = 0; k < reps && !bad && list_; ++k)
        bad = launch_bin() || launch_sort() || launch_list_build(false);
    (void)hipEventRecord(ev[1], stream_);
    for (int k = 0; k < reps && !bad; ++k) {
        energy_ = true;
        bad = list_ ? launch_pair_pass() : (launch_bin() || launch_forces_only(false));
    }
    (void)hipEventRecord(ev[2], stream_);
    if (bad) {
        cleanup();
        return check(hipGetLastError(), "time parts", err);
    }
    rc = check(hipEventSynchronize(ev[2]), "time", err);
    float t1 = 0.f, t2 = 0.f;
    (void)hipEventElapsedTime(&t1, ev[0], ev[1]);
    (void)hipEventElapsedTime(&t2, ev[1], ev[2]);
    cleanup();
    *ms_update = t1 / std::max(reps, 1);
    *ms_pass = t2 / std::max(reps, 1);
    if (rc) return rc;
    return sync(err);
}

// Pair energies and virials at the current positions (the forces are
// recomputed with the same bits: same lists, or the same deterministic sort).
int Engine::energy_pass(std::string &err) {
    energy_ = true;
    const int bad = list_ ? launch_pair_pass() : (launch_bin() || launch_forces_only(false));
    if (bad) return check(hipGetLastError(), "energy pass", err);
    return FM_OK;
}

int Engine::list_stats(int r, double *mean, int32_t *maxlen, int64_t *nbuild, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    if (mean) *mean = 0.0;
    if (maxlen) *maxlen = 0;
    if (nbuild) *nbuild = 0;
    if (!list_) return FM_OK;
    int rc = ensure_primed(err);
    if (rc) return rc;
    const RepDesc &D = reps_h_[r];
    const int n = D.p1 - D.p0;
    std::vector<int32_t> nn(n);
    Ctl c{};
    rc = check(hipMemcpyAsync(nn.data(), nnum_.p + D.p0, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                              stream_), "list stats", err);
    if (rc) return rc;
    rc = check(hipMemcpyAsync(&c, ctl_.p + r, sizeof(Ctl), hipMemcpyDeviceToHost, stream_),
               "list stats", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    double s = 0.0;
    int32_t m = 0;
    for (int32_t v : nn) {
        v &= kListLen;  // length | image flag | (tile index << 16)
        s += v;
        m = std::max(m, v);
    }
    if (mean) *mean = n ? s / n : 0.0;
    if (maxlen) *maxlen = m;
    if (nbuild) *nbuild = c.nbuild;
    return FM_OK;
}

}  // namespace fm
