// adjoint.hip -- discrete adjoint of the KDK step with checkpointed
// recomputation (no reference counterpart).
//
// Forward step n:  v' = v_n + h M^-1 f(x_n);  x_{n+1} = x_n + dt v';
//                  v_{n+1} = v' + h M^-1 f(x_{n+1}),   h = dt / 2.
// With xb = dL/dx, vb = dL/dv and the two half-kicks that share f(x_n)
// merged, the backward sweep from xb = dL/dx_N, vb = 0 is, for n = N-1..0:
//     vb += dt xb                          (drift of step n)
//     w   = c_n M^-1 vb,  c_n = dt (n > 0), h (n = 0)
//     xb += J(x_n) w,  theta_b += (df/dtheta)(x_n)^T w
// J = df/dx is symmetric, so J^T w = J w.  The wrap is the identity
// almost everywhere.  Positions x_n are recomputed per segment from
// checkpoints (x, v) taken every `ckpt` steps (the last segment is recorded
// during the forward pass, so with ckpt >= nsteps nothing is recomputed);
// parameter gradients accumulate in FP64 (per element for E, per force tile
// for eps, sigma).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"

namespace fm {

int Engine::adj_alloc(std::string &err) {
    if (R_ != 1) { err = "the adjoint is available for single systems"; return FM_ERR_STATE; }
    if (types_) { err = "the adjoint supports one particle type"; return FM_ERR_STATE; }
    if ((size_t)lds_cap_ * (sizeof(float4) + sizeof(float2)) > 65536) {
        err = "adjoint tiles exceed the LDS: choose a smaller tile width";
        return FM_ERR_STATE;
    }
    hipError_t he;
    const size_t ntiles = std::max<size_t>(tiles_h_.size(), 1);
    if (pacc_.n != 2 * ntiles && (he = pacc_.alloc(2 * ntiles)) != hipSuccess)
        return check(he, "alloc", err);
    if (w_.p) return FM_OK;
    const size_t N = N_;
#define FM_AL(buf, cnt) \
    if ((he = buf.alloc(cnt)) != hipSuccess) return check(he, "alloc " #buf, err)
    FM_AL(w_, N);
    FM_AL(jx_, N);
    FM_AL(jy_, N);
    FM_AL(xbx_, N);
    FM_AL(xby_, N);
    FM_AL(vbx_, N);
    FM_AL(vby_, N);
    FM_AL(cfa_, std::max(1, 3 * ntri_ + 2 * nbond_));
    FM_AL(dE_, std::max(1, ntri_));
#undef FM_AL
    // particles outside every force tile (LJ off) keep a zero pair part
    if ((he = hipMemsetAsync(jx_.p, 0, N * sizeof(float), stream_)) != hipSuccess)
        return check(he, "memset", err);
    if ((he = hipMemsetAsync(jy_.p, 0, N * sizeof(float), stream_)) != hipSuccess)
        return check(he, "memset", err);
    return FM_OK;
}

// Neighbour lists built at the configuration held at xy (x, then y: a
// history slot), the way the forward pass built them there: the same sort
// of the same positions, so the same lists.
int Engine::adj_lists_at(float *xy, std::string &err) {
    xalt_ = xy;
    yalt_ = xy + N_;
    const int bad = launch_bin() || launch_sort() || launch_list_build(false);
    xalt_ = yalt_ = nullptr;
    if (bad) return check(hipGetLastError(), "adjoint lists", err);
    return FM_OK;
}

// J w over the lists of the configuration in xalt_ / yalt_ (lists built at
// the forward pass's last rebuild at or before it).
int Engine::adj_pass_list(std::string &err) {
    AdjArgs A{};
    A.f = force_args();
    A.w = w_.p;
    A.jx = jx_.p;
    A.jy = jy_.p;
    A.cfa = cfa_.p;
    A.eE = eE_.p;
    A.dE = dE_.p;
    A.pacc = pacc_.p;
    if (launch_adj_forces(sigma_, true, block_, A, stream_) != hipSuccess)
        return check(hipGetLastError(), "adjoint forces", err);
    return FM_OK;
}

int Engine::adj_pass(std::string &err) {
    if (launch_bin()) return check(hipGetLastError(), "bin", err);
    if (launch_sort()) return check(hipGetLastError(), "sort", err);
    AdjArgs A{};
    A.f = force_args();
    A.w = w_.p;
    A.jx = jx_.p;
    A.jy = jy_.p;
    A.cfa = cfa_.p;
    A.eE = eE_.p;
    A.dE = dE_.p;
    A.pacc = pacc_.p;
    if (launch_adj_forces(sigma_, false, block_, A, stream_) != hipSuccess)
        return check(hipGetLastError(), "adjoint forces", err);
    return FM_OK;
}

int Engine::adj_kick(bool accumulate, bool drift, bool make_w, float c) {
    AdjKickArgs k{};
    k.reps = reps_.p;
    k.spans = kspans_.p;
    k.xbx = xbx_.p;
    k.xby = xby_.p;
    k.vbx = vbx_.p;
    k.vby = vby_.p;
    k.jx = jx_.p;
    k.jy = jy_.p;
    k.cfa = cfa_.p;
    k.goff = gather_ ? goff_.p : nullptr;
    k.gent = gent_.p;
    k.inv_m = inv_m_.p;
    k.w = w_.p;
    k.c = c;
    k.accumulate = accumulate;
    k.drift = drift;
    k.make_w = make_w;
    return launch_adj_kick((int)kspans_h_.size(), k, stream_) == hipSuccess ? 0 : -1;
}

int Engine::save_state(float *dst) {
    const size_t b = (size_t)N_ * sizeof(float);
    const float *src[4] = {x_.p, y_.p, vx_.p, vy_.p};
    for (int q = 0; q < 4; ++q)
        if (hipMemcpyAsync(dst + (size_t)q * N_, src[q], b, hipMemcpyDeviceToDevice, stream_) !=
            hipSuccess)
            return -1;
    return 0;
}

int Engine::load_state(const float *src) {
    const size_t b = (size_t)N_ * sizeof(float);
    float *dst[4] = {x_.p, y_.p, vx_.p, vy_.p};
    for (int q = 0; q < 4; ++q)
        if (hipMemcpyAsync(dst[q], src + (size_t)q * N_, b, hipMemcpyDeviceToDevice, stream_) !=
            hipSuccess)
            return -1;
    primed_ = false;
    obs_valid_ = false;
    return 0;
}

namespace {
double sum_pairs(const std::vector<double> &v, int which) {
    double s = 0.0;
    for (size_t k = which; k < v.size(); k += 2) s += v[k];
    return s;
}
}  // namespace

int Engine::adjoint_op(const float *wx, const float *wy, float *jx, float *jy, double *dE,
                       double *deps, double *dsig, std::string &err) {
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_order(err);
    if (rc) return rc;
    rc = adj_alloc(err);
    if (rc) return rc;
    const int N = N_;
    const std::vector<int32_t> &iid = iid_h_[0];
    std::vector<float2> wh(N);
    for (int u = 0; u < N; ++u) wh[iid[u]] = make_float2(wx[u], wy[u]);
    hipError_t he;
    if ((he = hipMemcpyAsync(w_.p, wh.data(), N * sizeof(float2), hipMemcpyHostToDevice,
                             stream_)) != hipSuccess)
        return check(he, "upload w", err);
    if ((he = hipMemsetAsync(dE_.p, 0, dE_.n * sizeof(double), stream_)) != hipSuccess)
        return check(he, "memset", err);
    if ((he = hipMemsetAsync(pacc_.p, 0, pacc_.n * sizeof(double), stream_)) != hipSuccess)
        return check(he, "memset", err);
    rc = adj_pass(err);
    if (rc) return rc;
    primed_ = false;  // the forward force buffers were not refreshed
    if (launch_total_force(N, jx_.p, jy_.p, cfa_.p, gather_ ? goff_.p : nullptr, gent_.p,
                           outx_.p, outy_.p, stream_) != hipSuccess)
        return check(hipGetLastError(), "total", err);
    std::vector<float> bx(N), by(N);
    std::vector<double> de(dE_.n), pa(pacc_.n);
    (void)hipMemcpyAsync(bx.data(), outx_.p, N * sizeof(float), hipMemcpyDeviceToHost, stream_);
    (void)hipMemcpyAsync(by.data(), outy_.p, N * sizeof(float), hipMemcpyDeviceToHost, stream_);
    (void)hipMemcpyAsync(de.data(), dE_.p, de.size() * sizeof(double), hipMemcpyDeviceToHost,
                         stream_);
    (void)hipMemcpyAsync(pa.data(), pacc_.p, pa.size() * sizeof(double), hipMemcpyDeviceToHost,
                         stream_);
    rc = sync(err);
    if (rc) return rc;
    for (int u = 0; u < N; ++u) {
        if (jx) jx[u] = bx[iid[u]];
        if (jy) jy[u] = by[iid[u]];
    }
    if (dE)
        for (int e = 0; e < ntri_; ++e) dE[ecaller_h_[e]] = de[e];
    if (deps) *deps = sum_pairs(pa, 0);
    if (dsig) *dsig = sum_pairs(pa, 1);
    return FM_OK;
}

// Graphs of recorded segments (adj_record_segment) carry this bit in their
// key (step-body graphs use bits 0..14); segments longer than
// kMaxGraphSegment run eagerly.
constexpr int kSegmentGraphKey = 1 << 20;
constexpr int kMaxGraphSegment = 256;

// ---------------------------------------------------------------------------
// Forward part of an adjoint: nsteps steps from the current state, with a
// checkpoint (x, v) at every segment start and the positions of the last
// segment recorded as they are produced.  The state advances as with step().
// The backward sweep (adjoint_sweep) then needs xb = dL/dx_N, vb = 0 seeded.
// ---------------------------------------------------------------------------
int Engine::adjoint_forward(int nsteps, int ckpt, std::string &err) {
    if (nsteps < 1) { err = "nsteps must be >= 1"; return FM_ERR_ARG; }
    adj_.valid = false;
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_order(err);
    if (rc) return rc;
    rc = adj_alloc(err);
    if (rc) return rc;
    const int N = N_;
    if (ckpt <= 0) {
        // keep as many stored configurations as a quarter of the free HBM
        // (at most 64 GiB) holds: the whole trajectory when it fits, so the
        // backward sweep recomputes nothing
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
        const size_t budget = std::min(fr / 4, (size_t)64 << 30);
        const size_t per = (size_t)2 * N * sizeof(float);
        const size_t fit = budget / per;
        ckpt = (int)std::max<size_t>(1, std::min<size_t>(fit, (size_t)nsteps));
    }
    ckpt = std::min(ckpt, nsteps);
    const int S = (nsteps + ckpt - 1) / ckpt;
    hipError_t he;
    if (ckpt_.n < (size_t)(S + 1) * 4 * N && (he = ckpt_.alloc((size_t)(S + 1) * 4 * N)) != hipSuccess)
        return check(he, "alloc checkpoints", err);
    if (hist_.n < (size_t)ckpt * 2 * N) {
        // the recording graphs hold the history's address: drop them with it
        for (auto it = graphs_.begin(); it != graphs_.end();) {
            if (it->first & kSegmentGraphKey) {
                (void)hipGraphExecDestroy(it->second);
                it = graphs_.erase(it);
            } else {
                ++it;
            }
        }
        if ((he = hist_.alloc((size_t)ckpt * 2 * N)) != hipSuccess)
            return check(he, "alloc history", err);
        if ((he = rblog_buf_.alloc((size_t)ckpt)) != hipSuccess)
            return check(he, "alloc rebuild log", err);
    }
    // ---- forward, checkpoint at every segment start and at the end; the
    // last segment records its positions as it goes (no energies needed) ----
    obs_valid_ = false;
    for (int s = 0; s < S; ++s) {
        // list mode: fresh neighbour lists at every segment start, as the
        // recomputation from its checkpoint will have (same summation order)
        // and as the backward sweep builds them there
        if (list_) primed_ = false;
        rc = ensure_primed(err);
        if (rc) return rc;
        if (save_state(ckpt_.p + (size_t)s * 4 * N)) return check(hipGetLastError(), "ckpt", err);
        const int len = std::min(ckpt, nsteps - s * ckpt);
        if (s == S - 1) {
            if (adj_record_segment(len)) return check(hipGetLastError(), "step", err);
        } else {
            // graphs of the step without observables: open, bodies, close
            int bad = launch_bodies(BodyKind{1, false, true, false, false, false, false});
            for (int k = 1; k < len && !bad; ++k)
                bad = launch_bodies(BodyKind{1, true, true, false, false, false, false});
            if (!bad) bad = launch_bodies(BodyKind{1, true, false, false, false, false, false});
            if (bad) return check(hipGetLastError(), "step", err);
        }
    }
    steps_ += nsteps;
    if (save_state(ckpt_.p + (size_t)S * 4 * N)) return check(hipGetLastError(), "ckpt", err);
    rc = sync(err);
    if (rc) return rc;
    adj_.valid = true;
    adj_.nsteps = nsteps;
    adj_.ckpt = ckpt;
    adj_.S = S;
    adj_.steps_at_end = steps_;
    return FM_OK;
}

// The positions x_0 .. x_{len-1} of a segment into history slots 0 .. len-1
// (slot 0 before the first drift; the drift of body k writes x_{k+1} into
// slot k + 1) while the segment is stepped from the current state, and the
// closing half-kick; in list mode the rebuild decision taken for x_k (body
// k's gated binning) into rblog_buf_[k] (x_0's lists are fresh).  Every
// launch of the segment is one graph, captured the first time a segment of
// this length is recorded (the slots are the same buffers for every
// segment).
int Engine::adj_record_segment(int len, bool close) {
    auto body = [&]() -> int {
        if (launch_copy2(hist_.p, x_.p, y_.p, N_, stream_) != hipSuccess) return -1;  // slot 0
        rec_ = len > 1 ? hist_.p + (size_t)2 * N_ : nullptr;
        int bad = launch_step_body(false, true);
        for (int k = 1; k < len && !bad; ++k) {
            // the drift of body k records x_{k+1} into slot k + 1
            rec_ = (k + 1 < len) ? hist_.p + (size_t)(k + 1) * 2 * N_ : nullptr;
            rblog_ = list_ ? rblog_buf_.p + k : nullptr;
            bad = launch_step_body(true, true);
        }
        rec_ = nullptr;
        rblog_ = nullptr;
        if (!bad && close) bad = launch_step_body(true, false);
        return bad;
    };
    // a long segment (the whole trajectory stored) runs once per forward
    // pass: captured, it would only cost the capture
    if (!use_graph_ || len > kMaxGraphSegment) return body();
    const int key = kSegmentGraphKey | (len << 1) | (close ? 1 : 0) | (par_ << 19);
    hipGraphExec_t ex = nullptr;
    for (auto &g : graphs_)
        if (g.first == key) ex = g.second;
    const int par0 = par_;
    const bool pend0 = bins_pending_;
    if (!ex) {
        if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
        const int bad = body();
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(stream_, &g);
        par_ = par0;  // nothing ran: the launch below sets the host state
        bins_pending_ = pend0;
        if (bad || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return -1;
        }
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return -1;
        graphs_.emplace_back(key, ex);
    }
    if (hipGraphLaunch(ex, stream_) != hipSuccess) return -1;
    // the host state the segment leaves: one drift per step of it (each
    // flips the list-mode parity), bins pending after a drift (cell walk)
    par_ = par0 ^ ((len & 1) && list_ ? 1 : 0);
    bins_pending_ = !list_ && !close;
    ene_valid_ = false;
    return 0;
}

// Backward sweep from the seeded xb = dL/dx_N (vb = 0): segment by segment,
// recomputing the positions of every segment but the last from its
// checkpoint.  Accumulates dE (per element) and the eps / sigma terms (per
// tile); restores the state at the end of the forward pass.
int Engine::adjoint_sweep(std::string &err) {
    const int N = N_, nsteps = adj_.nsteps, ckpt = adj_.ckpt, S = adj_.S;
    int rc;
    const float dt = reps_h_[0].dt, h = reps_h_[0].hdt;
    bool first = true;
    std::vector<int32_t> rb;
    for (int s = S - 1; s >= 0; --s) {
        const int len = std::min(ckpt, nsteps - s * ckpt);
        if (s < S - 1) {
            // recompute x_{sC} .. x_{sC+len-1} of this segment
            if (load_state(ckpt_.p + (size_t)s * 4 * N)) return check(hipGetLastError(), "restore", err);
            rc = ensure_primed(err);
            if (rc) return rc;
            if (adj_record_segment(len, false)) return check(hipGetLastError(), "step", err);
        }
        if (list_) {
            // where the forward pass built its lists in this segment: at x_0
            // and wherever the rebuild rule fired
            rb.assign(len, 0);
            if (len > 1) {
                rc = check(hipMemcpyAsync(rb.data() + 1, rblog_buf_.p + 1, (len - 1) * sizeof(int32_t),
                                          hipMemcpyDeviceToHost, stream_), "rebuild log", err);
                if (rc) return rc;
                rc = sync(err);
                if (rc) return rc;
            }
            rb[0] = 1;
        }
        int built = -1;  // the configuration whose lists are in place
        for (int k = len - 1; k >= 0; --k) {
            const int n = s * ckpt + k;
            if (first) {
                if (adj_kick(false, true, true, n > 0 ? dt : h)) return check(hipGetLastError(), "adj kick", err);
                first = false;
            }
            if (list_) {
                // the lists the forward pass used for x_k: those it built at
                // its last rebuild at or before k (every list is valid for the
                // configurations up to the next rebuild, by the rebuild rule)
                int b = k;
                while (!rb[b]) --b;
                if (b != built) {
                    rc = adj_lists_at(hist_.p + (size_t)b * 2 * N, err);
                    if (rc) return rc;
                    built = b;
                }
            }
            xalt_ = hist_.p + (size_t)k * 2 * N;
            yalt_ = xalt_ + N;
            rc = list_ ? adj_pass_list(err) : adj_pass(err);
            xalt_ = yalt_ = nullptr;
            if (rc) return rc;
            const bool more = n > 0;
            if (adj_kick(true, more, more, (n - 1 > 0) ? dt : h))
                return check(hipGetLastError(), "adj kick", err);
        }
    }
    // ---- restore the state at the end of the forward pass ----
    if (load_state(ckpt_.p + (size_t)S * 4 * N)) return check(hipGetLastError(), "restore", err);
    return ensure_primed(err);
}

int Engine::adj_grads(double *dE, double *deps, double *dsig, std::string &err) {
    std::vector<double> de(dE_.n), pa(pacc_.n);
    (void)hipMemcpyAsync(de.data(), dE_.p, de.size() * sizeof(double), hipMemcpyDeviceToHost, stream_);
    (void)hipMemcpyAsync(pa.data(), pacc_.p, pa.size() * sizeof(double), hipMemcpyDeviceToHost, stream_);
    const int rc = sync(err);
    if (rc) return rc;
    if (dE)
        for (int e = 0; e < ntri_; ++e) dE[ecaller_h_[e]] = de[e];
    if (deps) *deps = sum_pairs(pa, 0);
    if (dsig) *dsig = sum_pairs(pa, 1);
    return FM_OK;
}

int Engine::adj_zero() {
    const size_t fb = (size_t)N_ * sizeof(float);
    if (hipMemsetAsync(xbx_.p, 0, fb, stream_) != hipSuccess ||
        hipMemsetAsync(xby_.p, 0, fb, stream_) != hipSuccess ||
        hipMemsetAsync(vbx_.p, 0, fb, stream_) != hipSuccess ||
        hipMemsetAsync(vby_.p, 0, fb, stream_) != hipSuccess ||
        hipMemsetAsync(dE_.p, 0, dE_.n * sizeof(double), stream_) != hipSuccess ||
        hipMemsetAsync(pacc_.p, 0, pacc_.n * sizeof(double), stream_) != hipSuccess)
        return -1;
    return 0;
}

// Backward sweep of the last adjoint_forward with the terminal cotangent
// (dL/dx_N, dL/dy_N) in caller order (either may be null: zero).
int Engine::adjoint_backward(const float *dx, const float *dy, double *dE, double *deps,
                             double *dsig, std::string &err) {
    if (!adj_.valid || adj_.steps_at_end != steps_) {
        err = "no recorded forward pass: call fm_adjoint_forward first (the state has moved since)";
        return FM_ERR_STATE;
    }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_order(err);
    if (rc) return rc;
    const int N = N_;
    if (adj_zero()) return check(hipGetLastError(), "memset", err);
    const std::vector<int32_t> &iid = iid_h_[0];
    std::vector<float> bx(N, 0.0f), by(N, 0.0f);
    for (int u = 0; u < N; ++u) {
        if (dx) bx[iid[u]] = dx[u];
        if (dy) by[iid[u]] = dy[u];
    }
    const size_t fb = (size_t)N * sizeof(float);
    if (hipMemcpyAsync(xbx_.p, bx.data(), fb, hipMemcpyHostToDevice, stream_) != hipSuccess ||
        hipMemcpyAsync(xby_.p, by.data(), fb, hipMemcpyHostToDevice, stream_) != hipSuccess)
        return check(hipGetLastError(), "cotangent", err);
    rc = sync(err);  // the host buffers are reused
    if (rc) return rc;
    rc = adjoint_sweep(err);
    if (rc) return rc;
    adj_.valid = false;
    return adj_grads(dE, deps, dsig, err);
}

int Engine::adjoint_run(int nsteps, int ckpt, int nloss, const int32_t *ids, const float *tx,
                        const float *ty, double *loss, double *dE, double *deps, double *dsig,
                        std::string &err) {
    if (nsteps < 1 || nloss < 1 || !ids || !tx || !ty) {
        err = "nsteps and the loss set must be non-empty";
        return FM_ERR_ARG;
    }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_order(err);
    if (rc) return rc;
    const int N = N_;
    for (int k = 0; k < nloss; ++k)
        if (ids[k] < 0 || ids[k] >= N) { err = "loss particle out of range"; return FM_ERR_ARG; }
    rc = adjoint_forward(nsteps, ckpt, err);
    if (rc) return rc;
    // ---- the loss (1/K) sum |x_k - t_k|^2 and its gradient as the seed ----
    const std::vector<int32_t> &iid = iid_h_[0];
    std::vector<int32_t> lid(nloss);
    std::vector<float2> tgt(nloss);
    for (int k = 0; k < nloss; ++k) {
        lid[k] = iid[ids[k]];
        tgt[k] = make_float2(tx[k], ty[k]);
    }
    hipError_t he;
    if ((he = loss_ids_.alloc(nloss)) != hipSuccess) return check(he, "alloc", err);
    if ((he = loss_t_.alloc(nloss)) != hipSuccess) return check(he, "alloc", err);
    const int nlb = (nloss + 255) / 256;
    if ((he = loss_part_.alloc(nlb)) != hipSuccess) return check(he, "alloc", err);
    (void)hipMemcpyAsync(loss_ids_.p, lid.data(), nloss * sizeof(int32_t), hipMemcpyHostToDevice,
                         stream_);
    (void)hipMemcpyAsync(loss_t_.p, tgt.data(), nloss * sizeof(float2), hipMemcpyHostToDevice,
                         stream_);
    if (adj_zero()) return check(hipGetLastError(), "memset", err);
    LossArgs la{};
    la.reps = reps_.p;
    la.n = nloss;
    la.ids = loss_ids_.p;
    la.target = loss_t_.p;
    la.x = x_.p;
    la.y = y_.p;
    la.xbx = xbx_.p;
    la.xby = xby_.p;
    la.scale = 2.0f / (float)nloss;
    la.part = loss_part_.p;
    if (launch_adj_loss(la, stream_) != hipSuccess) return check(hipGetLastError(), "loss", err);
    std::vector<double> lp(nlb);
    (void)hipMemcpyAsync(lp.data(), loss_part_.p, nlb * sizeof(double), hipMemcpyDeviceToHost,
                         stream_);
    rc = sync(err);  // (lid, tgt: host buffers of the uploads above)
    if (rc) return rc;
    rc = adjoint_sweep(err);
    if (rc) return rc;
    adj_.valid = false;
    double l = 0.0;
    for (double v : lp) l += v;
    if (loss) *loss = l / nloss;
    return adj_grads(dE, deps, dsig, err);
}

}  // namespace fm
