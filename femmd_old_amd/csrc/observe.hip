// observe.hip -- observables beyond the per-step energies: elastomer cluster
// properties (femmd2D/src/feaElements.cpp:86-196), the cluster virial
// (femmd2D/src/interactions.cpp:370-382) and the velocity rescale to a
// temperature (femmd2D/src/md3dsystem.cpp:92-136).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "engine.hpp"

namespace fm {

namespace {
template <class T>
hipError_t put(DBuf<T> &b, const std::vector<T> &v, hipStream_t s) {
    hipError_t e = b.alloc(std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
}
}  // namespace

// Clusters are the elastomers in the order they were added; cluster c of
// replica r owns the caller ids [cl_first, cl_first + cl_n) of that replica.
int Engine::build_clusters(const std::vector<const HostSystem *> &reps, std::string &err) {
    std::vector<int32_t> nodes, cid(N_, -1);
    cl_rep_h_.clear();
    cl_off_h_.assign(1, 0);
    for (int r = 0; r < (int)reps.size(); ++r) {
        const HostSystem &S = *reps[r];
        for (size_t c = 0; c < S.cl_first.size(); ++c) {
            const int id = (int)cl_rep_h_.size();
            for (int32_t k = 0; k < S.cl_n[c]; ++k) {
                const int32_t i = iid_h_[r][S.cl_first[c] + k];
                nodes.push_back(i);
                cid[i] = id;
            }
            cl_rep_h_.push_back(r);
            cl_off_h_.push_back((int32_t)nodes.size());
        }
    }
    const int ncl = (int)cl_rep_h_.size();
    hipError_t he;
    if ((he = put(cl_rep_, cl_rep_h_, stream_)) != hipSuccess ||
        (he = put(cl_off_, cl_off_h_, stream_)) != hipSuccess ||
        (he = put(cl_nodes_, nodes, stream_)) != hipSuccess ||
        (he = put(cid_, cid, stream_)) != hipSuccess ||
        (he = cl_out_.alloc((size_t)std::max(ncl, 1) * kClusterOut)) != hipSuccess ||
        (he = cl_cm_.alloc(std::max(ncl, 1))) != hipSuccess) {
        (void)hipStreamSynchronize(stream_);  // nodes, cid: sources of the copies
        return check(he, "clusters", err);
    }
    // nodes and cid are the sources of asynchronous copies
    return check(hipStreamSynchronize(stream_), "clusters", err);
}

int Engine::ncluster(int r) const {
    int n = 0;
    for (int32_t v : cl_rep_h_) n += v == r;
    return n;
}

int Engine::run_cluster_props(float *ux, float *uy) {
    ClusterArgs a{};
    a.reps = reps_.p;
    a.ncl = (int)cl_rep_h_.size();
    a.crep = cl_rep_.p;
    a.off = cl_off_.p;
    a.nodes = cl_nodes_.p;
    a.x = x_.p;
    a.y = y_.p;
    a.vx = vx_.p;
    a.vy = vy_.p;
    a.mass = mass_.p;
    a.out = cl_out_.p;
    a.cm = cl_cm_.p;
    a.ux = ux;
    a.uy = uy;
    return launch_cluster_props(a, stream_) == hipSuccess ? 0 : -1;
}

int Engine::cluster_props(int r, double *out, float *ux, float *uy, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_order(err);
    if (rc) return rc;
    const bool unf = ux || uy;
    if (unf) {
        // unfolded positions of the cluster nodes, the wrapped positions of
        // every other particle
        if (hipMemcpyAsync(outx_.p, x_.p, N_ * sizeof(float), hipMemcpyDeviceToDevice, stream_) !=
                hipSuccess ||
            hipMemcpyAsync(outy_.p, y_.p, N_ * sizeof(float), hipMemcpyDeviceToDevice, stream_) !=
                hipSuccess)
            return check(hipGetLastError(), "cluster copy", err);
    }
    if (run_cluster_props(unf ? outx_.p : nullptr, unf ? outy_.p : nullptr))
        return check(hipGetLastError(), "cluster props", err);
    const int ncl = (int)cl_rep_h_.size();
    std::vector<double> o((size_t)std::max(ncl, 1) * kClusterOut);
    rc = check(hipMemcpyAsync(o.data(), cl_out_.p, o.size() * sizeof(double), hipMemcpyDeviceToHost,
                              stream_), "cluster props", err);
    if (rc) return rc;
    const RepDesc &D = reps_h_[r];
    const int n = D.p1 - D.p0;
    std::vector<float> bx, by;
    if (unf) {
        bx.resize(n);
        by.resize(n);
        (void)hipMemcpyAsync(bx.data(), outx_.p + D.p0, n * sizeof(float), hipMemcpyDeviceToHost,
                             stream_);
        (void)hipMemcpyAsync(by.data(), outy_.p + D.p0, n * sizeof(float), hipMemcpyDeviceToHost,
                             stream_);
    }
    rc = sync(err);
    if (rc) return rc;
    int k = 0;
    for (int c = 0; c < ncl; ++c) {
        if (cl_rep_h_[c] != r) continue;
        if (out)
            for (int q = 0; q < kClusterOut; ++q) out[(size_t)k * kClusterOut + q] = o[(size_t)c * kClusterOut + q];
        ++k;
    }
    if (unf) {
        const std::vector<int32_t> &iid = iid_h_[r];
        for (int u = 0; u < n; ++u) {
            if (ux) ux[u] = bx[iid[u] - D.p0];
            if (uy) uy[u] = by[iid[u] - D.p0];
        }
    }
    return FM_OK;
}

int Engine::cluster_virial(int r, double *vir, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_primed(err);
    if (rc) return rc;
    double v = 0.0;
    if (!cl_rep_h_.empty() && !tiles_h_.empty()) {
        // centres of mass, then the pairs of the current binning (a sort
        // that no list build follows: a fresh list epoch on the next use)
        if (run_cluster_props(nullptr, nullptr) || launch_bin() || launch_sort())
            return check(hipGetLastError(), "cluster virial", err);
        if (list_) primed_ = false;
        hipError_t he;
        if ((he = cl_vir_.alloc(tiles_h_.size())) != hipSuccess) return check(he, "alloc", err);
        ClusterVirialArgs A{};
        A.f = force_args();
        A.cid = cid_.p;
        A.cm = cl_cm_.p;
        A.out = cl_vir_.p;
        if (launch_cluster_virial(sigma_, types_, block_, A, stream_) != hipSuccess)
            return check(hipGetLastError(), "cluster virial", err);
        std::vector<double> t(tiles_h_.size());
        rc = check(hipMemcpyAsync(t.data(), cl_vir_.p, t.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, stream_), "cluster virial", err);
        if (rc) return rc;
        rc = sync(err);
        if (rc) return rc;
        for (size_t k = 0; k < t.size(); ++k)
            if (tiles_h_[k].rep == r) v += t[k];
    }
    if (vir) *vir = v;
    return FM_OK;
}

// femmd2D/src/md3dsystem.cpp:119-136 (all particles: T = KE / N) and
// :92-117 (a group: T = 2 KE_group / (3 N), N the whole system's count, as
// the reference has it); v *= sqrt(temp / T).
int Engine::set_temperature(int r, double temp, int n, const int32_t *ids, std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    if (!(temp >= 0.0) || !std::isfinite(temp)) { err = "temperature must be >= 0"; return FM_ERR_ARG; }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_order(err);
    if (rc) return rc;
    const RepDesc &D = reps_h_[r];
    const int np = D.p1 - D.p0;
    std::vector<int32_t> sel;
    if (n > 0) {
        if (!ids) { err = "ids required"; return FM_ERR_ARG; }
        sel.resize(n);
        for (int k = 0; k < n; ++k) {
            if (ids[k] < 0 || ids[k] >= np) { err = "particle id out of range"; return FM_ERR_ARG; }
            sel[k] = iid_h_[r][ids[k]];
        }
    } else {
        sel.resize(np);
        for (int k = 0; k < np; ++k) sel[k] = D.p0 + k;
    }
    const int m = (int)sel.size();
    hipError_t he;
    if ((he = put(sel_, sel, stream_)) != hipSuccess) return check(he, "ids", err);
    const int nb = (m + 255) / 256;
    if ((he = vel_part_.alloc(std::max(nb, 1))) != hipSuccess) return check(he, "alloc", err);
    if (launch_vel_ke(sel_.p, m, vx_.p, vy_.p, mass_.p, vel_part_.p, stream_) != hipSuccess)
        return check(hipGetLastError(), "kinetic energy", err);
    std::vector<double> part(nb);
    rc = check(hipMemcpyAsync(part.data(), vel_part_.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                              stream_), "kinetic energy", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    double ke = 0.0;
    for (double v : part) ke += v;
    const double cur = (n > 0) ? 2.0 * ke / (3.0 * np) : ke / np;
    if (!(cur > 0.0)) { err = "the selected particles are at rest"; return FM_ERR_STATE; }
    const float f = (float)std::sqrt(temp / cur);
    if (launch_vel_scale(sel_.p, m, vx_.p, vy_.p, f, stream_) != hipSuccess)
        return check(hipGetLastError(), "rescale", err);
    // the kinetic energy of the last formed observables is stale: the next
    // query forms them again from the new velocities
    obs_valid_ = false;
    adj_.valid = false;
    return sync(err);
}

// ---------------------------------------------------------------------------
// Allocation-free velocity rescale.  The same kernels and the same order of
// sums as set_temperature above, with the selected internal ids, the
// kinetic-energy partials and their total in one file-scope device array
// (`scratch`, kNoallocBytes per device, held until the process ends) instead
// of buffers allocated per call.  Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {
constexpr int kNoallocIds = 1 << 20;    // selected particles per call
constexpr int kNoallocParts = 1 << 16;  // kinetic-energy partials (256 particles each)
constexpr size_t kNoallocPartOff = (size_t)kNoallocIds * sizeof(int32_t);
constexpr size_t kNoallocTotalOff = kNoallocPartOff + (size_t)kNoallocParts * sizeof(double);
constexpr size_t kNoallocBytes = kNoallocTotalOff + 256;
constexpr int kNoallocDevices = 64;

// The host-side name of this array is a stand-in, not a device address:
// host code uses the address hipGetSymbolAddress returns for each device.
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoallocDevices] = {};

// This device's copy of `scratch`, looked up once per device with that
// device current (the caller has made it current).
hipError_t scratch_of(int device, unsigned char **out) {
    if (device < 0 || device >= kNoallocDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[device]);
    return hipSuccess;
}

// out = part[0] + part[1] + ... in index order (one thread: the order of the
// host loop of set_temperature, so the total has the same bits)
__global__ void k_sum_in_order(const double *part, int n, double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += part[k];
    *out = s;
}
}  // namespace

int Engine::set_temperature_noalloc(int r, double temp, int n, const int32_t *ids,
                                    std::string &err) {
    if (r < 0 || r >= R_) { err = "replica out of range"; return FM_ERR_ARG; }
    if (!(temp >= 0.0) || !std::isfinite(temp)) { err = "temperature must be >= 0"; return FM_ERR_ARG; }
    const RepDesc &D = reps_h_[r];
    const int np = D.p1 - D.p0;
    const int m = n > 0 ? n : np;
    const int nb = (m + 255) / 256;
    if (n > kNoallocIds) {
        err = "fm_set_temperature_noalloc: at most 1048576 ids per call";
        return FM_ERR_ARG;
    }
    if (nb > kNoallocParts) {
        err = "fm_set_temperature_noalloc: at most 16777216 particles per call";
        return FM_ERR_ARG;
    }
    if (n > 0 && !ids) { err = "ids required"; return FM_ERR_ARG; }
    for (int k = 0; k < n; ++k)
        if (ids[k] < 0 || ids[k] >= np) { err = "particle id out of range"; return FM_ERR_ARG; }
    int rc = check(hipSetDevice(device_), "hipSetDevice", err);
    if (rc) return rc;
    rc = ensure_order(err);
    if (rc) return rc;
    unsigned char *base = nullptr;
    rc = check(scratch_of(device_, &base), "hipGetSymbolAddress(scratch)", err);
    if (rc) return rc;
    int32_t *sel = reinterpret_cast<int32_t *>(base);
    double *part = reinterpret_cast<double *>(base + kNoallocPartOff);
    double *total = reinterpret_cast<double *>(base + kNoallocTotalOff);
    // the selected caller ids as internal ids, staged through a buffer on
    // the stack; all particles of a single system need no ids (internal id
    // k, in the order set_temperature lists them)
    const int32_t *dsel = nullptr;
    if (n > 0 || D.p0 != 0) {
        int32_t chunk[2048];
        for (int k0 = 0; k0 < m; k0 += 2048) {
            const int c = std::min(2048, m - k0);
            for (int k = 0; k < c; ++k)
                chunk[k] = n > 0 ? iid_h_[r][ids[k0 + k]] : D.p0 + k0 + k;
            rc = check(hipMemcpyAsync(sel + k0, chunk, c * sizeof(int32_t), hipMemcpyHostToDevice,
                                      stream_), "ids", err);
            if (rc) return rc;
            rc = check(hipStreamSynchronize(stream_), "ids", err);  // chunk is reused
            if (rc) return rc;
        }
        dsel = sel;
    }
    if (launch_vel_ke(dsel, m, vx_.p, vy_.p, mass_.p, part, stream_) != hipSuccess)
        return check(hipGetLastError(), "kinetic energy", err);
    hipLaunchKernelGGL(k_sum_in_order, dim3(1), dim3(64), 0, stream_, part, nb, total);
    rc = check(hipGetLastError(), "kinetic energy sum", err);
    if (rc) return rc;
    double ke = 0.0;
    rc = check(hipMemcpyAsync(&ke, total, sizeof(double), hipMemcpyDeviceToHost, stream_),
               "kinetic energy", err);
    if (rc) return rc;
    rc = sync(err);
    if (rc) return rc;
    const double cur = (n > 0) ? 2.0 * ke / (3.0 * np) : ke / np;
    if (!(cur > 0.0)) { err = "the selected particles are at rest"; return FM_ERR_STATE; }
    const float f = (float)std::sqrt(temp / cur);
    if (launch_vel_scale(dsel, m, vx_.p, vy_.p, f, stream_) != hipSuccess)
        return check(hipGetLastError(), "rescale", err);
    obs_valid_ = false;
    adj_.valid = false;
    return sync(err);
}

}  // namespace fm
