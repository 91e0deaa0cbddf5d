// engine.hpp -- host side of the step engine (C++17, HIP runtime).
//
// HostSystem is the caller-ordered description of one system, filled by the
// fm_add_* calls (the reference's Particles / Bonds / FeaElements host
// vectors, femmd2D/src/particles.h:41-126).  Engine owns the device image of
// one or more HostSystems (a batch of replicas; a plain system is a batch of
// one) and runs the step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "../../include/femmd.h"

namespace fm {

// the next timed launch (kick, binning, sort, list build, forces) records
// its dispatch's start and end on these events; (nullptr, nullptr) disarms
void arm_kernel_timer(hipEvent_t start, hipEvent_t stop);

struct HostSystem {
    double Lx = 0, Ly = 0, dt = 0.005;
    int device = 0;
    // particles, caller order
    std::vector<float> x, y, vx, vy, mass, sigma;
    std::vector<int32_t> type, excl;
    std::vector<float> rx, ry;  // reference (unwrapped) positions
    // pair table (reference default: WCA, femmd2D/src/interactions.cpp:25-28)
    int32_t ntypes = 1;
    std::vector<float> eps{1.0f}, rc{1.12246204830937f}, ush{1.0f};
    bool lj_on = true;
    // elements
    std::vector<int32_t> cl_first, cl_n; // elastomer clusters: caller id ranges
    std::vector<int32_t> tri;            // 3 per element, caller ids
    std::vector<float> tE, tnu;          // per element, < 0 = global
    int32_t model = FM_NEO_HOOKEAN;
    float E = 0.0f, nu = 0.3f;
    // bonds
    std::vector<int32_t> bi, bj;
    std::vector<float> bk, br0;
    int32_t next_group = 0;
    // neighbour-list skin; 0 = cell walk every step.  The reference's CPU
    // list uses 0.4 (femmd2D/src/nn.h:63); 0.6 rebuilds less often for a
    // slightly longer list and measured faster here.  Any skin gives the same
    // interacting sets.
    float skin = 0.6f;
    // launch tuning
    int32_t tile_w = 0, block = 256, use_graph = 1;
    // steps between re-orderings of the internal order (0: never)
    int32_t reorder_every = 1000;
    // observables formed and accumulated every step (1) or per run (0)
    int32_t observe_every = 1;
};

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void **)&p, count * sizeof(T));
    }
};

class Engine {
public:
    ~Engine();
    // Builds the device image of the given systems.  On failure returns a
    // negative FM_ERR_* and fills err.
    static int build(const std::vector<const HostSystem *> &reps, std::unique_ptr<Engine> &out,
                     std::string &err);

    int step(int nsteps, std::string &err);
    int compute_forces(std::string &err);
    int energies(int r, fm_energies *e, std::string &err);
    int get_state(int r, float *x, float *y, float *vx, float *vy, std::string &err);
    int set_state(int r, const float *x, const float *y, const float *vx, const float *vy,
                  std::string &err);
    int get_forces(int r, float *fx, float *fy, std::string &err);
    int get_cells(int r, int32_t *ncx, int32_t *ncy, int32_t *key, int32_t *start,
                  int32_t *perm, std::string &err);
    int pair_sets(int r, int32_t *cnt, int64_t *s1, int64_t *s2, std::string &err);
    int time_forces(int reps, double *ms, std::string &err);
    // The two parts of time_forces timed apart, per repetition: the
    // neighbour-list update of every replica (binning, sort, list build;
    // 0 for the cell walk) and the pair pass with its energies (the cell
    // walk: binning, sort and pair pass).
    int time_parts(int reps, double *ms_update, double *ms_pass, std::string &err);
    int profile(int nsteps, double *out, std::string &err);
    // neighbour-list statistics of replica r: mean and max list length,
    // rebuilds so far (0 / 0 / 0 when the engine walks cells every step)
    int list_stats(int r, double *mean, int32_t *maxlen, int64_t *nbuild, std::string &err);
    bool list_mode() const { return list_; }
    // adjoint (adjoint.hip): J(x) w and (df/dtheta)^T w at the current
    // positions, caller order; and a checkpointed forward + backward sweep
    int adjoint_op(const float *wx, const float *wy, float *jx, float *jy, double *dE,
                   double *deps, double *dsig, std::string &err);
    int adjoint_run(int nsteps, int ckpt, int nloss, const int32_t *ids, const float *tx,
                    const float *ty, double *loss, double *dE, double *deps, double *dsig,
                    std::string &err);
    // the forward part (nsteps steps recorded for the backward sweep; the
    // state advances) and the backward sweep of a terminal cotangent
    // dL/dx_N, dL/dy_N (caller order)
    int adjoint_forward(int nsteps, int ckpt, std::string &err);
    int adjoint_backward(const float *dx, const float *dy, double *dE, double *deps,
                         double *dsig, std::string &err);
    int ntri() const { return ntri_; }
    int set_tuning(int tile_w, int block, int use_graph, std::string &err);
    // internal order: re-sort the movable particles to the last counting
    // sort now, or every `every` steps of step() (0: never)
    int reorder(std::string &err);
    void set_reorder_interval(int every) { reorder_every_ = every; }
    int reorder_interval() const { return reorder_every_; }
    int64_t reorders() const { return nreorder_; }
    uint32_t errors();
    int clear_errors();
    int sync(std::string &err);
    int nrep() const { return R_; }
    int nparticles(int r) const { return reps_h_[r].p1 - reps_h_[r].p0; }
    int64_t steps() const { return steps_; }
    // bucket tuning for batches: per replica tile width
    int retile(const std::vector<int> &tile_w, int block, std::string &err);
    int ncells_of(int r) const { return reps_h_[r].ncx * reps_h_[r].ncy; }
    int set_dt(int r, double dt, std::string &err);
    const RepDesc &rep(int r) const { return reps_h_[r]; }
    // elastomer cluster observables (replica r): kClusterOut doubles per
    // cluster; unfolded positions in caller order (may be null)
    int ncluster(int r) const;
    int cluster_props(int r, double *out, float *ux, float *uy, std::string &err);
    int cluster_virial(int r, double *vir, std::string &err);
    // velocity rescale to temperature temp (the reference's setTemperature)
    int set_temperature(int r, double temp, int n, const int32_t *ids, std::string &err);
    // the same rescale with no device allocation (scratch in a file-scope
    // device array; see observe.hip)
    int set_temperature_noalloc(int r, double temp, int n, const int32_t *ids, std::string &err);
    int tile_width(int r) const { return tile_w_[r]; }
    // running averages of the per-step observables (Properties::accumulate,
    // femmd2D/src/properties.cpp:39-86)
    int averages(int r, fm_averages *out, std::string &err);
    int reset_averages(std::string &err);
    int set_observe_every(int every, std::string &err);
    int observe_every() const { return obs_every_; }
    int block() const { return block_; }
    double density_of(int r) const;

private:
    Engine() = default;
    // gated: only replicas flagged for a rebuild.  In list mode every
    // binning records the reference positions of the displacement test and
    // every sort writes the sorted-order state; a sort that no list build
    // follows must leave primed_ false.
    int launch_bin(bool gated = false);
    // One step body: kick2 (second half-kick of the previous step, with the
    // kinetic-energy partials when energy is set), drift (first half-kick,
    // drift, wrap), then the neighbour update and the force pass (pair
    // energies and virials when energy is set).  l1 / l2: the previous step's
    // observables are formed (level 1 in the binning or sort, level 2 in the
    // force pass) and accumulated.
    int launch_step_body(bool kick2, bool drift, bool energy = false, bool l1 = false,
                         bool l2 = false);
    // A run of `count` identical bodies, or the closing half-kick plus the
    // observables of the last step (close): replayed from a graph captured
    // on first use (use_graph_), else launched one by one.
    struct BodyKind {
        int count;
        bool kick2, drift, energy, l1, l2, close;
        int key() const {
            return count | (kick2 << 8) | (drift << 9) | (energy << 10) | (l1 << 11) | (l2 << 12) |
                   (close << 13);
        }
    };
    int launch_bodies(const BodyKind &k);
    int launch_bodies_eager(const BodyKind &k);
    BodyKind kind_open() const;
    BodyKind kind_first() const;
    BodyKind kind_body(int count) const;
    BodyKind kind_close(bool after_body) const;
    // the graph of kind k, captured (not launched) if missing; null when
    // graphs are off or the capture failed
    hipGraphExec_t graph_of(const BodyKind &k);
    // captures every kind of body a run launches in the current mode, so
    // that the first run of any length already replays graphs only
    int prepare_graphs();
    int launch_forces_only(bool debug);
    int launch_sort(bool gated = false);
    int launch_list_build(bool gated);
    int launch_pair_pass();  // list mode: list pair pass + elements + bonds
    ForceArgs force_args();
    KickArgs kick_args(bool drift);
    ObsArgs obs_args();
    // list mode: the sorted-order positions and displacement slots come in
    // pairs; par_ names the buffer that holds the current positions (and the
    // slots of the drift that produced them), the other one receives the
    // next drift's
    int par_ = 0;
    // cell walk: a drift's binning not yet consumed by a sort (a body ends
    // with the next step's binning pending); a fresh binning clears it first
    bool bins_pending_ = false;
    float2 *xys_cur() const { return xys_.p + (size_t)par_ * N_; }
    float2 *xys_nxt() const { return xys_.p + (size_t)(par_ ^ 1) * N_; }
    uint32_t *dmax_cur() const { return dmax_.p + (size_t)par_ * R_ * kDispSlots * kDispStride; }
    uint32_t *dmax_nxt() const {
        return dmax_.p + (size_t)(par_ ^ 1) * R_ * kDispSlots * kDispStride;
    }
    // the observables of the current state (no accumulation), when the last
    // formed step is not the current state
    int form_observables(std::string &err);
    // positions the binning / force kernels read: the live state, or a
    // stored configuration during the adjoint sweep
    float *xcur() const { return xalt_ ? xalt_ : x_.p; }
    float *ycur() const { return yalt_ ? yalt_ : y_.p; }
    int adj_alloc(std::string &err);
    int adj_pass(std::string &err);  // bin + sort + adjoint force kernel at xcur
    // list mode backward sweep at xcur: lists rebuilt (fresh, or gated on the
    // displacement from the last build) + adjoint kernel over the lists
    int adj_lists_at(float *xy, std::string &err);
    int adj_pass_list(std::string &err);
    int adj_kick(bool accumulate, bool drift, bool make_w, float c);
    // a segment of len steps from the current state with its positions
    // recorded into the history slots (and the closing half-kick)
    int adj_record_segment(int len, bool close = true);
    int adjoint_sweep(std::string &err);
    int adj_grads(double *dE, double *deps, double *dsig, std::string &err);
    int adj_zero();
    struct AdjRun {
        bool valid = false;
        int nsteps = 0, ckpt = 0, S = 0;
        int64_t steps_at_end = 0;
    } adj_;
    int save_state(float *dst);
    int load_state(const float *src);
    int build_tiles(const std::vector<int> &tile_w, std::string &err);
    int build_clusters(const std::vector<const HostSystem *> &reps, std::string &err);
    int run_cluster_props(float *ux, float *uy);
    int ensure_primed(std::string &err);
    int launch_reorder_now();
    // the host maps uid_h_ / iid_h_ follow the device's uid after a
    // re-ordering (refreshed on first use)
    int ensure_order(std::string &err);
    int check(hipError_t e, const char *what, std::string &err);
    int drop_graph();

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    int R_ = 0, N_ = 0, ncells_ = 0, ntri_ = 0, nbond_ = 0, cap_ = 32, lds_cap_ = 2048;
    int block_ = 256, use_graph_ = 1;
    bool sigma_ = false, types_ = false, gather_ = false;
    bool list_ = false;  // neighbour-list mode
    float skin_ = 0.0f;
    int nbr_cap_ = 0;
    bool primed_ = false;
    // pair energies and virials are accumulated only by force passes with
    // energy_ set (the last step of a run and every explicit query);
    // ene_valid_: the partials belong to the current positions
    bool energy_ = true, ene_valid_ = false;
    int energy_pass(std::string &err);
    // per-step observables: formed and accumulated every step (1, the
    // reference's doStep) or only for the last step of each run (0)
    int obs_every_ = 1;
    bool obs_valid_ = false;  // obs_ holds the totals of the current state
    bool l1_ = false, l2_ = false;  // flags of the body being launched
    std::vector<char> lj_on_;
    int64_t steps_ = 0;
    int reorder_every_ = 1000;
    int64_t since_reorder_ = 0, nreorder_ = 0;
    bool order_stale_ = false;
    int nmov_ = 0;
    int model_ = 0;
    double rho_cp_ = 1.16;

    std::vector<RepDesc> reps_h_;
    std::vector<int32_t> uid_h_;               // internal -> caller id (per replica)
    std::vector<std::vector<int32_t>> iid_h_;  // per replica: caller -> internal
    std::vector<int> tile_w_;                  // per replica tile width in use
    std::vector<Span> kspans_h_, espans_h_, bspans_h_;
    std::vector<Tile> tiles_h_;
    std::vector<int32_t> rng_h_;  // per replica 8 ints (energy ranges)

    DBuf<RepDesc> reps_;
    DBuf<PairEntry> ptab_;
    DBuf<float> x_, y_, vx_, vy_, fx_, fy_, mass_, inv_m_, sig_;
    DBuf<int32_t> grp_, typ_, uid_;
    DBuf<uint8_t> mov_;                // per internal id: movable
    DBuf<int32_t> movids_, ro_bsum_, ro_oldof_;
    DBuf<int32_t> fspan_, ftile_, fblock_, nf_;  // batches: compact rebuild work lists
    DBuf<int4> rinfo_;
    DBuf<uint8_t> mark_;
    DBuf<unsigned long long> sort_status2_;     // status words of the listed sort
    DBuf<int2> brep_;                   // per sort block: replica range
    DBuf<uint32_t> ro_tmp_;
    DBuf<int32_t> count_, tmp_, start_, perm_, sorted_typ_;
    DBuf<unsigned long long> sort_status_, sort_ticket_;
    int nsort_ = 0;
    int lb_grid_ = 0;  // gated list-build grid (resident workgroups)
    DBuf<float4> sorted_;
    DBuf<Span> kspans_, espans_, bspans_;
    DBuf<int32_t> nbr_, nnum_, any_;
    DBuf<Ctl> ctl_;
    DBuf<float> sigs_;
    DBuf<uint32_t> dmax_;
    DBuf<int32_t> spos_;
    DBuf<float2> xys_;
    DBuf<Tile> tiles_;
    DBuf<int32_t> t0_, t1_, t2_, bi_, bj_, goff_, gent_, gpos_;
    DBuf<float4> dmbm_;
    DBuf<float2> alm_, cf_;
    DBuf<float> bk_, br0_;
    DBuf<double> ke_part_, lj_part_, fea_part_, bond_part_;
    // node kick: the number of particles it integrates (per replica
    // [pg, p1)), its spans (per replica [nrng[2r], nrng[2r+1])) and their
    // kinetic-energy partials; the force pass's partials of the particles it
    // integrates, per tile
    int nnode_ = 0;
    DBuf<int32_t> nrng_;
    DBuf<Span> nspans_;
    std::vector<Span> nspans_h_;
    DBuf<double> ke_node_, ke_tile_;
    DBuf<double> obs_chunk_, obs_;  // observables: level-1 chunks, per-replica records
    std::vector<double> obs_h_;
    DBuf<int32_t> rng_;
    DBuf<uint32_t> err_;
    DBuf<float> outx_, outy_;
    DBuf<int32_t> dbg_cnt_;
    DBuf<long long> dbg_s1_, dbg_s2_;

    // profiling: a start and an end event per timed kernel of a step body
    // (kick, binning, sort, list build, forces; set only while profiling)
    static constexpr int kProfKernels = 5;
    hipEvent_t *prof_ev_ = nullptr;
    void time_next(int k) {
        if (prof_ev_) arm_kernel_timer(prof_ev_[2 * k], prof_ev_[2 * k + 1]);
    }
    // captured runs of bodies, keyed by BodyKind::key()
    std::vector<std::pair<int, hipGraphExec_t>> graphs_;

    // adjoint state
    float *xalt_ = nullptr, *yalt_ = nullptr;
    float *rec_ = nullptr;  // drift of the next step body records x, y here
    int32_t *rblog_ = nullptr;  // its gated binning records the rebuild decision here
    std::vector<int32_t> ecaller_h_;  // internal element -> caller element index
    DBuf<float> eE_;                  // Young's modulus per element
    DBuf<float2> w_, cfa_;
    DBuf<float> jx_, jy_, xbx_, xby_, vbx_, vby_;
    DBuf<double> dE_, pacc_, loss_part_;
    DBuf<float> ckpt_, hist_;
    DBuf<int32_t> rblog_buf_;  // per recorded configuration: lists rebuilt there

    // elastomer clusters
    std::vector<int32_t> cl_rep_h_, cl_off_h_;  // replica, node range per cluster
    DBuf<int32_t> cl_rep_, cl_off_, cl_nodes_, cid_, sel_;
    DBuf<double> cl_out_, cl_vir_, vel_part_;
    DBuf<float2> cl_cm_;
    DBuf<int32_t> loss_ids_;
    DBuf<float2> loss_t_;
};

// launchers (kernels.hip)
hipError_t launch_kick(int mode, bool gather, bool energy, int nblk, const KickArgs &a,
                       hipStream_t s);
hipError_t launch_sort(const ScanArgs &a, hipStream_t s);
hipError_t launch_forces(bool sigma, bool types, bool debug, bool list, bool energy, int kick,
                         int block, const ForceArgs &a, hipStream_t s);
hipError_t launch_list_build(bool sigma, int block, const ForceArgs &a, hipStream_t s);
size_t list_build_lds_bytes(int lds_cap, int cell_cap, int nbr_cap, int block, bool buffered);
int list_build_resident(bool sigma, int block, const ForceArgs &a, int device);
hipError_t launch_ctl_reset(Ctl *ctl, int nrep, uint32_t *dmax, int nslots, hipStream_t s);
size_t forces_lds_bytes(int lds_cap, bool types, bool debug);
hipError_t launch_total_force(int n, const float *fx, const float *fy, const float2 *cf,
                              const int32_t *goff, const int32_t *gent, float *ox, float *oy,
                              hipStream_t s);
hipError_t launch_observe(const ObsArgs &o, bool accumulate, hipStream_t s);
hipError_t launch_fill_int(int32_t *p, int32_t v, size_t n, hipStream_t s);
// dst[0, n) = x, dst[n, 2n) = y (graph-safe copy kernel)
hipError_t launch_copy2(float *dst, const float *x, const float *y, int n, hipStream_t s);
hipError_t launch_adj_forces(bool sigma, bool list, int block, const AdjArgs &A, hipStream_t s);
hipError_t launch_adj_kick(int nblk, const AdjKickArgs &a, hipStream_t s);
hipError_t launch_adj_loss(const LossArgs &a, hipStream_t s);
hipError_t launch_cluster_props(const ClusterArgs &a, hipStream_t s);
hipError_t launch_cluster_virial(bool sigma, bool types, int block, const ClusterVirialArgs &A,
                                 hipStream_t s);
hipError_t launch_vel_ke(const int32_t *ids, int n, const float *vx, const float *vy,
                         const float *mass, double *part, hipStream_t s);
hipError_t launch_vel_scale(const int32_t *ids, int n, float *vx, float *vy, float f, hipStream_t s);
hipError_t launch_reorder(const ReorderArgs &a, hipStream_t s);
hipError_t launch_decide(const DecideArgs &d, hipStream_t s);

}  // namespace fm
