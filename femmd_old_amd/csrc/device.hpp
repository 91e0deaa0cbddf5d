// device.hpp -- data layout shared by the host engine and the HIP kernels.
//
// All per-particle state is structure-of-arrays in HBM, in the engine's
// internal order (per replica: particles sorted by their cell at commit time,
// which keeps every cell's members close in memory).  A batch of replicas is
// stored back to back: replica r owns particles [p0, p1), cells
// [cell0, cell0 + ncx*ncy] (the last one always empty: it closes the
// replica's cell starts), elements and bonds of its own ranges.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fm {

// Largest tile row (cells) and its slot table: 3 rows of (W + 2) cells.
constexpr int kMaxTileW = 64;
constexpr int kMaxSlots = 3 * (kMaxTileW + 2);
// Cells per workgroup of the single-pass counting sort (one per thread).
constexpr int kSortCells = 256;
constexpr int kKickThreads = 256;
constexpr int kElemThreads = 256;
constexpr int kMaxPairTypes = 16;  // particle types of a pair table
constexpr int kMaxListCap = 512;  // neighbour-list capacity per particle
// Largest per-cell capacity: a row of 3 cells holds < 256 candidates, so the
// list build buffers an entry as its one-byte offset in that row.
constexpr int kMaxCellCap = 60;
static_assert(3 * kMaxCellCap < 256, "one-byte list-build offsets");
constexpr int kMaxListGroups = kMaxListCap / 4;
// A list slot's word: length in the low bits, kListImg when the particle's
// neighbours include a halo cell across the periodic boundary (its entries
// need the minimum image), the particle's index in its tile above bit 16.
constexpr int kListLen = 0x7fff, kListImg = 0x8000;

// Per-replica geometry and integration constants.
struct alignas(16) RepDesc {
    float Lx, Ly;      // periodic box
    float hLx, hLy;    // 0.5 * box (exact)
    float fncx, fncy;  // (float)ncx, (float)ncy for the cell key
    int32_t ncx, ncy;
    int32_t cell0;     // first global cell of this replica
    int32_t p0, p1;    // particle range
    float dt, hdt;     // step and half step
    int32_t pt;        // offset of this replica's pair table
    int32_t ntypes;
    float rc2_uniform; // ((rc*rc)*s)*s for single-type uniform-sigma systems
    float s2_uniform;  // sigma^2 for uniform sigma
    float s_uniform;   // sigma for uniform sigma
    float skin;        // neighbour-list skin (0: no list, cells every step)
    float rl2_uniform; // list radius^2, (rc*s + skin)^2 with a rounding margin
    int32_t ks0, nks;  // this replica's kick spans [ks0, ks0 + nks)
    float dslack;      // rounding allowance of the displacement test
    float im_uniform;  // 1/m shared by every particle of the replica (the
                       // kick then loads no inv_m); 0: masses differ
    int32_t pg;        // particles [pg, p1) have element or bond corners (or
                       // the replica has no pair forces): the node kick
                       // integrates them, the force pass the others
    float rc_list;     // list mode: the largest cutoff of the pair table (in
                       // sigma units); lists hold every pair within
                       // rc_list * sigma_ij + skin
    int32_t pad_[3];
};

// Per-step observables (femmd2D/src/properties.cpp:15-86, evaluated and
// accumulated after every step as Integrator::doStep does, integrator.cpp:18).
// Per replica, kObsStride doubles:
//   [0, 8)   the last formed step's totals {KE2 = sum m v^2, max v^2, U_lj,
//            W_lj, U_fea, W_fea, U_bond, W_bond}
//   [8]      steps accumulated since the last reset
//   [9, 17)  sums and [17, 25) sums of squares of the averaged quantities
//            {uKinetic, uPotential, uTotal, pressure, virial, feaVirial,
//            uBonded, uFea} (Prop::accum, femmd2D/src/properties.h:6-13)
// A step's totals are formed in two levels with a fixed order (no float
// atomics): kObsChunks chunks per replica, each the fixed-order sum of one
// slice of the per-workgroup partials (ke, lj, fea, bond), written by one
// workgroup of 256 threads; then the chunks in chunk order.
constexpr int kObsRaw = 8, kObsAvg = 8, kObsStride = 32, kObsChunks = 16;
constexpr int kObsRed = 4 * kObsRaw;  // LDS doubles of a level-1 reduction
struct ObsArgs {
    const RepDesc *reps;
    const int32_t *rng;        // 8 per replica: kick, tile, element, bond ranges
    const int32_t *nrng;       // 2 per replica: node-kick span range
    int32_t ke_kick;           // kinetic energy from ke_part over the kick spans
                               // (a query); else from ke_tile over the tiles
                               // and ke_node over the node-kick spans (steps)
    const double *ke_part;     // 2 per kick span: sum m v^2, max v^2
    const double *ke_tile;     // 2 per tile: the same, particles kicked by the pair pass
    const double *ke_node;     // 2 per node-kick span: the same, element / bond particles
    const double *lj_part;     // 2 per tile: U, W
    const double *fea_part;    // 2 per element span: U, W
    const double *bond_part;   // 2 per bond span: U, W
    double *chunk;             // kObsChunks * kObsRaw per replica (level 1)
    int32_t nchunk;            // level-1 chunks per replica (1..kObsChunks; of
                               // the replica's partials in order)
    double *obs;               // kObsStride per replica
    int32_t nrep;
};

// Displacement maxima of a kick: each workgroup takes an atomic max into one
// of kDispSlots slots of its replica (spread to limit contention, one slot
// per 128-byte line); the binning reads the slots, the force pass clears them.
constexpr int kDispSlots = 16;
constexpr int kDispStride = 32;

// Per-replica control words of the neighbour-list mode.
struct alignas(16) Ctl {
    int32_t flag;    // rebuild the list in this step
    float disp;      // displacement bound since the last build
    uint32_t nbuild; // builds so far
    int32_t pad;
};

// One pair-table entry: 48 eps, 4 eps, additive shift, cutoff in sigma units.
struct alignas(16) PairEntry {
    float e48, e4, shift, rc;
};

// Work tables: one workgroup per entry.
struct alignas(16) Tile {  // force pass, LJ role
    int32_t rep, cy, cx0;
    int32_t w_flags;  // W in bits 0..15, bit 16 = needs minimum image,
                      // bit 17 in x, bit 18 in y
};
struct alignas(16) Span {  // kick / element / bond workgroups
    int32_t rep, begin, end, pad;
};

// Kernel argument bundles (passed by value).
struct KickArgs {
    const RepDesc *reps;
    const Span *spans;
    float *x, *y, *vx, *vy;
    const float *fx, *fy;      // pair forces (internal order)
    const float2 *cf;          // element / bond corner forces, in gather order:
                               // particle p's corners at [goff[p], goff[p + 1])
    const int32_t *goff;       // gather CSR offsets (n + 1), may be null
    const float *mass, *inv_m;
    int32_t *count;            // per-cell counters (zero on entry)
    int32_t *tmp;              // cell * cap + slot -> particle
    int32_t cap;
    double *ke_part;           // 2 per workgroup: sum m v^2, max v^2
    uint32_t *err;
    Ctl *ctl;                  // list mode (gated binning): rebuild decision
    int32_t *any;              // list mode: set when some replica rebuilds
    uint32_t *dmax;            // list mode: per replica kDispSlots slots of the
                               // step's max squared displacement (float bits),
                               // kDispStride words apart
    const int32_t *spos;       // list mode: particle -> sorted position
    float2 *xys;               // list mode: positions in sorted order (out)
    float *rec;                // if set, the drift also writes x to rec[p] and
                               // y to rec[n + p] (trajectory recording)
    int32_t n;                 // particles (stride of rec)
    int32_t nspans;            // spans (grid-stride binning)
    const int32_t *fspan;      // gated batches: the rebuilding replicas' spans
    const int32_t *nfs;        // (decided by k_decide), and their count
    int32_t one_rep;           // gated binning of a single replica (decided first)
    int32_t *rblog;            // if set (single replica), the gated binning
                               // writes its rebuild decision there
    int32_t gather_lds;        // gather corner forces through LDS in
                               // coalesced chunks (uniform)
    ObsArgs obs;               // per-step observables
    int32_t obs_l1;            // the binning forms the previous step's level-1 chunks
};

// Rebuild decision of a batch (k_flags, k_decide): per replica the flag, and
// the compact lists of the rebuilding replicas' kick spans and tiles that
// the gated binning and list build walk instead of every span and tile.
// It also lists the sort blocks that hold a cell of a rebuilding replica (in
// increasing order), restarts the sort's per-step ticket counter and
// advances its step epoch.
struct DecideArgs {
    const RepDesc *reps;
    Ctl *ctl;
    const uint32_t *dmax;
    const int32_t *rng;        // 8 per replica: spans [0,1), tiles [2,3), ...
    int32_t nrep;
    int32_t *any;
    int32_t *fspan, *ftile;    // out: span / tile indices
    int32_t *fblock;           // out: sort blocks
    int32_t nsort;
    int32_t *nf;               // out: [0] spans, [1] tiles, [2] sort blocks,
                               // [3] sort ticket counter (reset), [4] step epoch
    int4 *rinfo;               // per replica (k_flags): spans, tiles to list
                               // (0 if it does not rebuild), first span, first tile
    uint8_t *mark;             // per sort block: holds a rebuilding replica's
                               // cell (set by k_flags, cleared by k_decide)
};

// Batches of more than this many replicas take the rebuild decision in one
// small kernel before the gated binning, which with the list build then
// walks compact lists of the rebuilding replicas' spans and tiles; smaller
// systems let every binning workgroup take it for the replicas it meets.
constexpr int kDecideMinReps = 8;

// Single-pass counting sort: workgroup b scans cells [b * kSortCells, ...),
// takes its carry-in from its predecessors' status words (decoupled
// look-back) and scatters the particles of its own cells.
struct ScanArgs {
    const RepDesc *reps;
    int32_t nrep;
    int32_t ncells;
    int32_t cap;
    int32_t nsort;             // sort workgroups per launch
    const int2 *brep;          // per sort block: first and last replica of its cells
    int32_t *count;            // per-cell counters, zeroed once read
    const int32_t *tmp;        // cell * cap + slot -> particle
    unsigned long long *status;  // nsort status words (value, state, launch)
    unsigned long long *ticket;  // workgroup tickets, never reset
    const int32_t *fblock;     // listed (gated batches): the blocks to sort,
    int32_t *nf;               // their count nf[2], tickets nf[3], epoch nf[4]
    unsigned long long *status2; // status words of the listed sort
    const Ctl *ctl;            // list mode gating (null: always run)
    const int32_t *any;
    int32_t *start;            // per cell: first sorted position (replica based)
    int32_t *perm;             // sorted position -> particle
    const int32_t *uid;        // particle -> its index in the caller's order
                               // (the in-cell rank key)
    const float *x, *y;
    const int32_t *grp;        // exclusion group per particle
    const float *sig;          // per-particle sigma (may be null)
    const int32_t *typ;        // per-particle type (may be null)
    float4 *sorted;            // sorted position -> {x, y, grp bits, sigma}
    int32_t *sorted_typ;       // sorted position -> type (may be null)
    int32_t *spos;             // list mode: particle -> sorted position
    float2 *xys;               // list mode: positions in sorted order
    float *sigs;               // list mode: sigma in sorted order (may be null)
    uint32_t *err;
    ObsArgs obs;               // per-step observables
    int32_t obs_l1;            // (cell walk) form the previous step's level-1 chunks
};

struct ForceArgs {
    const RepDesc *reps;
    const PairEntry *ptab;
    // pair role
    const Tile *tiles;
    int32_t ntiles;
    int32_t lds_cap;           // staged particles that fit in LDS
    const int32_t *start;
    const int32_t *perm;
    const float4 *sorted;
    const int32_t *sorted_typ;
    float *fx, *fy;            // pair forces out, internal order
    double *lj_part;           // 2 per tile: U, W
    // element role
    const Span *espans;
    int32_t nespans;
    const int32_t *t0, *t1, *t2;
    const float4 *dmbm;        // per element: Dm {00,01,10,11}, Bm {00,01,10,11}
    const float2 *alm;         // A0 * lambda, A0 * mu
    int32_t model;             // 0 NH, 1 SVK (legacy is SVK with doubled moduli)
    const float *x, *y;        // current positions, internal order
    float2 *cf;                // corner forces out, in gather order
    const int32_t *gpos;       // corner slot (3e + k, then bonds 2l + k) ->
                               // its entry in the per-particle gather order
    double *fea_part;          // 2 per element workgroup: U, W
    // bond role
    const Span *bspans;
    int32_t nbspans;
    const int32_t *bi, *bj;
    const float *bk, *br0;
    int32_t cf_bond0;          // first corner slot of the bonds
    double *bond_part;         // 2 per bond workgroup: U, W
    // neighbour-list mode (pair role over the tiles of the list build)
    int32_t nbr_cap;           // list capacity per particle (multiple of 8)
    void *nbr;                 // neighbours of the particle at sorted position
                               // p as uint16 byte offsets of float2 slots of
                               // its tile's LDS staging, four per uint64
                               // group, group g at element g * n + p
    int32_t *nnum;             // per list slot q: length | kListImg when its
                               // entries need the minimum image | (k << 16),
                               // k the tile index of the particle whose list
                               // it is
    int32_t lbuf;              // list build: stage entries in LDS (coalesced
                               // column writes) when set
    int32_t gated_grid;        // list build grid when gated (resident workgroups)
    int32_t cell_cap;          // slots per cell (cell counts are capped at it)
    const float2 *xys;         // list mode: positions in sorted order
    const float *sigs;         // list mode: sigma in sorted order (sigma path)
    const Ctl *ctl;            // list build gating (null: always build)
    const int32_t *ftile;      // gated batches: the rebuilding replicas' tiles
    const int32_t *nft;        // and their count
    int32_t *any;              // reset by the force pass after the rebuild
    uint32_t *dmax;            // displacement slots, cleared by the force pass
    Ctl *ctl_w;                // displacement bounds, updated by the force pass
    int32_t nrep;
    int32_t n;                 // particles (stride of nbr)
    // diagnostics (pair-set dump); null in production
    const int32_t *uid;        // internal -> caller id
    int32_t *dbg_cnt;
    long long *dbg_s1, *dbg_s2;
    uint32_t *err;
    ObsArgs obs;               // per-step observables
    int32_t obs_l2;            // combine the previous step's chunks and accumulate
    // The kick of the pass (kick modes of k_forces): each tile integrates its
    // centre particles that have no element or bond corners (the others are
    // left to the node kick, which needs their gathered corner forces):
    // second half-kick of the previous step with the kinetic energy and, in
    // the step mode, first half-kick, drift and wrap.
    float *xw, *yw;            // internal positions written (the same arrays as x, y)
    float *vx, *vy;
    const float *inv_m, *mass;
    const int32_t *goff;       // gather CSR offsets (null: no particle has corners)
    float2 *xys_out;           // list mode: the drifted positions, sorted order
                               // (the other buffer of the pair: tiles still
                               // stage their neighbours from xys)
    uint32_t *dmax_out;        // list mode: slots of this drift's maximum squared
                               // displacement (dmax: the previous drift's)
    float *rec;                // trajectory recording (as KickArgs)
    double *ke_tile;           // 2 per tile: sum m v^2, max v^2 of the kicked particles
    int32_t *count, *tmp;      // cell walk: binning of the drifted positions
};

// Adjoint pass: J(x) w (J = df/dx, the negative Hessian of the potential)
// and the parameter terms (df/dtheta)^T w at one configuration.
struct AdjArgs {
    ForceArgs f;          // forward inputs (binning, tiles, elements, bonds)
    const float2 *w;      // w per particle, internal order
    float *jx, *jy;       // pair part of J w, internal order
    float2 *cfa;          // element / bond corner parts of J w
    const float *eE;      // Young's modulus per element (forces are linear in E)
    double *dE;           // per element: += w . f_e / E_e
    double *pacc;         // per tile: += sum_i w_i . f_i / eps, += w . df/dsigma
};

// Backward half-kicks / drift: xb += J w (accumulate), vb += dt xb (drift),
// w = c inv_m vb (make_w); one workgroup per kick span.
struct AdjKickArgs {
    const RepDesc *reps;
    const Span *spans;
    float *xbx, *xby, *vbx, *vby;
    const float *jx, *jy;
    const float2 *cfa;
    const int32_t *goff, *gent;
    const float *inv_m;
    float2 *w;
    float c;
    int32_t accumulate, drift, make_w;
};

// Loss = (1/K) sum_k |x_k - t_k|^2 over K particles (minimum image);
// seeds xb = dL/dx (other particles keep xb = 0).
struct LossArgs {
    const RepDesc *reps;
    int32_t n;
    const int32_t *ids;   // internal ids of the loss particles
    const float2 *target;
    const float *x, *y;
    float *xbx, *xby;
    float scale;          // 2 / K
    double *part;         // per workgroup partial sum of |x - t|^2
};

// Elastomer cluster observables (one workgroup per cluster).
struct ClusterArgs {
    const RepDesc *reps;
    int32_t ncl;
    const int32_t *crep;       // replica of each cluster
    const int32_t *off;        // cluster -> first entry of nodes (ncl + 1)
    const int32_t *nodes;      // internal ids of the clusters' nodes
    const float *x, *y, *vx, *vy, *mass;
    double *out;               // kClusterOut per cluster
    float2 *cm;                // centre of mass per cluster (the virial pass)
    float *ux, *uy;            // unfolded positions, internal order (may be null)
};
// per cluster: mass, centre of mass x, y, its velocity x, y, centroid x, y
constexpr int kClusterOut = 7;

// Cluster virial: sum over interacting pairs (i, j) of particles of
// clusters of 1/2 f_ij . (R_c(i) - R_c(j)) (minimum image), per tile.
struct ClusterVirialArgs {
    ForceArgs f;
    const int32_t *cid;        // cluster of each particle (-1: not in one)
    const float2 *cm;
    double *out;               // per tile
};

// Re-ordering of the internal order (particles with no element, bond or
// cluster -- "movable" -- re-assigned to the movable internal slots in the
// order of the last counting sort).  Sorted slots are processed in chunks of
// kReorderChunk (256 lanes x 4 consecutive slots).
constexpr int kReorderChunk = 1024;
constexpr int kReorderArrays = 12;
struct ReorderArgs {
    int n;                     // particles (all replicas)
    int nblk;                  // chunks of kReorderChunk sorted slots
    const uint8_t *mov;        // per internal id: movable
    const int32_t *movids;     // the movable internal ids, ascending
    int nmov;                  // their count
    int32_t *perm;             // sorted slot -> internal id (rewritten)
    int32_t *spos;             // internal id -> sorted slot (list mode; rewritten)
    int32_t *bsum;             // per chunk: movable count, then its exclusive prefix
    int32_t *oldof;            // new internal id -> old internal id
    int narr;                  // per-particle 4-byte arrays to permute
    uint32_t *arr[kReorderArrays];
    uint32_t *tmp;             // narr * n words
};

}  // namespace fm
