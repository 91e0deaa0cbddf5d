// kernels.hip -- CDNA4 (gfx950) kernels of the femmd2D step.
//
// A run of steps (one HIP graph node each):
//   k_kick       first half-kick, drift and wrap of every particle from the
//                stored forces (the start of a run; cell walk: binning)
// then per step (femmd2D/src/integrator.cpp:8-29):
//   k_bin        list mode: rebuild decision (displacement bound > skin/2)
//                and binning of the replicas that rebuild (gated)
//   k_sort       single-pass counting sort of the binning (decoupled
//                look-back): cell starts, sorted order, sorted copies
//   k_list_build list mode: Verlet lists of the rebuilding replicas' tiles
//   k_forces     one launch with three workgroup roles: LJ cell tiles
//                (LDS-staged neighbours, lists or the 3x3 cell walk), CST
//                elements and harmonic bonds (corner forces into a buffer in
//                gather order, no atomics); each tile then gives its centre
//                particles that have no element or bond corners their second
//                half-kick, and the first half-kick, drift and wrap of the
//                next step (cell walk: binning of the new positions)
//   k_kick       the node kick: the same for the particles with corners,
//                their corner forces gathered through a per-particle CSR in
//                a fixed order
// and at the end of a run the second half-kick only, then the observables.
// Energies, virials and the kinetic energy are per-workgroup FP64 partials
// summed in a fixed order (k_bin / k_sort level 1, k_forces level 2).
//
// The file is compiled with -ffp-contract=off: every FP32 expression that
// must agree bit for bit with the reference's real=float arithmetic (wrap,
// cell key, minimum image, cutoff predicate) is evaluated in its source
// order; fused multiply-adds are written explicitly where wanted.
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "device.hpp"

namespace fm {

// Grids of the gated list-mode kernels: a few workgroups per CU loop over
// the work, so a skipped step costs little (the binning workgroups also
// each make the rebuild decision, so that grid is the smallest).
constexpr int kGatedBinGrid = 1024;
// Workgroups of the batch rebuild decision (each writes a share of the
// compact span and tile lists).
constexpr int kDecideGrid = 16;
constexpr int kGatedGrid = 2048;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// femmd2D/src/def.h:112 pbcCalc: r - floor(r / L) * L, FP32, IEEE division.
// For 0 < r < L the expression is r itself: r / L rounds to at most
// 1 - 2^-24 (the largest float below L over L is at most that), so its
// floor is 0; the division is then skipped (0 and -0 take the full path,
// which maps -0 to +0).
__device__ __forceinline__ float wrap1(float r, float L) {
    if (r > 0.0f && r < L) return r;
    float q = r / L;
    float f = floorf(q);
    float p = f * L;
    return r - p;
}

// femmd2D/src/nnCuda.cu:9-17 cellId: ix = (int)(x * ncx / Lx), ix == ncx
// clamped to ncx - 1.  The extra clamp into [0, ncx) only matters for
// non-finite input, which is flagged separately.
__device__ __forceinline__ int cell_key(float x, float y, const RepDesc &R) {
    float tx = x * R.fncx;
    tx = tx / R.Lx;
    int ix = (int)tx;
    if (ix == R.ncx) ix -= 1;
    float ty = y * R.fncy;
    ty = ty / R.Ly;
    int iy = (int)ty;
    if (iy == R.ncy) iy -= 1;
    ix = min(max(ix, 0), R.ncx - 1);
    iy = min(max(iy, 0), R.ncy - 1);
    return ix + iy * R.ncx;
}

// p - q + k L for k in {-1, 0, 1}, with the coordinate near L moved by L
// first: that move is exact (Sterbenz), so only one rounding remains at the
// scale of the small difference.  Used for force values across the
// periodic boundary; the image decision itself is taken by the caller.
__device__ __forceinline__ float shifted_diff(float p, float q, int k, float L) {
    if (k > 0) return p - (q - L);
    if (k < 0) return (p - L) - q;
    return p - q;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_sum_int(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
// Inclusive prefix sum over the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// Deterministic block reduction of two doubles (fixed wave order); the
// result is valid in thread 0.  red must hold 2 * (blockDim / 64) doubles.
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red) {
    a = wave_sum(a);
    b = wave_sum(b);
    int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) { red[2 * w] = a; red[2 * w + 1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int k = 0; k < nw; ++k) { sa += red[2 * k]; sb += red[2 * k + 1]; }
        a = sa;
        b = sb;
    }
}

// ---------------------------------------------------------------------------
// Per-step observables (femmd2D/src/properties.cpp:15-86).  Level 1: chunk c
// of replica r is the fixed-order sum of slice c of each of its partial
// ranges (kick spans, tiles, element spans, bond spans), formed by one
// workgroup of exactly 256 threads (the binning, sort and observe kernels all
// run 256), so its bits do not depend on which kernel forms it.  Level 2:
// one thread sums the replica's chunks in chunk order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void obs_slice(int b, int e, int c, int nc, int &lo, int &hi) {
    const long long len = e - b;
    lo = b + (int)(len * c / nc);
    hi = b + (int)(len * (c + 1) / nc);
}

// Level 1 for chunk c of replica r: out[0, 8) of the chunk (see kObsRaw).
// Called by every thread of a 256-thread workgroup (uniformly).
__device__ void obs_chunk(const ObsArgs &o, int r, int c, double *red) {
    const int32_t *g = o.rng + 8 * r;
    const int tid = threadIdx.x;
    double v[kObsRaw] = {0, 0, 0, 0, 0, 0, 0, 0};
    int lo, hi;
    if (o.ke_kick) {  // a query: the partials of k_ke over the kick spans
        obs_slice(g[0], g[1], c, o.nchunk, lo, hi);
        for (int k = lo + tid; k < hi; k += 256) {
            v[0] += o.ke_part[2 * k];
            v[1] = fmax(v[1], o.ke_part[2 * k + 1]);
        }
    } else {  // a step: the particles the node kick integrated
        obs_slice(o.nrng[2 * r], o.nrng[2 * r + 1], c, o.nchunk, lo, hi);
        for (int k = lo + tid; k < hi; k += 256) {
            v[0] += o.ke_node[2 * k];
            v[1] = fmax(v[1], o.ke_node[2 * k + 1]);
        }
    }
    obs_slice(g[2], g[3], c, o.nchunk, lo, hi);
    for (int k = lo + tid; k < hi; k += 256) {
        v[2] += o.lj_part[2 * k];
        v[3] += o.lj_part[2 * k + 1];
        if (!o.ke_kick) {  // and those the pair pass integrated, per tile
            v[0] += o.ke_tile[2 * k];
            v[1] = fmax(v[1], o.ke_tile[2 * k + 1]);
        }
    }
    obs_slice(g[4], g[5], c, o.nchunk, lo, hi);
    for (int k = lo + tid; k < hi; k += 256) {
        v[4] += o.fea_part[2 * k];
        v[5] += o.fea_part[2 * k + 1];
    }
    obs_slice(g[6], g[7], c, o.nchunk, lo, hi);
    for (int k = lo + tid; k < hi; k += 256) {
        v[6] += o.bond_part[2 * k];
        v[7] += o.bond_part[2 * k + 1];
    }
    // the chunk's seven sums and its max v^2: each wavefront's in a fixed
    // butterfly, then the four wavefronts in order (one exchange through
    // LDS; red holds kObsRed doubles)
    const double m = v[1];
#pragma unroll
    for (int k = 0; k < kObsRaw; ++k) v[k] = (k == 1) ? wave_max_d(m) : wave_sum(v[k]);
    __syncthreads();  // earlier readers of red are done
    if ((tid & 63) == 0)
        for (int k = 0; k < kObsRaw; ++k) red[(tid >> 6) * kObsRaw + k] = v[k];
    __syncthreads();
    if (tid == 0) {
        double *out = o.chunk + ((size_t)r * kObsChunks + c) * kObsRaw;
        for (int k = 0; k < kObsRaw; ++k) {
            double t = red[k];
            for (int w = 1; w < 4; ++w)
                t = (k == 1) ? fmax(t, red[w * kObsRaw + k]) : t + red[w * kObsRaw + k];
            out[k] = t;
        }
    }
    __syncthreads();  // red is free again
}

// The level-1 jobs (replica, chunk) of one step, spread over a kernel's
// workgroups (256 threads each; uniform per workgroup).
__device__ void obs_level1(const ObsArgs &o, double *red) {
    const int njob = o.nrep * o.nchunk;
    for (int q = blockIdx.x; q < njob; q += gridDim.x) obs_chunk(o, q / o.nchunk, q % o.nchunk, red);
}

// Level 2 for replica r (one thread): the step's totals from its chunks in
// chunk order, the averaged quantities of femmd2D/src/properties.cpp:15-36
// and, when accumulate is set, Prop::accum of each (properties.h:10).
__device__ void obs_finish(const ObsArgs &o, int r, bool accumulate) {
    const double *ch = o.chunk + (size_t)r * kObsChunks * kObsRaw;
    double t[kObsRaw] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < o.nchunk; ++c) {
        const double *q = ch + c * kObsRaw;
        for (int k = 0; k < kObsRaw; ++k) t[k] = (k == 1) ? fmax(t[1], q[1]) : t[k] + q[k];
    }
    double *ob = o.obs + (size_t)r * kObsStride;
    for (int k = 0; k < kObsRaw; ++k) ob[k] = t[k];
    if (!accumulate) return;
    const RepDesc &R = o.reps[r];
    const double n = (double)(R.p1 - R.p0), V = (double)R.Lx * (double)R.Ly;
    const double uk = 0.5 * t[0] / n;
    const double up = (t[2] + t[4] + t[6]) / n;
    const double a[kObsAvg] = {uk, up, uk + up, 0.5 * t[0] / V + 0.5 * (t[3] + t[5]) / V,
                               t[3], t[5], t[6] / n, t[4] / n};
    ob[kObsRaw] += 1.0;
    for (int k = 0; k < kObsAvg; ++k) {
        ob[kObsRaw + 1 + k] += a[k];
        ob[kObsRaw + 1 + kObsAvg + k] += a[k] * a[k];
    }
}

// Bin one particle per lane into its cell's slot list (the reference's
// atomic cell fill, femmd2D/src/nnCuda.cu:69-82, re-designed): runs of
// equal cells among consecutive lanes -- the common case, since particles
// are stored near their cell -- take one integer atomicAdd for the whole
// run and hand out slots by lane offset.  Slot order within a cell is not
// deterministic; the scatter below orders each cell by particle index.
// All 64 lanes of the wavefront must call this (ballot / shuffles).
__device__ __forceinline__ void bin_wave_to(int32_t *count, int32_t *tmp, int cap, uint32_t *err,
                                            bool active, int p, int c) {
    const int lane = threadIdx.x & 63;
    const int ck = active ? c : -1 - lane;  // inactive lanes never join a run
    const int prev = __shfl_up(ck, 1);
    const bool head = (lane == 0) || (ck != prev);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int h = 63 - __clzll((long long)(heads & upto));
    const unsigned long long after = heads & ~upto;
    const int next = after ? (__ffsll((unsigned long long)after) - 1) : 64;
    int base = 0;
    if (head && active) base = atomicAdd(&count[c], next - lane);
    base = __shfl(base, h);
    if (active) {
        const int slot = base + (lane - h);
        if (slot < cap) tmp[(size_t)c * cap + slot] = p;
        else atomicOr(err, 1u);  // FM_DEV_CELL_OVERFLOW (the particle is dropped)
    }
}
__device__ __forceinline__ void bin_wave(const KickArgs &a, bool active, int p, int c) {
    bin_wave_to(a.count, a.tmp, a.cap, a.err, active, p, c);
}

// The velocity-Verlet kick of one particle (femmd2D/src/integrator.cpp:
// 100-135, integratorCuda.cu:110-206), shared by k_kick and the pair pass's
// kick so both give the same bits.  ax, ay = f * inv_m.
//   kKick2: v += dt/2 a (second half of the previous step); with kEnergy
//           m |v|^2 (m from 1/m: one rounding from the stored mass; a pinned
//           particle has inv_m = 0 and zero velocity) and |v|^2
//   kDrift: v += dt/2 a, x += dt v, wrap (non-finite coordinates flagged
//           and reset); dxp, dyp: the step's displacement
template <bool kKick2, bool kDrift, bool kEnergy>
__device__ __forceinline__ void kick_particle(const RepDesc &R, float ax, float ay, float im,
                                              float &vx, float &vy, float &x, float &y,
                                              double &ke, float &vv, float &dxp, float &dyp,
                                              uint32_t *err) {
    if (kKick2) {
        vx = vx + R.hdt * ax;
        vy = vy + R.hdt * ay;
        if (kEnergy) {
            vv = vx * vx + vy * vy;
            const float m = im > 0.0f ? 1.0f / im : 0.0f;
            ke = (double)(m * vv);
        }
    }
    if (kDrift) {
        vx = vx + R.hdt * ax;
        vy = vy + R.hdt * ay;
        dxp = R.dt * vx;
        dyp = R.dt * vy;
        x = x + dxp;
        y = y + dyp;
        x = wrap1(x, R.Lx);
        y = wrap1(y, R.Ly);
        if (!(isfinite(x) && isfinite(y))) {
            atomicOr(err, 8u);  // FM_DEV_NONFINITE
            x = 0.0f;
            y = 0.0f;
        }
    }
}

// ---------------------------------------------------------------------------
// k_kick: integrator + binning (femmd2D/src/integrator.cpp:100-135,
// femmd2D/src/nnCuda.cu:69-82 re-designed).  One workgroup covers one span
// of one replica: particles [begin, end) (all particles: the first
// half-kick of a run; the replica's [pg, p1): the node kick, of the
// particles with element or bond corners, which the force pass does not
// kick).
//   kKick2: second half-kick of the previous step (kinetic-energy and max
//           v^2 partials of the workgroup with kEnergy)
//   kDrift: first half-kick, drift, wrap; bin (cell walk every step) or
//           mirror x, y into sorted order and take the step's largest
//           displacement (neighbour lists)
// a = (f_pair + sum of the particle's corner forces) * inv_m; the corner
// forces of particle p are the consecutive entries [goff[p], goff[p + 1]).
// ---------------------------------------------------------------------------
// The corner forces of a span's particles are one contiguous run of the
// gather order, [goff[begin], goff[end]): the workgroup reads it in
// coalesced chunks through LDS and each lane adds its own particle's
// entries in order (the same sums in the same order as reading its run
// directly, which leaves each lane a strided walk of its own).  It pays
// where the node kick is throughput-bound (many workgroups); with few, the
// two barriers per chunk lengthen each workgroup's path and the direct
// walk is faster.
constexpr int kGatherChunk = 2048;
__device__ __forceinline__ void gather_corners(const KickArgs &a, const Span &sp, int p,
                                               bool active, float &fx, float &fy) {
    __shared__ float2 s_cf[kGatherChunk];
    const int E0 = a.goff[sp.begin], E1 = a.goff[sp.end];
    const int e0 = active ? a.goff[p] : 0, e1 = active ? a.goff[p + 1] : 0;
    for (int c = E0; c < E1; c += kGatherChunk) {  // uniform per workgroup
        const int n = min(kGatherChunk, E1 - c);
        __syncthreads();  // the previous chunk's readers are done
        for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) s_cf[i] = a.cf[c + i];
        __syncthreads();
        const int lo = max(e0, c), hi = min(e1, c + n);
        for (int e = lo; e < hi; ++e) {
            const float2 v = s_cf[e - c];
            fx += v.x;
            fy += v.y;
        }
    }
}

template <bool kKick2, bool kDrift, bool kGather, bool kList, bool kEnergy>
__global__ __launch_bounds__(kKickThreads) void k_kick(KickArgs a) {
    __shared__ double red[2 * (kKickThreads / 64)];
    __shared__ float redm[kKickThreads / 64];
    const Span sp = a.spans[blockIdx.x];
    const RepDesc R = a.reps[sp.rep];
    const int p = sp.begin + (int)threadIdx.x;
    const bool active = p < sp.end;
    double ke = 0.0;
    float vvmax = 0.0f, d2 = 0.0f;
    int cell = 0;
    float fx = 0.0f, fy = 0.0f;
    if (active) {
        fx = a.fx[p];
        fy = a.fy[p];
    }
    if (kGather) {
        if (a.gather_lds) {  // uniform: set for node kicks of many workgroups
            gather_corners(a, sp, p, active, fx, fy);
        } else if (active) {
            for (int e = a.goff[p]; e < a.goff[p + 1]; ++e) {
                const float2 c = a.cf[e];
                fx += c.x;
                fy += c.y;
            }
        }
    }
    if (active) {
        float im;  // (the same value either way: no load when uniform)
        if (R.im_uniform > 0.0f) im = R.im_uniform;
        else im = a.inv_m[p];
        const float ax = fx * im, ay = fy * im;
        float vx = a.vx[p], vy = a.vy[p];
        float x = 0.0f, y = 0.0f, dxp = 0.0f, dyp = 0.0f;
        if (kDrift) {
            x = a.x[p];
            y = a.y[p];
        }
        kick_particle<kKick2, kDrift, kEnergy>(R, ax, ay, im, vx, vy, x, y, ke, vvmax, dxp, dyp,
                                               a.err);
        if (kDrift) {
            a.x[p] = x;
            a.y[p] = y;
            if (a.rec) {
                a.rec[p] = x;
                a.rec[a.n + p] = y;
            }
            if (!kList) cell = R.cell0 + cell_key(x, y, R);
            if (kList) {
                a.xys[a.spos[p]] = make_float2(x, y);
                d2 = dxp * dxp + dyp * dyp;  // this step's displacement
            }
        }
        a.vx[p] = vx;
        a.vy[p] = vy;
    }
    if (kDrift && !kList) bin_wave(a, active, p, cell);
    if (kDrift && kList) {
        // the workgroup's largest displacement, one atomic max into a slot of
        // the replica (relaxed: the binning reads the slots after this
        // kernel has completed)
        __shared__ float redd[kKickThreads / 64];
        const float m = wave_max(d2);
        if ((threadIdx.x & 63) == 0) redd[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            float mm = redd[0];
            for (int k = 1; k < kKickThreads / 64; ++k) mm = fmaxf(mm, redd[k]);
            const int slot = blockIdx.x % kDispSlots;
            atomicMax(a.dmax + ((size_t)sp.rep * kDispSlots + slot) * kDispStride,
                      __float_as_uint(mm));
        }
    }
    if (kKick2 && kEnergy) {
        double vm = 0.0;
        // max via a float wave reduction (exact, order independent)
        float m = wave_max(vvmax);
        int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) redm[w] = m;
        block_sum2(ke, vm, red);
        if (threadIdx.x == 0) {
            float mm = 0.0f;
            for (int k = 0; k < kKickThreads / 64; ++k) mm = fmaxf(mm, redm[k]);
            a.ke_part[2 * blockIdx.x] = ke;
            a.ke_part[2 * blockIdx.x + 1] = (double)mm;
        }
    }
}

// Kinetic-energy partials of the current velocities (no kick).
__global__ __launch_bounds__(kKickThreads) void k_ke(KickArgs a) {
    __shared__ double red[2 * (kKickThreads / 64)];
    __shared__ float redm[kKickThreads / 64];
    const Span sp = a.spans[blockIdx.x];
    const int p = sp.begin + (int)threadIdx.x;
    double ke = 0.0, vm = 0.0;
    float vvmax = 0.0f;
    if (p < sp.end) {
        const float vx = a.vx[p], vy = a.vy[p];
        const float vv = vx * vx + vy * vy;
        ke = (double)(a.mass[p] * vv);
        vvmax = vv;
    }
    float m = wave_max(vvmax);
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = m;
    block_sum2(ke, vm, red);
    if (threadIdx.x == 0) {
        float mm = 0.0f;
        for (int k = 0; k < kKickThreads / 64; ++k) mm = fmaxf(mm, redm[k]);
        a.ke_part[2 * blockIdx.x] = ke;
        a.ke_part[2 * blockIdx.x + 1] = (double)mm;
    }
}

// Binning of the current positions without moving them (used after
// fm_set_state / fm_commit and by fm_compute_forces).
// List mode displacement bound of replica r after this step: the running
// sum of per-step maximum displacements since the last build (the
// reference's dispHi, femmd2D/src/md3dsystem.cpp:177, here summed from the
// drift actually applied, before the decision) plus a rounding allowance
// for the stored positions.  Called by one wavefront; every caller reads
// the same slots and bound, so all get the same value.
__device__ __forceinline__ float list_disp(const uint32_t *dmax, const Ctl *ctl, int r,
                                           const RepDesc &R) {
    const int lane = threadIdx.x & 63;
    const uint32_t v = lane < kDispSlots ? dmax[((size_t)r * kDispSlots + lane) * kDispStride] : 0u;
    const float mm = wave_max(__uint_as_float(v));
    return ctl[r].disp + sqrtf(mm) * 1.0001f + R.dslack;
}

// List mode rebuild decision for replica R (femmd2D/src/nn.cpp:114-134:
// rebuild when the bound exceeds skin / 2), made by every wavefront that
// bins one of its spans; the workgroup holding the replica's first span
// records it.  The force pass of the step then folds this step into the
// bound (or resets it after a rebuild).
__device__ int list_decide(const KickArgs &a, const Span &sp, const RepDesc &R) {
    const float d = list_disp(a.dmax, a.ctl, sp.rep, R);
    const int flag = d > 0.5f * R.skin;
    if (threadIdx.x == 0 && sp.begin == R.p0) {
        Ctl *c = a.ctl + sp.rep;
        c->flag = flag;
        if (flag) {
            c->nbuild += 1u;
            *a.any = 1;  // every writer stores the same value
        }
    }
    return flag;
}

// The rebuild decision of a batch: each lane takes a replica (its bound as
// list_disp forms it), then the rebuilding replicas' kick spans and tiles
// are written as two compact lists (inclusive scans
// over the replicas, in replica order; each entry found by a binary search
// of those scans in LDS), which the gated binning and list build walk
// instead of every span and tile.  The sort blocks holding a cell of a
// rebuilding replica are marked in LDS and listed in increasing order.
__device__ __forceinline__ int block_scan_incl(int v, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int is = wave_incl_scan(v);
    if (lane == 63) wsum[wv] = is;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < wv; ++k) base += wsum[k];
    total = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) total += wsum[k];
    __syncthreads();
    return base + is;
}

// One wavefront per replica: the bound as list_disp forms it, the flag, and
// what the replica adds to the work lists.
__global__ __launch_bounds__(256) void k_flags(DecideArgs d) {
    const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (r >= d.nrep) return;  // uniform per wavefront
    const RepDesc &R = d.reps[r];
    const float bound = list_disp(d.dmax, d.ctl, r, R);
    if ((threadIdx.x & 63) == 0) {
        const int flag = bound > 0.5f * R.skin;
        d.ctl[r].flag = flag;
        int4 info = make_int4(0, 0, d.rng[8 * r], d.rng[8 * r + 2]);
        if (flag) {
            d.ctl[r].nbuild += 1u;
            *d.any = 1;  // every writer stores the same value
            info.x = d.rng[8 * r + 1] - d.rng[8 * r];
            info.y = d.rng[8 * r + 3] - d.rng[8 * r + 2];
            if (d.fblock) {
                // the replica's cells, with its closing empty cell
                const int b0 = R.cell0 / kSortCells;
                const int b1 = (R.cell0 + R.ncx * R.ncy) / kSortCells;
                for (int b = b0; b <= b1; ++b) d.mark[b] = 1;
            }
        }
        d.rinfo[r] = info;
    }
}

// The work lists of a batch: the rebuilding replicas' kick spans and tiles
// (inclusive scans over the replicas, in replica order; each entry found by
// a binary search of those scans in LDS), written by kDecideGrid workgroups,
// and the marked sort blocks in increasing order (by the first workgroup,
// which clears the marks for the next step).
__global__ __launch_bounds__(1024) void k_decide(DecideArgs d) {
    __shared__ int inS[1024], inT[1024], fS[1024], fT[1024], wsum[16];
    const int t = threadIdx.x;
    // every workgroup forms the same scans and writes its share of the
    // span and tile lists; workgroup 0 alone lists the sort blocks (it
    // clears their marks) and writes the counts
    const int g = blockIdx.x, G = gridDim.x;
    int baseS = 0, baseT = 0;
    for (int c0 = 0; c0 < d.nrep; c0 += 1024) {
        const int r = c0 + t;
        const int4 info = r < d.nrep ? d.rinfo[r] : make_int4(0, 0, 0, 0);
        int totS, totT;
        inS[t] = block_scan_incl(info.x, wsum, totS);
        inT[t] = block_scan_incl(info.y, wsum, totT);
        fS[t] = info.z;
        fT[t] = info.w;
        __syncthreads();
        // entry v of this chunk's list belongs to the first replica whose
        // inclusive count exceeds v
        auto fill = [&](const int *in, const int *first, int tot, int base, int32_t *out) {
            for (int v = g * 1024 + t; v < tot; v += G * 1024) {
                int lo = 0, hi = 1023;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (in[mid] > v) hi = mid; else lo = mid + 1;
                }
                const int ex = lo > 0 ? in[lo - 1] : 0;
                out[base + v] = first[lo] + (v - ex);
            }
        };
        fill(inS, fS, totS, baseS, d.fspan);
        fill(inT, fT, totT, baseT, d.ftile);
        baseS += totS;
        baseT += totT;
        __syncthreads();
    }
    if (g != 0) return;  // (uniform per workgroup)
    int nfb = 0;
    if (d.fblock) {
        for (int c0 = 0; c0 < d.nsort; c0 += 1024) {
            const int b = c0 + t;
            const int m = b < d.nsort ? d.mark[b] : 0;
            int tot;
            const int ib = block_scan_incl(m, wsum, tot);
            if (m) {
                d.fblock[nfb + ib - 1] = b;
                d.mark[b] = 0;
            }
            nfb += tot;
        }
    }
    if (t == 0) {
        d.nf[0] = baseS;
        d.nf[1] = baseT;
        d.nf[2] = nfb;
        d.nf[3] = 0;                  // the listed sort's tickets
        d.nf[4] = d.nf[4] + 1;        // and its step epoch
    }
}

__global__ __launch_bounds__(kKickThreads) void k_bin(KickArgs a) {
  __shared__ double ored[kObsRed];
  // the previous step's observables, level 1 (before any early exit)
  if (a.obs_l1) obs_level1(a.obs, ored);
  int last_rep = -1, flag = 1;
  if (a.ctl && !a.fspan && a.one_rep) {
      // a single replica: the decision comes first, from the replica's
      // descriptor and displacement slots loaded together (no span read
      // before them), so that a step without a rebuild leaves at once
      const RepDesc &R0 = a.reps[0];
      flag = list_disp(a.dmax, a.ctl, 0, R0) > 0.5f * R0.skin;
      if (blockIdx.x == 0 && threadIdx.x == 0) {  // as list_decide records it
          a.ctl->flag = flag;
          if (flag) {
              a.ctl->nbuild += 1u;
              *a.any = 1;
          }
          if (a.rblog) *a.rblog = flag;  // (adjoint recording)
      }
      if (!flag) return;
      last_rep = 0;
  }
  // decided (batches): only the rebuilding replicas' spans, listed
  const int nsp = a.fspan ? *a.nfs : a.nspans;
  for (int b = blockIdx.x; b < nsp; b += gridDim.x) {
    const Span sp = a.spans[a.fspan ? a.fspan[b] : b];
    const RepDesc R = a.reps[sp.rep];
    if (a.ctl && !a.fspan) {  // list mode: bin only the replicas that rebuild this step
        if (sp.rep != last_rep) {  // uniform per workgroup
            flag = list_decide(a, sp, R);
            last_rep = sp.rep;
        }
        if (!flag) {
            // on to this workgroup's first span of the next replica
            const int e = R.ks0 + R.nks, g = (int)gridDim.x;
            if (e > b + g) b += ((e - b - 1) / g) * g;
            continue;
        }
    }
    const int p = sp.begin + (int)threadIdx.x;
    const bool active = p < sp.end;
    int c = 0;
    if (active) {
        float x = a.x[p], y = a.y[p];
        if (!(isfinite(x) && isfinite(y))) {
            atomicOr(a.err, 8u);
            x = 0.0f;
            y = 0.0f;
        }
        c = R.cell0 + cell_key(x, y, R);
    }
    bin_wave(a, active, p, c);
  }
}


// ---------------------------------------------------------------------------
// Counting sort (femmd2D/src/nn.cpp:8-20 cell fill order, re-designed as a
// single pass).  Workgroup b (taken in launch order from a ticket counter)
// owns cells [b * kSortCells, (b + 1) * kSortCells): it reads and clears
// their capped counts, scans them, and publishes a status word before it
// looks back: an inclusive prefix at once when the block holds the first
// cell of a replica (every replica's sorted positions start at its own p0,
// so the scan restarts there), else its aggregate, replaced by the inclusive
// prefix once its look-back is done.  Look-back by the first wavefront
// reads up to 64 predecessors at a time.  Each status word is one 64-bit
// relaxed atomic (value, state, launch epoch), so no fences are needed;
// the cell starts and the scattered particles are read by later kernels
// only.  A workgroup waits only on lower tickets, whose workgroups are
// already running and publish before they wait.  Gated (list mode), a block
// with no cell of a rebuilding replica is skipped, and a block whose first
// cell's replica does not rebuild does not look back.  Listed (gated
// batches): the blocks to sort come from k_decide's list, taken in ticket
// order from a per-step counter, with the step's epoch and a status array
// of their own; workgroups beyond the list leave before taking a ticket.
// ---------------------------------------------------------------------------
constexpr unsigned long long kLbAggregate = 1ull, kLbInclusive = 2ull;
constexpr int kSortIds = 4096;  // slot entries a sort workgroup copies to LDS
constexpr int kSortBatch = 4;   // particles per lane in flight in the scatter
constexpr int kLbSpinLimit = 1 << 22;

__device__ __forceinline__ unsigned long long lb_word(uint32_t epoch, unsigned long long state,
                                                      int v) {
    return ((unsigned long long)(epoch & 0x3fffffffu) << 34) | (state << 32) | (uint32_t)v;
}

// Replica of cell c of sort block b: a binary search over the block's
// replicas (brep, usually one or two).
__device__ __forceinline__ int rep_of_cell(const ScanArgs &a, int b, int c) {
    const int2 br = a.brep[b];
    int lo = br.x, hi = br.y + 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.reps[mid].cell0 <= c) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kSortCells) void k_sort(ScanArgs a) {
    __shared__ int s_ex[kSortCells + 1];  // local exclusive prefix of the counts
    __shared__ int s_st[kSortCells];      // cell starts
    __shared__ int s_lim[kSortCells];     // end of each cell's replica range
    __shared__ int s_w[kSortCells / 64];
    __shared__ unsigned long long s_tk;
    __shared__ int s_carry, s_head;
    __shared__ int s_ids[kSortIds];       // the block's slot entries' caller indices, cell order
    __shared__ uint8_t s_cl[kSortIds];    // their cell (local index)
    __shared__ double s_ored[kObsRed];
    // (cell walk) the previous step's observables, level 1, before the
    // ticket: no workgroup waits on one that has not taken its ticket
    if (a.obs_l1) obs_level1(a.obs, s_ored);
    if (a.any && !*a.any) return;  // list mode, no replica rebuilds (uniform)
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // listed (gated batches): workgroups beyond the listed blocks leave at
    // once, so only the listed blocks take tickets; they take the blocks in
    // ticket order, with the step's epoch and a status array of their own
    const bool listed = a.fblock != nullptr;
    if (listed && (int)blockIdx.x >= a.nf[2]) return;
    if (t == 0)
        s_tk = listed ? (unsigned long long)atomicAdd(a.nf + 3, 1) : atomicAdd(a.ticket, 1ull);
    __syncthreads();
    const unsigned long long tk = s_tk;
    const int b = listed ? a.fblock[tk] : (int)(tk % (unsigned long long)a.nsort);
    const uint32_t epoch =
        listed ? (uint32_t)a.nf[4] : (uint32_t)(tk / (unsigned long long)a.nsort);
    unsigned long long *const status = listed ? a.status2 : a.status;
    const int c0 = b * kSortCells, c = c0 + t;
    if (a.ctl && !listed) {
        // gated (list mode): a block holding no cell of a rebuilding replica
        // has nothing to count, scan or scatter, and no rebuilding replica
        // looks back past its own first block, so it is skipped whole
        __syncthreads();
        if (t == 0) {
            const int2 br = a.brep[b];
            int f = 0;
            for (int q = br.x; q <= br.y; ++q) f |= a.ctl[q].flag;
            s_head = f;
        }
        __syncthreads();
        if (!s_head) return;
    }
    int n = 0;
    if (c < a.ncells) {
        n = min(a.count[c], a.cap);
        a.count[c] = 0;  // ready for the next binning
    }
    const int r = rep_of_cell(a, b, min(c, a.ncells - 1));
    const RepDesc &R = a.reps[r];
    s_lim[t] = R.p1;
    const int hl = R.cell0 - c0;  // replica's first cell, local (< 0: earlier block)
    // block exclusive scan
    const int incl = wave_incl_scan(n);
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    int wb = 0;
    for (int k = 0; k < w; ++k) wb += s_w[k];
    const int ex = wb + incl - n;
    s_ex[t] = ex;
    if (t == kSortCells - 1) s_ex[kSortCells] = ex + n;
    __syncthreads();
    const int tot = s_ex[kSortCells];
    // publish (the last cell's replica decides: its first cell in this block
    // means the block's end value is known without look-back)
    if (t == kSortCells - 1) {
        s_head = hl >= 0;
        const unsigned long long v = (hl >= 0)
            ? lb_word(epoch, kLbInclusive, R.p0 + tot - s_ex[hl])
            : lb_word(epoch, kLbAggregate, tot);
        __hip_atomic_store(status + b, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // look-back, needed when the block's first cell continues a replica
    // begun in an earlier block (uniform: read from the first cell)
    // (not for a replica that does not rebuild: its starts are not written)
    if (t == 0) s_carry = (hl >= 0 || (a.ctl && !a.ctl[r].flag)) ? -1 : 0;
    __syncthreads();
    const bool need = s_carry == 0;
    if (need && w == 0) {
        int carry = 0;
        int j = b - 1;
        int spins = 0;
        while (true) {
            const int q = j - lane;
            unsigned long long v = 0;
            if (q >= 0) v = __hip_atomic_load(status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const bool cur = (uint32_t)(v >> 34) == (epoch & 0x3fffffffu);
            const unsigned long long st = (v >> 32) & 3ull;
            // q < 0 cannot be reached before an inclusive word (block 0
            // always holds replica 0's first cell); treated as inclusive 0
            const bool valid = (q < 0) || (cur && st != 0ull);
            const bool inc = (q < 0) || (cur && st == kLbInclusive);
            const unsigned long long im = __ballot(inc), vm = __ballot(valid);
            const int k = im ? (__ffsll(im) - 1) : 63;
            const unsigned long long want = (k == 63) ? ~0ull : ((2ull << k) - 1ull);
            if ((vm & want) != want) {
                if (++spins > kLbSpinLimit) {  // never expected: report, end the wait
                    if (lane == 0) atomicOr(a.err, 32u);  // FM_DEV_SORT_STALL
                    break;
                }
                continue;
            }
            carry += wave_sum_int((lane <= k && q >= 0) ? (int)(uint32_t)v : 0);
            if (im) break;
            j -= 64;
        }
        if (lane == 0) {
            s_carry = carry;
            if (!s_head)  // a block holding a replica's first cell published already
                __hip_atomic_store(status + b, lb_word(epoch, kLbInclusive, carry + tot),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    const int st = (hl >= 0) ? R.p0 + ex - s_ex[hl] : s_carry + ex;
    s_st[t] = st;
    if (c < a.ncells) {
        // list mode: only the rebuilding replicas' cell starts change (the
        // others' lists index staging laid out by their own last starts)
        if (!a.ctl || a.ctl[r].flag) a.start[c] = st;
        if (c == a.ncells - 1) a.start[a.ncells] = st + n;
    }
    __syncthreads();
    // scatter: the block's particles in cell order; a particle's sorted
    // position is its cell's start + its rank among the cell's members by
    // their index in the caller's order, as the reference's cell fill walks
    // the particles (femmd2D/src/nn.cpp:8-20), whatever the internal order.
    // The block's slot entries' caller indices are first gathered into LDS
    // (loads batched per lane), so the ranks read LDS; a block with more
    // particles than the LDS copy holds reads the slot lists from memory.
    const bool in_lds = tot <= kSortIds;
    auto cell_of = [&](int q) {
        int lo = 0, hi = kSortCells;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_ex[mid] <= q) lo = mid; else hi = mid;
        }
        return lo;
    };
    if (in_lds) {
        for (int q0 = t; q0 < tot; q0 += kSortBatch * kSortCells) {
            int v[kSortBatch], cl[kSortBatch];
#pragma unroll
            for (int u = 0; u < kSortBatch; ++u) {
                const int q = q0 + u * kSortCells;
                if (q < tot) {
                    cl[u] = cell_of(q);
                    v[u] = a.tmp[(size_t)(c0 + cl[u]) * a.cap + (q - s_ex[cl[u]])];
                }
            }
#pragma unroll
            for (int u = 0; u < kSortBatch; ++u) {
                const int q = q0 + u * kSortCells;
                if (q < tot) {
                    s_ids[q] = a.uid[v[u]];
                    s_cl[q] = (uint8_t)cl[u];
                }
            }
        }
        __syncthreads();
    }
    for (int q0 = t; q0 < tot; q0 += kSortBatch * kSortCells) {
        int ii[kSortBatch], pos[kSortBatch];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) {
            const int q = q0 + u * kSortCells;
            ii[u] = -1;
            if (q < tot) {
                const int cl = in_lds ? (int)s_cl[q] : cell_of(q);
                const int b = s_ex[cl], m = s_ex[cl + 1] - b;
                // (caller indices are unique within a replica, whose cells
                // these are: a strict order)
                const int *src = a.tmp + (size_t)(c0 + cl) * a.cap;
                const int i = src[q - b];  // the particle (internal index)
                int rank = 0;
                if (in_lds) {
                    const int ui = s_ids[q];
                    for (int k = 0; k < m; ++k) rank += (s_ids[b + k] < ui);
                } else {
                    const int ui = a.uid[i];
                    for (int k = 0; k < m; ++k) rank += (a.uid[src[k]] < ui);
                }
                ii[u] = i;
                pos[u] = s_st[cl] + rank;
                // counts that add up past the replica's particles (a binning
                // on top of one not yet sorted) are reported, never written
                if (pos[u] >= s_lim[cl]) {
                    atomicOr(a.err, 1u);  // FM_DEV_CELL_OVERFLOW
                    ii[u] = -1;
                }
            }
        }
        float4 e[kSortBatch];
        int ty[kSortBatch];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) {
            if (ii[u] >= 0) {
                const int i = ii[u];
                e[u].x = a.x[i];
                e[u].y = a.y[i];
                e[u].z = __int_as_float(a.grp[i]);
                e[u].w = a.sig ? a.sig[i] : 0.0f;
                ty[u] = a.typ ? a.typ[i] : 0;
            }
        }
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) {
            if (ii[u] >= 0) {
                const int i = ii[u], ps = pos[u];
                a.perm[ps] = i;
                a.sorted[ps] = e[u];
                if (a.typ) a.sorted_typ[ps] = ty[u];
                if (a.spos) {  // list mode: the sorted-order state the list pass reads
                    a.spos[i] = ps;
                    a.xys[ps] = make_float2(e[u].x, e[u].y);
                    if (a.sigs) a.sigs[ps] = e[u].w;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Force pass.
// ---------------------------------------------------------------------------

// Workgroup id -> tile id so that each XCD (workgroups are dealt round-robin
// over the 8 XCDs) walks a contiguous band of cell rows, keeping the
// neighbouring rows it re-reads in its own L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int nx = 8;
    const int q = n / nx, r = n % nx;
    const int x = b % nx, k = b / nx;
    return x * q + min(x, r) + k;
}

struct PairLds {
    float4 *ent;   // staged particles {x, y, grp bits, sigma}
    int *typ;      // staged types (typed path)
    int *sp;       // staged sorted positions (diagnostic path)
    int *cstart;   // per slot: first sorted position
    int *off;      // per slot: first LDS entry (S + 1)
    PairEntry *pt; // pair table of the replica
    double *red;   // 3 per wavefront
    float *redf;   // 2 per wavefront
};

// The kick a tile gives its centre particles (the kick modes of k_forces):
// accumulators of the kicked particles' kinetic energy, max v^2 and max
// squared displacement.
struct TileKick {
    double ke = 0.0;
    float vv = 0.0f, d2 = 0.0f;
};

// Kick of the tile's centre particle id (sorted position p) at its staged
// position (x, y) (the same bits as x[id]) with pair force (fxs, fys).
// kKick 1: second half-kick only (the closing kick of a run); 2: second and
// first half-kick, drift and wrap.  A particle with element or bond corners
// is left to the node kick, which adds its gathered corner forces: its pair
// force is stored for it (and every particle's in mode 0 and 1, for the next
// run's first half-kick).  Returns the cell of the drifted position (cell
// walk, mode 2), -1 when there is none to bin.
template <int kKick, bool kList, bool kEnergy>
__device__ __forceinline__ int tile_kick(const ForceArgs &a, const RepDesc &R, int id, int p,
                                         float x, float y, float fxs, float fys, TileKick &tk) {
    const bool node = id >= R.pg;  // (internal order: these come last)
    if (kKick != 2 || node) {
        a.fx[id] = fxs;
        a.fy[id] = fys;
    }
    if (kKick == 0 || node) return -1;
    const float im = R.im_uniform > 0.0f ? R.im_uniform : a.inv_m[id];
    const float ax = fxs * im, ay = fys * im;
    float vx = a.vx[id], vy = a.vy[id], dxp = 0.0f, dyp = 0.0f, vv = 0.0f;
    double ke = 0.0;
    kick_particle<true, kKick == 2, kEnergy>(R, ax, ay, im, vx, vy, x, y, ke, vv, dxp, dyp, a.err);
    a.vx[id] = vx;
    a.vy[id] = vy;
    if (kEnergy) {
        tk.ke += ke;
        tk.vv = fmaxf(tk.vv, vv);
    }
    if (kKick != 2) return -1;
    a.xw[id] = x;
    a.yw[id] = y;
    if (a.rec) {
        a.rec[id] = x;
        a.rec[a.n + id] = y;
    }
    if (kList) {
        a.xys_out[p] = make_float2(x, y);
        tk.d2 = fmaxf(tk.d2, dxp * dxp + dyp * dyp);
        return -1;
    }
    return R.cell0 + cell_key(x, y, R);
}

// End of a tile: fixed-order sums of U, W and the kinetic energy, maxima of
// v^2 and of the squared displacement (valid in thread 0).
__device__ void tile_reduce(double &u, double &w, TileKick &tk, const PairLds &L) {
    u = wave_sum(u);
    w = wave_sum(w);
    double ke = wave_sum(tk.ke);
    float vv = wave_max(tk.vv), d2 = wave_max(tk.d2);
    const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        L.red[3 * wv] = u;
        L.red[3 * wv + 1] = w;
        L.red[3 * wv + 2] = ke;
        L.redf[2 * wv] = vv;
        L.redf[2 * wv + 1] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double su = 0.0, sw = 0.0, sk = 0.0;
        float mv = 0.0f, md = 0.0f;
        for (int k = 0; k < nw; ++k) {
            su += L.red[3 * k];
            sw += L.red[3 * k + 1];
            sk += L.red[3 * k + 2];
            mv = fmaxf(mv, L.redf[2 * k]);
            md = fmaxf(md, L.redf[2 * k + 1]);
        }
        u = su;
        w = sw;
        tk.ke = sk;
        tk.vv = mv;
        tk.d2 = md;
    }
}

// Writes a tile's partials: U, W (kEnergy), the kicked particles' kinetic
// energy and max v^2 (kEnergy with a kick) and its largest displacement
// into one of the replica's slots for the next rebuild decision (list mode,
// step kick).  Thread 0.
template <int kKick, bool kList, bool kEnergy>
__device__ __forceinline__ void tile_partials(const ForceArgs &a, int tile, int rep, double u,
                                              double w, const TileKick &tk) {
    if (threadIdx.x != 0) return;
    if (kEnergy) {
        a.lj_part[2 * tile] = u;
        a.lj_part[2 * tile + 1] = w;
    }
    if (kEnergy && kKick != 0) {
        a.ke_tile[2 * tile] = tk.ke;
        a.ke_tile[2 * tile + 1] = (double)tk.vv;
    }
    if (kList && kKick == 2)
        atomicMax(a.dmax_out + ((size_t)rep * kDispSlots + tile % kDispSlots) * kDispStride,
                  __float_as_uint(tk.d2));
}

// LJ over one tile (femmd2D/src/interactionsCuda.cu:43-100 pair maths,
// full-list accumulation of femmd2D/src/interactions.cpp:347-389: each
// ordered pair adds to f_i only, U and W take 1/2).
template <bool kWrap, bool kSigma, bool kTypes, bool kDebug, bool kEnergy, int kKick>
__device__ void pair_tile_body(const ForceArgs &a, const RepDesc &R, int rep, int W,
                               int tile, PairLds &L) {
    const int row = W + 2;
    const int c_lo = L.off[row + 1], c_hi = L.off[row + W + 1];
    const int nc = c_hi - c_lo;
    const int sbase = L.cstart[row + 1];
    const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
    const int nt = R.ntypes;
    double Ublk = 0.0, Wblk = 0.0;
    TileKick tk;
    // every lane of a wavefront takes part in each pass (the binning of the
    // drifted positions is a wavefront operation)
    for (int k0 = 0; k0 < nc; k0 += blockDim.x) {
        const int k = k0 + (int)threadIdx.x;
        int id = 0, cell = -1;
        if (k < nc) {
        const int me = c_lo + k;
        // slot of this centre particle within row 1 (1..W)
        int lo = row + 1, hi = row + W + 1;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (L.off[mid] <= me) lo = mid; else hi = mid;
        }
        const int col = lo - row;
        const float4 pi = L.ent[me];
        const float xi = pi.x, yi = pi.y;
        const float xim = xi - Lx, yim = yi - Ly;
        const int gi = __float_as_int(pi.z);
        const int gix = gi >= 0 ? gi : -2;  // -1 never matches -2
        const float si = pi.w;
        const int ti = kTypes ? L.typ[me] : 0;
        const int a0 = L.off[col - 1], b0 = L.off[col + 2];
        const int a1 = L.off[row + col - 1], b1 = L.off[row + col + 2];
        const int a2 = L.off[2 * row + col - 1], b2 = L.off[2 * row + col + 2];
        const int total = (b0 - a0) + (b1 - a1) + (b2 - a2);
        int j = a0;
        if (j == b0) j = a1;
        if (j == b1) j = a2;
        // single-type accumulators: A = sum s12, B = sum s6 (s6 = (sigma/r)^6)
        // give U = 4 eps (A - B) + n shift and W = 48 eps (A - B/2); the
        // typed path accumulates eps-weighted terms per pair instead
        float fx = 0.f, fy = 0.f, accA = 0.f, accB = 0.f, ush = 0.f;
        int nint = 0;
        long long s1 = 0, s2 = 0;
        const unsigned rc2b_u = __float_as_uint(R.rc2_uniform);
        // the next candidate's LDS entry is read one iteration ahead
        float4 pn = L.ent[total > 0 ? j : 0];
        for (int t = 0; t < total; ++t) {
            const float4 pj = pn;
            int jn = j + 1;
            if (jn == b0) jn = a1;
            if (jn == b1) jn = a2;
            pn = L.ent[(t + 1 < total) ? jn : j];
            const int gj = __float_as_int(pj.z);
            // femmd2D/src/interactionsCuda.cu:84-90, FP32, reference order
            float dx = xi - pj.x;
            float dy = yi - pj.y;
            int sx = 0, sy = 0;
            if (kWrap) {
                if (dx >= hLx) { dx -= Lx; sx = 1; } else if (dx < -hLx) { dx += Lx; sx = -1; }
                if (dy >= hLy) { dy -= Ly; sy = 1; } else if (dy < -hLy) { dy += Ly; sy = -1; }
            }
            const float dr2 = dx * dx + dy * dy;
            float s2ij, e48 = 0.f, e4 = 0.f, shift = 0.f;
            unsigned rc2b;
            if (kTypes) {
                const PairEntry pe = L.pt[ti * nt + (kTypes ? L.typ[j] : 0)];
                e48 = pe.e48; e4 = pe.e4; shift = pe.shift;
                const float sij = kSigma ? (si + pj.w) * 0.5f : R.s_uniform;  // (no sigma staged)
                rc2b = __float_as_uint(pe.rc * pe.rc * sij * sij);
                s2ij = sij * sij;
            } else if (kSigma) {
                const float sij = (si + pj.w) * 0.5f;
                const float rc = L.pt[0].rc;
                rc2b = __float_as_uint(rc * rc * sij * sij);
                s2ij = sij * sij;
            } else {
                rc2b = rc2b_u;
                s2ij = R.s2_uniform;
            }
            // the cutoff predicate dr2 < rc2 on the bit patterns of
            // non-negative floats; the centre's own slot is left out by index
            // (i != j of femmd2D/src/nn.cpp:44-46: a distinct particle on the
            // same spot interacts, as in the reference; the 3 x 3 stencil
            // holds every particle once)
            const bool hit = (__float_as_uint(dr2) < rc2b) & (j != me) & (gj != gix);
            if (hit) {
                // force from the image-exact difference: the far coordinate
                // is shifted by L before subtracting (exact by Sterbenz)
                // (same unfused expression as dr2: an unshifted pair gets the
                // same bits on interior and boundary tiles, so results do not
                // depend on the tile shape)
                float dxa = dx, dya = dy, r2a = dr2;
                if (kWrap) {
                    if (sx > 0) dxa = xim - pj.x; else if (sx < 0) dxa = xi - (pj.x - Lx);
                    if (sy > 0) dya = yim - pj.y; else if (sy < 0) dya = yi - (pj.y - Ly);
                    r2a = dxa * dxa + dya * dya;
                }
                const float inv = __builtin_amdgcn_rcpf(r2a);
                const float sr2 = s2ij * inv;
                const float sr4 = sr2 * sr2;
                const float sr6 = sr4 * sr2;
                const float s12 = sr6 * sr6;
                const float tv = __builtin_fmaf(-0.5f, sr6, s12);  // s6 (s6 - 1/2)
                if (kTypes) {
                    const float fc = e48 * tv * inv;
                    fx = __builtin_fmaf(fc, dxa, fx);
                    fy = __builtin_fmaf(fc, dya, fy);
                    if (kEnergy) {
                        accA = __builtin_fmaf(e4, s12 - sr6, accA);  // energy
                        accB = __builtin_fmaf(e48, tv, accB);        // virial
                        ush += shift;
                    }
                } else {
                    const float fc = tv * inv;
                    fx = __builtin_fmaf(fc, dxa, fx);
                    fy = __builtin_fmaf(fc, dya, fy);
                    if (kEnergy) {
                        accA += s12;
                        accB += sr6;
                    }
                }
                if (kEnergy || kDebug) ++nint;
                if (kDebug) {
                    const long long v = (long long)a.uid[a.perm[L.sp[j]]] + 1;
                    s1 += v;
                    s2 += v * v;
                }
            }
            j = jn;
        }
        double Ui, Wi;
        if (!kTypes) {
            const PairEntry pe = L.pt[0];
            fx *= pe.e48;
            fy *= pe.e48;
            const double A = accA, B = accB;
            Ui = (double)pe.e4 * (A - B) + (double)nint * pe.shift;
            Wi = (double)pe.e48 * (A - 0.5 * B);
        } else {
            Ui = (double)accA + (double)ush;
            Wi = (double)accB;
        }
        id = a.perm[sbase + k];
        cell = tile_kick<kKick, false, kEnergy>(a, R, id, sbase + k, xi, yi, fx, fy, tk);
        if (kEnergy) {
            Ublk += 0.5 * Ui;
            Wblk += 0.5 * Wi;
        }
        if (kDebug) {
            const int u = a.uid[id];
            a.dbg_cnt[u] = nint;
            a.dbg_s1[u] = s1;
            a.dbg_s2[u] = s2;
        }
        }
        if (kKick == 2) bin_wave_to(a.count, a.tmp, a.cell_cap, a.err, cell >= 0, id, cell);
    }
    if (kEnergy || kKick == 2) {
        tile_reduce(Ublk, Wblk, tk, L);
        tile_partials<kKick, false, kEnergy>(a, tile, rep, Ublk, Wblk, tk);
    }
}

// Stages the 3 x (W+2) cells of a tile into LDS: rows cy-1, cy, cy+1 and
// columns cx0-1 .. cx0+W with the periodic wrap of femmd2D/src/nnCuda.cu:
// 26-37; sorted positions are contiguous per cell.  With wsrc set (the
// adjoint pass) the per-particle vector w is staged beside each entry.
// Returns the number of staged particles, or -1 if they exceed the LDS.
// Layout of a tile's staging: the 3 x (W+2) cells of rows cy-1, cy, cy+1 and
// columns cx0-1 .. cx0+W with the periodic wrap of femmd2D/src/nnCuda.cu:
// 26-37.  Fills L.cstart (first sorted position per slot), L.off (first LDS
// entry per slot, S + 1) and the pair table; returns the number of staged
// particles, or -1 if they exceed the LDS capacity.
__device__ int stage_layout(const ForceArgs &a, const Tile &T, const RepDesc &R, int W,
                            PairLds &L) {
    const int row = W + 2, S = 3 * row;
    const int tid = threadIdx.x;
    for (int s = tid; s < S; s += blockDim.x) {
        const int r = s / row, col = s - r * row;
        int cy = T.cy + r - 1;
        if (cy < 0) cy += R.ncy; else if (cy >= R.ncy) cy -= R.ncy;
        int cx = T.cx0 + col - 1;
        if (cx < 0) cx += R.ncx; else if (cx >= R.ncx) cx -= R.ncx;
        const int c = R.cell0 + cy * R.ncx + cx;
        const int st = a.start[c];
        L.cstart[s] = st;
        L.off[s + 1] = a.start[c + 1] - st;  // count, scanned below
    }
    if (L.pt)  // (the list build needs no pair table)
        for (int k = tid; k < R.ntypes * R.ntypes; k += blockDim.x) L.pt[k] = a.ptab[R.pt + k];
    __syncthreads();
    // exclusive scan of S <= 198 counts by the first wavefront (4 per lane)
    if (tid < 64) {
        int v[4], s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = 4 * tid + q;
            v[q] = (idx < S) ? L.off[idx + 1] : 0;
            s += v[q];
        }
        const int incl = wave_incl_scan(s);
        int run = incl - s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = 4 * tid + q;
            if (idx < S) {
                L.off[idx] = run;
                run += v[q];
            }
        }
        if (tid == 63) L.off[S] = incl;
    }
    __syncthreads();
    const int M = L.off[S];
    return M > a.lds_cap ? -1 : M;
}

// Copies every staged entry e (LDS index) from its sorted position p:
// st(e, ld(p)).  The cells of one staged row are consecutive in sorted order
// except where the column index wraps to 0 (at most twice), so each row is
// up to three contiguous runs; the rows themselves are consecutive in LDS.
// Lanes stride over all M entries of the tile, kStageBatch per lane at a
// time with every load issued before the first store.
constexpr int kStageBatch = 4;
// Position of the list-padding slot staged after a tile's particles: far
// from every particle (squared distances stay finite and beyond any cutoff).
constexpr float kSentinel = 1.0e18f;
template <class Ld, class St>
__device__ void stage_copy(const Tile &T, const RepDesc &R, int W, const PairLds &L, Ld &&ld,
                           St &&st) {
    using V = decltype(ld(0));
    const int row = W + 2;
    // slots (within a row) where the column wraps to 0
    int b1 = (T.cx0 == 0) ? 1 : (R.ncx - T.cx0 + 1);
    int b2 = b1 + R.ncx;
    if (b1 >= row) b1 = row;
    if (b2 >= row) b2 = row;
    const int r1 = L.off[row], r2 = L.off[2 * row], M = L.off[3 * row];
    const int bd = blockDim.x;
    for (int e0 = (int)threadIdx.x; e0 < M; e0 += kStageBatch * bd) {
        V v[kStageBatch];
#pragma unroll
        for (int q = 0; q < kStageBatch; ++q) {
            const int e = e0 + q * bd;
            if (e < M) {
                const int s0 = (e >= r2) ? 2 * row : ((e >= r1) ? row : 0);
                const int o1 = L.off[s0 + b1], o2 = L.off[s0 + b2];
                const int ps = s0 + ((e >= o2) ? b2 : ((e >= o1) ? b1 : 0));
                v[q] = ld(L.cstart[ps] + (e - L.off[ps]));
            }
        }
#pragma unroll
        for (int q = 0; q < kStageBatch; ++q) {
            const int e = e0 + q * bd;
            if (e < M) st(e, v[q]);
        }
    }
}

// Stages a tile's particles from the sorted arrays ({x, y, grp bits,
// sigma} and types; with wsrc set, the adjoint's per-particle vector beside
// them).  Returns the number of staged particles, or -1 if they exceed the
// LDS capacity.
template <bool kTypes, bool kDebug, bool kPid = false>
__device__ int stage_tile(const ForceArgs &a, const Tile &T, const RepDesc &R, int W,
                          PairLds &L, const float2 *wsrc, float2 *lw) {
    const int M = stage_layout(a, T, R, W, L);
    if (M < 0) return -1;
    struct Ent {
        float4 v;
        int t, sp;
        float2 w;
    };
    stage_copy(
        T, R, W, L,
        [&](int p) {
            Ent u;
            u.v = a.sorted[p];
            u.t = kTypes ? a.sorted_typ[p] : 0;
            u.sp = kDebug ? p : (kPid ? a.perm[p] : 0);  // sorted position / internal id
            if (wsrc) u.w = wsrc[a.perm[p]];
            return u;
        },
        [&](int e, const Ent &u) {
            L.ent[e] = u.v;
            if (kTypes) L.typ[e] = u.t;
            if (kDebug || kPid) L.sp[e] = u.sp;
            if (wsrc) lw[e] = u.w;
        });
    __syncthreads();
    return M;
}

template <bool kSigma, bool kTypes, bool kDebug, bool kEnergy, int kKick>
__device__ void pair_tile(const ForceArgs &a, int tile, PairLds &L) {
    const Tile T = a.tiles[tile];
    const RepDesc R = a.reps[T.rep];
    const int W = T.w_flags & 0xffff;
    const bool wrapT = (T.w_flags >> 16) & 1;
    if (stage_tile<kTypes, kDebug>(a, T, R, W, L, nullptr, nullptr) < 0) {
        atomicOr(a.err, 2u);  // FM_DEV_TILE_OVERFLOW (its particles are not moved)
        tile_partials<kKick, false, kEnergy>(a, tile, T.rep, 0.0, 0.0, TileKick{});
        return;
    }
    if (wrapT)
        pair_tile_body<true, kSigma, kTypes, kDebug, kEnergy, kKick>(a, R, T.rep, W, tile, L);
    else
        pair_tile_body<false, kSigma, kTypes, kDebug, kEnergy, kKick>(a, R, T.rep, W, tile, L);
}

// ---------------------------------------------------------------------------
// Neighbour-list mode (femmd2D/src/nn.cpp:22-62, re-designed): a tile of
// cells of side >= rc*sigma + skin is staged as in the pair pass; every
// centre particle records the partners within rc*sigma + skin (exclusion
// rule applied here, once) by the byte offset of their slot in the list
// pass's LDS staging (8 bytes per slot, so a tile stages at most 8191
// particles), 16 bits each, four per group, groups interleaved over
// particles (group g of sorted
// position p at g * n + p) so that a wavefront of consecutive positions reads
// one group in one coalesced load.  Until the next rebuild the sorted order
// and the tiles stay fixed, so the list pass restages each tile the same way
// from the current positions and gathers the neighbours from LDS.
// ---------------------------------------------------------------------------

// Stores group g (four LDS indices j[0..3]) of sorted position p.
__device__ __forceinline__ void list_store_group(const ForceArgs &a, int g, int p,
                                                 const int *j) {
    const uint64_t v = (uint64_t)(uint16_t)j[0] | ((uint64_t)(uint16_t)j[1] << 16) |
                       ((uint64_t)(uint16_t)j[2] << 32) | ((uint64_t)(uint16_t)j[3] << 48);
    reinterpret_cast<uint64_t *>(a.nbr)[(size_t)g * a.n + p] = v;
}

// Stores slot k alone (list build without the LDS staging buffer).
__device__ __forceinline__ void list_store_slot(const ForceArgs &a, int k, int p, int j) {
    reinterpret_cast<uint16_t *>(a.nbr)[((size_t)(k >> 2) * a.n + p) * 4 + (k & 3)] =
        (uint16_t)j;
}

// List build of one staged tile, entries written straight to the list (used
// when the LDS holds no list buffer: one 2-byte store per entry).
// Which centre columns (1 .. W in the staging) of a tile need the minimum
// image: those next to a halo column across the periodic boundary in x, and
// all of them when a halo row crosses it in y or the box has fewer than 5
// cells along an axis (elsewhere a neighbour is less than 2 cells <= 0.4 L
// away in the same periodic image, so the image shift is zero).  Every entry
// of such a list takes it, not only those in the halo across the boundary:
// the centre and its neighbours in the boundary cells can wrap between two
// builds.  A particle that wrapped is then more than a cell width minus
// skin / 2 > rc from every centre two or more cells from the boundary, so
// the plain difference rejects it there as the minimum image would.
struct ImgCols {
    bool all, left, right;
    __device__ bool needs(int col, int W) const {
        return all || (left && col == 1) || (right && col == W);
    }
};
__device__ __forceinline__ ImgCols img_cols(const Tile &T, const RepDesc &R, int W) {
    ImgCols c;
    c.all = R.ncx < 5 || R.ncy < 5 || T.cy == 0 || T.cy >= R.ncy - 1;
    c.left = T.cx0 == 0;
    c.right = T.cx0 + W >= R.ncx;
    return c;
}

template <bool kWrap, bool kSigma>
__device__ void list_tile_body_direct(const ForceArgs &a, const Tile &T, const RepDesc &R, int W,
                                      PairLds &L) {
    const ImgCols ic = img_cols(T, R, W);
    const int row = W + 2;
    const int c_lo = L.off[row + 1], c_hi = L.off[row + W + 1];
    const int nc = c_hi - c_lo;
    const int sbase = L.cstart[row + 1];
    const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
    const int bd = blockDim.x;
    const int cap = a.nbr_cap;
    for (int k = threadIdx.x; k < nc; k += bd) {
        const int me = c_lo + k;
        int lo = row + 1, hi = row + W + 1;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (L.off[mid] <= me) lo = mid; else hi = mid;
        }
        const int col = lo - row;
        const float4 pi = L.ent[me];
        const float xi = pi.x, yi = pi.y;
        const int gi = __float_as_int(pi.z);
        const int gix = gi >= 0 ? gi : -2;
        const float si = pi.w;
        const int p = sbase + k;
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int j0 = L.off[r * row + col - 1], j1 = L.off[r * row + col + 2];
            for (int j = j0; j < j1; ++j) {
                const float4 pj = L.ent[j];
                float dx = xi - pj.x;
                float dy = yi - pj.y;
                if (kWrap) {
                    float ox = (dx >= hLx) ? Lx : 0.0f;
                    ox = (dx < -hLx) ? -Lx : ox;
                    float oy = (dy >= hLy) ? Ly : 0.0f;
                    oy = (dy < -hLy) ? -Ly : oy;
                    dx = dx - ox;
                    dy = dy - oy;
                }
                const float dr2 = dx * dx + dy * dy;
                float rl2;
                if (kSigma) {
                    const float rl = R.rc_list * ((si + pj.w) * 0.5f) + R.skin;
                    rl2 = (rl * rl) * 1.00002f;
                } else {
                    rl2 = R.rl2_uniform;
                }
                bool hit = (dr2 < rl2) & (__float_as_int(pj.z) != gix);
                if (r == 1) hit &= (j != me);
                if (hit) {
                    if (cnt < cap) list_store_slot(a, cnt, p, 8 * j);
                    ++cnt;
                }
            }
        }
        const int m = min(cnt, cap);
        // groups of four, padded with the sentinel slot (staged far away:
        // the pair pass's cutoff test rejects it); cap is a multiple of 8
        for (int kk = m; kk < ((m + 3) & ~3); ++kk) list_store_slot(a, kk, p, 8 * a.lds_cap);
        // list slot p holds particle k of the tile
        a.nnum[p] = m | (kWrap && ic.needs(col, W) ? kListImg : 0) | (k << 16);
        if (cnt > cap) atomicOr(a.err, 16u);  // FM_DEV_LIST_OVERFLOW
    }
}

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// Rank of this lane's list among the active lanes of the workgroup: the
// lists of particles that need the minimum image after all others, each
// part longest first (by groups of four entries), ties by lane: a counting
// sort over (image, length, wavefront) with the ties inside a wavefront
// ordered by ballots.  hist holds 2 (capg + 1) (blockDim / 64) ints.
// Deterministic.
__device__ int list_rank(int n4, bool img, bool act, int capg, int *hist) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nw = blockDim.x >> 6, nk = 2 * (capg + 1), tot = nw * nk;
    const int key = (img ? capg + 1 : 0) + capg - min(n4, capg);  // 0 = longest
    __syncthreads();  // earlier readers of hist are done
    for (int i = tid; i < tot; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    if (act) atomicAdd(&hist[key * nw + w], 1);
    __syncthreads();
    if (w == 0) {  // exclusive scan in (length, wavefront) order
        const int per = (tot + 63) >> 6, b0 = lane * per;
        int sum = 0;
        for (int i = 0; i < per; ++i)
            if (b0 + i < tot) sum += hist[b0 + i];
        int run = wave_incl_scan(sum) - sum;
        for (int i = 0; i < per; ++i)
            if (b0 + i < tot) {
                const int v = hist[b0 + i];
                hist[b0 + i] = run;
                run += v;
            }
    }
    __syncthreads();
    unsigned long long eq = __ballot(act);
#pragma unroll
    for (int bit = 0; bit < 9; ++bit) {  // key < 2 (kMaxListGroups + 1) <= 512
        const bool kb = (key >> bit) & 1;
        const unsigned long long m = __ballot(kb);
        eq &= kb ? m : ~m;
    }
    const unsigned long long before = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    return act ? hist[key * nw + w] + __popcll(eq & before) : -1;
}

// List build of one staged tile through the LDS list buffer: slot k of lane
// t at buf[k * (blockDim + 4) + t] holds the entry's offset in its row's run
// of candidates (< 3 cell capacities <= 180: one byte; the lane's counts at
// the two row ends tell the row): half the LDS of two-byte entries, so more
// workgroups share a CU, which this latency-bound loop needs.  Each
// row's candidate loop runs a wavefront-uniform trip count (the longest run
// among the lanes; a lane past its own run is masked), and every candidate
// is written to the lane's next buffer slot while only hits advance it, so
// the loop has no branches.  Reads past a lane's run stay inside the
// staging's headroom of 3 cells.
template <bool kWrap, bool kSigma>
__device__ void list_tile_body_buf(const ForceArgs &a, const Tile &T, const RepDesc &R, int W,
                                   PairLds &L, uint8_t *buf, int *hist) {
    const ImgCols ic = img_cols(T, R, W);
    const int row = W + 2;
    const int c_lo = L.off[row + 1], c_hi = L.off[row + W + 1];
    const int nc = c_hi - c_lo;
    const int sbase = L.cstart[row + 1];
    const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
    const int bd = blockDim.x;  // held in a register across the loops
    const int bs = bd + 4;      // buffer row stride (bytes): a slot's rows rotate over the banks
    const int tid = threadIdx.x;
    const int cap = a.nbr_cap, capm1 = a.nbr_cap - 1;
    // every lane of a wavefront takes part in each pass (uniform loop)
    for (int k0 = 0; k0 < nc; k0 += bd) {
        const int k = k0 + tid;
        const bool act = k < nc;
        const int me = c_lo + min(k, nc - 1);
        int lo = row + 1, hi = row + W + 1;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (L.off[mid] <= me) lo = mid; else hi = mid;
        }
        const int col = lo - row;
        const float4 pi = L.ent[me];
        const float xi = pi.x, yi = pi.y;
        const int gi = __float_as_int(pi.z);
        const int gix = gi >= 0 ? gi : -2;
        const float si = pi.w;
        int cnt = 0;
        int jr[3], cr[2];  // each row's first candidate; the count after rows 0 and 1
        // the three staged rows, each a contiguous run of candidates
        // (cells col-1 .. col+1); the centre particle is in row 1 only
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int j0 = L.off[r * row + col - 1];
            jr[r] = j0;
            const int len = act ? L.off[r * row + col + 2] - j0 : 0;
            const int lmax = __builtin_amdgcn_readfirstlane(wave_max_int(len));
#pragma unroll 2
            for (int t = 0; t < lmax; ++t) {
                const int j = j0 + t;
                const float4 pj = L.ent[j];
                float dx = xi - pj.x;
                float dy = yi - pj.y;
                if (kWrap) {
                    // minimum image as two selects per axis
                    float ox = (dx >= hLx) ? Lx : 0.0f;
                    ox = (dx < -hLx) ? -Lx : ox;
                    float oy = (dy >= hLy) ? Ly : 0.0f;
                    oy = (dy < -hLy) ? -Ly : oy;
                    dx = dx - ox;
                    dy = dy - oy;
                }
                // the margins of the list radius cover the rounding of this
                // fused form
                const float dr2 = __builtin_fmaf(dx, dx, dy * dy);
                float rl2;
                if (kSigma) {
                    const float rl = R.rc_list * ((si + pj.w) * 0.5f) + R.skin;
                    rl2 = (rl * rl) * 1.00002f;
                } else {
                    rl2 = R.rl2_uniform;
                }
                bool hit = (t < len) & (dr2 < rl2) & (__float_as_int(pj.z) != gix);
                if (r == 1) hit &= (j != me);
                // (24-bit multiply: full rate; cnt <= cap <= 512, bs <= 260)
                buf[__umul24(min(cnt, capm1), bs) + tid] = (uint8_t)t;
                cnt += hit ? 1 : 0;
            }
            if (r < 2) cr[r] = cnt;
        }
        const int m = min(cnt, cap);
        // the chunk's lists are stored longest first (list slot q = sbase +
        // k0 + rank), so that the pass's wavefronts take lists of similar
        // length; nnum[q] = length | (tile index k << 16)
        const bool img = kWrap && ic.needs(col, W);
        const int rank = list_rank((m + 3) >> 2, img, act, cap >> 2, hist);
        if (act) {
            const int p = sbase + k0 + rank;  // list slot
            // groups of four, padded with the sentinel slot (staged far
            // away: the pair pass's cutoff test rejects it); cap is a
            // multiple of 8.
            for (int g = 0; g < ((m + 3) >> 2); ++g) {
                int jj[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int kk = 4 * g + t;
                    const int j0 = kk < cr[0] ? jr[0] : (kk < cr[1] ? jr[1] : jr[2]);
                    jj[t] = kk < m ? 8 * (j0 + (int)buf[__umul24(kk, bs) + tid]) : 8 * a.lds_cap;
                }
                list_store_group(a, g, p, jj);
            }
            a.nnum[p] = m | (img ? kListImg : 0) | (k << 16);
            if (cnt > cap) atomicOr(a.err, 16u);  // FM_DEV_LIST_OVERFLOW
        }
    }
}

template <bool kSigma>
__global__ __launch_bounds__(256) void k_list_build(ForceArgs a) {
    extern __shared__ float4 dyn[];
    __shared__ int s_cstart[kMaxSlots];
    __shared__ int s_off[kMaxSlots + 1];
    PairLds L;
    L.ent = dyn;
    L.typ = nullptr;
    L.sp = nullptr;
    L.cstart = s_cstart;
    L.off = s_off;
    L.pt = nullptr;  // the list radius comes from the replica (rc_list)
    L.red = nullptr;
    // list buffer after the staging and its read headroom of 3 cells
    uint8_t *buf = reinterpret_cast<uint8_t *>(dyn + a.lds_cap + 3 * a.cell_cap);
    // the ranking histogram after the list buffer (list_build_lds_bytes)
    int *hist = reinterpret_cast<int *>(buf + (size_t)a.nbr_cap * (blockDim.x + 4));
    if (a.ctl && a.any && !*a.any) return;  // gated, and no replica rebuilds
    // gated batches walk the listed tiles of the rebuilding replicas
    const int nt = a.ftile ? *a.nft : a.ntiles;
    for (int b = blockIdx.x; b < nt; b += gridDim.x) {
        const Tile T = a.tiles[a.ftile ? a.ftile[xcd_remap(b, nt)] : xcd_remap(b, nt)];
        if (a.ctl && !a.ftile && !a.ctl[T.rep].flag) continue;  // uniform per workgroup
        const RepDesc R = a.reps[T.rep];
        const int W = T.w_flags & 0xffff;
        if (stage_tile<false, false>(a, T, R, W, L, nullptr, nullptr) < 0) {
            if (threadIdx.x == 0) atomicOr(a.err, 2u);  // FM_DEV_TILE_OVERFLOW
            __syncthreads();
            continue;
        }
        const bool wrapT = (T.w_flags >> 16) & 1;
        if (a.lbuf) {
            if (wrapT) list_tile_body_buf<true, kSigma>(a, T, R, W, L, buf, hist);
            else list_tile_body_buf<false, kSigma>(a, T, R, W, L, buf, hist);
        } else {
            if (wrapT) list_tile_body_direct<true, kSigma>(a, T, R, W, L);
            else list_tile_body_direct<false, kSigma>(a, T, R, W, L);
        }
        __syncthreads();  // the next tile restages the LDS
    }
}

// LJ over the neighbour lists: one lane per particle of a kick span, the
// pair maths and accumulation of pair_tile_body (FP32 cutoff predicate on
// the current positions, so the interacting set is the same as the cell
// walk's whenever the list is complete).
// LJ over the neighbour lists, one workgroup per tile of the list build: the
// tile's current positions (and sigmas) are staged in LDS in the build's
// layout, each lane takes a centre particle and gathers its listed
// neighbours from LDS (the pair maths and accumulation of pair_tile_body;
// the FP32 cutoff predicate on the current positions, so the interacting
// set is the cell walk's whenever the list is complete).
// Energies and virials (kEnergy) from two FP32 sums per particle, three
// vector instructions per list entry: T = sum_hit tv with tv = s12 - s6 / 2
// (the force's own factor) and Q = sum over all evaluated entries of
// (hit ? s6 : kq), kq = shift / (2 eps).  With A = sum s12, B = sum s6 and
// n interacting pairs, T = A - B / 2 and Q = B + kq (L - n) for L evaluated
// entries, so U_i = 4 eps (A - B) + n shift = 4 eps T - 2 eps (Q - kq L)
// and W_i = 48 eps (A - B / 2) = 48 eps T (femmd2D/src/interactionsCuda.cu:
// 91-96).  kCount (eps = 0, or the dump): n is counted instead.
template <int kWrap, bool kSigma, bool kTypes, bool kEnergy, bool kUnit, bool kDebug, bool kCount,
          int kKick>
__device__ void list_tile_pass_body(const ForceArgs &a, const RepDesc &R, int rep, int W, int tile,
                                    const PairLds &L, const float2 *lp, const float *ls,
                                    const int *lid, const int *lt) {
    const int row = W + 2;
    const int c_lo = L.off[row + 1], c_hi = L.off[row + W + 1];
    const int nc = c_hi - c_lo;
    const int sbase = L.cstart[row + 1];
    const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
    const PairEntry pe = L.pt[0];
    const float rc2_u = R.rc2_uniform;
    // (not used when kCount: eps may be 0 there; nor with a type table,
    // whose pairs accumulate their own eps-weighted terms)
    const float kq = (kCount || kTypes) ? 0.0f : 2.0f * pe.shift / pe.e4;
    const int nt = R.ntypes;
    const uint32_t Nu = (uint32_t)a.n;
    const int capg = a.nbr_cap >> 2;
    const uint64_t *nb = reinterpret_cast<const uint64_t *>(a.nbr);
    double Ublk = 0.0, Wblk = 0.0;
    TileKick tk;
    for (int l = threadIdx.x; l < nc; l += blockDim.x) {
        // list slot q holds the list of the tile's particle k (slots of a
        // chunk of blockDim particles are ordered longest list first)
        const int q = sbase + l;
        const int nq = a.nnum[q];
        const int k = nq >> 16;
        const int me = c_lo + k, p = sbase + k;
        const float2 pi = lp[me];
        const float xi = pi.x, yi = pi.y;
        const float xim = xi - Lx, yim = yi - Ly;
        const float si = kSigma ? ls[me] : 0.0f;
        const int ti = kTypes ? lt[me] : 0;
        const int n4 = ((nq & kListLen) + 3) >> 2;  // padded with the sentinel slot
        // in a tile that needs the minimum image, only the wavefronts
        // holding a list flagged for it take the image path (the build
        // ranks those lists after the others); elsewhere the shift is 0
        const bool wimg = kWrap != 0 && __ballot((nq & kListImg) != 0) != 0ull;
        // single type: T = sum_hit tv, Q as described above; with a type
        // table: T = sum_hit e4 (s12 - s6), Q = sum_hit e48 tv, ush = sum_hit
        // shift (femmd2D/src/interactions.cpp:347-389 per type pair)
        float fx = 0.f, fy = 0.f, accT = 0.f, accQ = 0.f, ush = 0.f;
        int nint = 0;
        long long d1 = 0, d2 = 0;  // diagnostic dump: sums of (id + 1), (id + 1)^2
        // one list entry, without branches: the force of a non-interacting
        // entry (or of the padding entry) is computed and discarded by a
        // select
        // wrapc: the axes that may need the minimum image (bit 0 x, bit 1
        // y); an axis left out has no shift in this tile, and its plain
        // difference has the bits the image path would give
        auto term = [&](auto wrapc, float2 pj, float sj, int tj, int jo) {
            constexpr int kW = decltype(wrapc)::value;
            constexpr bool kWX = (kW & 1) != 0, kWY = (kW & 2) != 0;
            const float xj = pj.x, yj = pj.y;
            const float ex = xi - xj, ey = yi - yj;
            float ox = 0.0f, oy = 0.0f;
            if (kWX) {  // minimum image as two selects per axis
                ox = (ex >= hLx) ? Lx : 0.0f;
                ox = (ex < -hLx) ? -Lx : ox;
            }
            if (kWY) {
                oy = (ey >= hLy) ? Ly : 0.0f;
                oy = (ey < -hLy) ? -Ly : oy;
            }
            const float dx = ex - ox;
            const float dy = ey - oy;
            const float dr2 = dx * dx + dy * dy;
            float s2ij, rc2, e48 = 0.0f, e4 = 0.0f, shift = 0.0f;
            if (kTypes) {
                const PairEntry q = L.pt[ti * nt + tj];
                e48 = q.e48;
                e4 = q.e4;
                shift = q.shift;
                const float sij = kSigma ? (si + sj) * 0.5f : R.s_uniform;
                rc2 = q.rc * q.rc * sij * sij;
                s2ij = sij * sij;
            } else if (kSigma) {
                const float sij = (si + sj) * 0.5f;
                rc2 = pe.rc * pe.rc * sij * sij;
                s2ij = sij * sij;
            } else {
                rc2 = rc2_u;
                s2ij = R.s2_uniform;
            }
            // femmd2D/src/interactionsCuda.cu:88-90 (a listed entry is never
            // the particle itself: the padding is the far sentinel slot)
            const bool hit = dr2 < rc2;
            float dxa = dx, dya = dy, r2a = dr2;
            if (kW) {
                // image-exact difference: the far coordinate is shifted by L
                // before subtracting (same bits as dx where no shift applies)
                if (kWX) dxa = (ox > 0.0f) ? (xim - xj) : ((ox < 0.0f) ? (xi - (xj - Lx)) : dx);
                if (kWY) dya = (oy > 0.0f) ? (yim - yj) : ((oy < 0.0f) ? (yi - (yj - Ly)) : dy);
                r2a = dxa * dxa + dya * dya;
            }
            const float inv = __builtin_amdgcn_rcpf(r2a);
            const float sr2 = kUnit ? inv : s2ij * inv;  // sigma^2 = 1: the same bits
            const float sr4 = sr2 * sr2;
            const float sr6 = sr4 * sr2;
            const float s12 = sr6 * sr6;
            const float tv = __builtin_fmaf(-0.5f, sr6, s12);
            // (a non-interacting entry has r2a > 0, so its rcp is finite)
            const float tvh = hit ? tv : 0.0f;
            const float fc = kTypes ? e48 * tvh * inv : tvh * inv;
            fx = __builtin_fmaf(fc, dxa, fx);
            fy = __builtin_fmaf(fc, dya, fy);
            if (kCount) nint += hit ? 1 : 0;
            if (kEnergy && kTypes) {
                accT = __builtin_fmaf(e4, hit ? s12 - sr6 : 0.0f, accT);
                accQ = __builtin_fmaf(e48, tvh, accQ);
                ush += hit ? shift : 0.0f;
            } else if (kEnergy) {
                accT += tvh;
                if (!kCount) accQ += hit ? sr6 : kq;
                else accQ += hit ? sr6 : 0.0f;
            }
            if (kDebug && hit) {
                const long long v = (long long)lid[jo >> 3] + 1;
                d1 += v;
                d2 += v * v;
            }
        };
        auto group = [&](auto wrapc, uint64_t cur) {
            const int j0 = (int)(cur & 0xffffu), j1 = (int)((cur >> 16) & 0xffffu);
            const int j2 = (int)((cur >> 32) & 0xffffu), j3 = (int)(cur >> 48);
            // entries are byte offsets of float2 staging slots
            const char *lpb = reinterpret_cast<const char *>(lp);
            const float2 q0 = *reinterpret_cast<const float2 *>(lpb + j0);
            const float2 q1 = *reinterpret_cast<const float2 *>(lpb + j1);
            const float2 q2 = *reinterpret_cast<const float2 *>(lpb + j2);
            const float2 q3 = *reinterpret_cast<const float2 *>(lpb + j3);
            float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
            if (kSigma) {
                s0 = ls[j0 >> 3];
                s1 = ls[j1 >> 3];
                s2 = ls[j2 >> 3];
                s3 = ls[j3 >> 3];
            }
            int t0 = 0, t1 = 0, t2 = 0, t3 = 0;
            if (kTypes) {
                t0 = lt[j0 >> 3];
                t1 = lt[j1 >> 3];
                t2 = lt[j2 >> 3];
                t3 = lt[j3 >> 3];
            }
            term(wrapc, q0, s0, t0, j0);
            term(wrapc, q1, s1, t1, j1);
            term(wrapc, q2, s2, t2, j2);
            term(wrapc, q3, s3, t3, j3);
        };
        // the next group is loaded while this one is evaluated (the group
        // index is clamped inside the array; a group past the list is loaded
        // and not used); the empty asm keeps the compiler from moving the
        // load down to its use in the next iteration
        auto groups = [&](auto wrapc) {
            uint64_t gv = nb[(uint32_t)q];
            for (int g = 0; g < n4; ++g) {
                const uint64_t cur = gv;
                gv = nb[(uint32_t)min(g + 1, capg - 1) * Nu + (uint32_t)q];
                __asm__ volatile("" ::: "memory");
                group(wrapc, cur);
            }
        };
        if (wimg) groups(std::integral_constant<int, kWrap>{});
        else groups(std::integral_constant<int, 0>{});
        const int id = a.perm[p];
        if (kTypes)
            tile_kick<kKick, true, kEnergy>(a, R, id, p, xi, yi, fx, fy, tk);
        else
            tile_kick<kKick, true, kEnergy>(a, R, id, p, xi, yi, fx * pe.e48, fy * pe.e48, tk);
        if (kEnergy && kTypes) {
            Ublk += 0.5 * ((double)accT + (double)ush);
            Wblk += 0.5 * (double)accQ;
        } else if (kEnergy) {
            const double T = accT, Q = accQ, e4 = pe.e4;
            // kCount: Q = B exactly; otherwise Q - kq L = B - kq n
            const double Ui = kCount ? e4 * (T - 0.5 * Q) + (double)nint * pe.shift
                                     : e4 * T - 0.5 * e4 * (Q - (double)kq * (4.0 * n4));
            Ublk += 0.5 * Ui;
            Wblk += 0.5 * ((double)pe.e48 * T);
        }
        if (kDebug) {
            const int u = a.uid[id];
            a.dbg_cnt[u] = nint;
            a.dbg_s1[u] = d1;
            a.dbg_s2[u] = d2;
        }
    }
    if (kEnergy || kKick == 2) {
        tile_reduce(Ublk, Wblk, tk, L);
        tile_partials<kKick, true, kEnergy>(a, tile, rep, Ublk, Wblk, tk);
    }
}

template <bool kSigma, bool kTypes, bool kEnergy, bool kDebug, int kKick>
__device__ void list_tile_pass(const ForceArgs &a, int tile, PairLds &L, float2 *lp,
                               float *ls, int *lid, int *lt) {
    const Tile T = a.tiles[tile];
    const RepDesc R = a.reps[T.rep];
    const int W = T.w_flags & 0xffff;
    if (stage_layout(a, T, R, W, L) < 0) {
        atomicOr(a.err, 2u);  // FM_DEV_TILE_OVERFLOW (its particles are not moved)
        tile_partials<kKick, true, kEnergy>(a, tile, T.rep, 0.0, 0.0, TileKick{});
        return;
    }
    stage_copy(
        T, R, W, L,
        [&](int p) {
            const float2 q = a.xys[p];
            return make_float4(q.x, q.y, kSigma ? a.sigs[p] : 0.0f,
                               kDebug ? __int_as_float(a.uid[a.perm[p]]) : 0.0f);
        },
        [&](int e, const float4 &u) {
            lp[e] = make_float2(u.x, u.y);
            if (kSigma) ls[e] = u.z;
            if (kDebug) lid[e] = __float_as_int(u.w);  // caller id (diagnostic dump)
        });
    if (kTypes)  // the types in the same staging order (sorted positions)
        stage_copy(
            T, R, W, L, [&](int p) { return make_float4(__int_as_float(a.sorted_typ[p]), 0.f, 0.f, 0.f); },
            [&](int e, const float4 &u) { lt[e] = __float_as_int(u.x); });
    if (threadIdx.x == 0) {  // the padding entries' slot, far from any particle
        lp[a.lds_cap] = make_float2(kSentinel, kSentinel);
        if (kSigma) ls[a.lds_cap] = 1.0f;
        if (kDebug) lid[a.lds_cap] = -1;
        if (kTypes) lt[a.lds_cap] = 0;
    }
    __syncthreads();
    // the axes whose minimum image the tile's lists may need: x alone for
    // a tile at the left or right edge away from the top and bottom rows
    // (it needs no y image), else both (bit 0 x, bit 1 y)
    const int wm = ((T.w_flags >> 17) & 3) == 1 ? 1 : (((T.w_flags >> 16) & 1) ? 3 : 0);
#define FM_LPB(WR, U, C) \
    list_tile_pass_body<WR, kSigma, kTypes, kEnergy, U, kDebug, C, kKick>(a, R, T.rep, W, tile, L, lp, ls, lid, lt)
    // hits are counted for the dump, and for the energy when eps = 0
    if (kDebug || (!kTypes && kEnergy && L.pt[0].e4 == 0.0f)) {
        if (wm) FM_LPB(3, false, true);
        else FM_LPB(0, false, true);
    } else if (!kSigma && R.s2_uniform == 1.0f) {  // (sigma^2 = 1: the same bits)
        switch (wm) {
            case 1: FM_LPB(1, true, false); break;
            case 3: FM_LPB(3, true, false); break;
            default: FM_LPB(0, true, false);
        }
    } else {
        switch (wm) {
            case 1: FM_LPB(1, false, false); break;
            case 3: FM_LPB(3, false, false); break;
            default: FM_LPB(0, false, false);
        }
    }
#undef FM_LPB
}

// CST element forces (femmd2D/src/interactions.cpp:532-689), re-derived in
// displacement form: with U = Ds - Dm and G = U Bm = F - I, every term of P
// is O(|G|), so FP32 keeps relative accuracy at small strain.  Corner forces
// go to their entries of the per-particle gather order, cf[gpos[3e + k]]; no
// atomics.
__device__ void elem_block(const ForceArgs &a, int b, double *red) {
    const Span sp = a.espans[b];
    const RepDesc R = a.reps[sp.rep];
    double U = 0.0, Wv = 0.0;
    for (int e = sp.begin + (int)threadIdx.x; e < sp.end; e += blockDim.x) {
        const int n0 = a.t0[e], n1 = a.t1[e], n2 = a.t2[e];
        const float x0 = a.x[n0], y0 = a.y[n0];
        const float x1 = a.x[n1], y1 = a.y[n1];
        const float x2 = a.x[n2], y2 = a.y[n2];
        // half-box unwrap of nodes 1, 2 against node 0
        // (femmd2D/src/interactions.cpp:554-568): shifts s1, s2 in {-1,0,1}
        const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
        const float dx01 = x0 - x1, dx02 = x0 - x2, dy01 = y0 - y1, dy02 = y0 - y2;
        const int s1x = (dx01 > hLx) - (dx01 < -hLx);
        const int s2x = (dx02 > hLx) - (dx02 < -hLx);
        const int s1y = (dy01 > hLy) - (dy01 < -hLy);
        const int s2y = (dy02 > hLy) - (dy02 < -hLy);
        const float Ds00 = shifted_diff(x0, x2, -s2x, Lx);
        const float Ds10 = shifted_diff(y0, y2, -s2y, Ly);
        const float Ds01 = shifted_diff(x1, x2, s1x - s2x, Lx);
        const float Ds11 = shifted_diff(y1, y2, s1y - s2y, Ly);
        const float4 Dm = a.dmbm[2 * e];
        const float4 B = a.dmbm[2 * e + 1];
        const float2 alm = a.alm[e];
        const float al = alm.x, am = alm.y;
        // Dm = {Dm00, Dm01, Dm10, Dm11}; B = {B00, B01, B10, B11}
        const float U00 = Ds00 - Dm.x, U01 = Ds01 - Dm.y;
        const float U10 = Ds10 - Dm.z, U11 = Ds11 - Dm.w;
        const float G00 = U00 * B.x + U01 * B.z;
        const float G10 = U10 * B.x + U11 * B.z;
        const float G01 = U00 * B.y + U01 * B.w;
        const float G11 = U10 * B.y + U11 * B.w;
        const float F00 = 1.0f + G00, F11 = 1.0f + G11;
        float P00, P01, P10, P11, ue;
        const float trG = G00 + G11;
        const float detG = G00 * G11 - G01 * G10;
        const float tJ = trG + detG;  // J - 1
        if (!(tJ > -1.0f)) atomicOr(a.err, 4u);  // FM_DEV_INVERTED
        if (a.model == 0) {
            // Neo-Hookean: P = [mu (J F - cof F) + lambda lnJ cof F] / J
            const float lJ = log1pf(tJ);
            const float iJ = 1.0f / (1.0f + tJ);
            const float A00 = (G00 - G11) + tJ * F00;
            const float A01 = G01 + tJ * G01 + G10;
            const float A10 = G10 + tJ * G10 + G01;
            const float A11 = (G11 - G00) + tJ * F11;
            const float ll = al * lJ;
            P00 = (am * A00 + ll * F11) * iJ;
            P01 = (am * A01 - ll * G10) * iJ;
            P10 = (am * A10 - ll * G01) * iJ;
            P11 = (am * A11 + ll * F00) * iJ;
            const float g2 = G00 * G00 + G01 * G01 + G10 * G10 + G11 * G11;
            ue = am * (trG - lJ) + 0.5f * am * g2 + 0.5f * al * lJ * lJ;
        } else {
            // SVK: E = (G + G^T + G^T G)/2, S = 2 mu E + lambda tr(E) I, P = F S
            const float E00 = G00 + 0.5f * (G00 * G00 + G10 * G10);
            const float E11 = G11 + 0.5f * (G01 * G01 + G11 * G11);
            const float E01 = 0.5f * (G01 + G10) + 0.5f * (G00 * G01 + G10 * G11);
            const float tr = E00 + E11;
            const float S00 = 2.0f * am * E00 + al * tr;
            const float S11 = 2.0f * am * E11 + al * tr;
            const float S01 = 2.0f * am * E01;
            P00 = F00 * S00 + G01 * S01;
            P01 = F00 * S01 + G01 * S11;
            P10 = G10 * S00 + F11 * S01;
            P11 = G10 * S01 + F11 * S11;
            ue = am * (E00 * E00 + E11 * E11 + 2.0f * E01 * E01) + 0.5f * al * tr * tr;
        }
        // H = -P' Bm^T (P' = A0 P, A0 folded into al, am)
        const float f0x = -(P00 * B.x + P01 * B.y);
        const float f0y = -(P10 * B.x + P11 * B.y);
        const float f1x = -(P00 * B.z + P01 * B.w);
        const float f1y = -(P10 * B.z + P11 * B.w);
        a.cf[a.gpos[3 * e]] = make_float2(f0x, f0y);
        a.cf[a.gpos[3 * e + 1]] = make_float2(f1x, f1y);
        a.cf[a.gpos[3 * e + 2]] = make_float2(-f0x - f1x, -f0y - f1y);
        U += (double)ue;
        Wv += (double)(f0x * Ds00 + f0y * Ds10) + (double)(f1x * Ds01 + f1y * Ds11);
    }
    block_sum2(U, Wv, red);
    if (threadIdx.x == 0) {
        a.fea_part[2 * b] = U;
        a.fea_part[2 * b + 1] = Wv;
    }
}

// Harmonic bonds (femmd2D/src/interactions.cpp:143-173, per-bond k, r0).
__device__ void bond_block(const ForceArgs &a, int b, double *red) {
    const Span sp = a.bspans[b];
    const RepDesc R = a.reps[sp.rep];
    double U = 0.0, Wv = 0.0;
    for (int l = sp.begin + (int)threadIdx.x; l < sp.end; l += blockDim.x) {
        const int i = a.bi[l], j = a.bj[l];
        const float xi = a.x[i], yi = a.y[i], xj = a.x[j], yj = a.y[j];
        float dx = xi - xj, dy = yi - yj;
        const int sx = (dx >= R.hLx) - (dx < -R.hLx);
        const int sy = (dy >= R.hLy) - (dy < -R.hLy);
        dx = shifted_diff(xi, xj, -sx, R.Lx);
        dy = shifted_diff(yi, yj, -sy, R.Ly);
        // r - r0 = (r^2 - r0^2) / (r + r0) with r^2 - r0^2 formed from
        // error-free products and sums, so a bond near rest keeps its
        // relative accuracy (r - r0 in FP32 would cancel to nothing)
        const float p1 = dx * dx, e1 = __builtin_fmaf(dx, dx, -p1);
        const float p2 = dy * dy, e2 = __builtin_fmaf(dy, dy, -p2);
        const float r2 = p1 + p2;
        const float bb = r2 - p1;
        const float t2 = (p1 - (r2 - bb)) + (p2 - bb);
        const float r0 = a.br0[l];
        const float h0 = r0 * r0, l0 = __builtin_fmaf(r0, r0, -h0);
        const float q = (r2 - h0) + ((t2 + e1 + e2) - l0);
        const float r = sqrtf(r2);
        const float k = a.bk[l];
        const float d = q / (r + r0);
        const float kdr = -k * d / r;
        const float fx = kdr * dx, fy = kdr * dy;
        a.cf[a.gpos[a.cf_bond0 + 2 * l]] = make_float2(fx, fy);
        a.cf[a.gpos[a.cf_bond0 + 2 * l + 1]] = make_float2(-fx, -fy);
        U += 0.5 * (double)(k * d * d);
        Wv += (double)(kdr * r2);
    }
    block_sum2(U, Wv, red);
    if (threadIdx.x == 0) {
        a.bond_part[2 * b] = U;
        a.bond_part[2 * b + 1] = Wv;
    }
}

// The force pass (femmd2D/src/interactions.cpp:83-109).  kKick: 0 forces
// only; 1 forces and the closing second half-kick of a run; 2 forces, the
// second half-kick of this step and the first half-kick, drift and wrap of
// the next (femmd2D/src/integrator.cpp:8-29), for every particle without
// element or bond corners (tile_kick).
// (8 wavefronts per SIMD: measured faster than the 6 its registers would
// otherwise allow, at the cost of one spilled register)
template <bool kSigma, bool kTypes, bool kDebug, bool kList, bool kEnergy, int kKick>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_forces(ForceArgs a) {
    extern __shared__ float4 dyn[];
    __shared__ int s_cstart[kMaxSlots];
    __shared__ int s_off[kMaxSlots + 1];
    __shared__ PairEntry s_pt[kMaxPairTypes * kMaxPairTypes];
    __shared__ double s_red[3 * 4];
    __shared__ float s_redf[2 * 4];
    int b = blockIdx.x;
    // the previous step's observables, level 2 (one thread per replica)
    if (a.obs_l2 && threadIdx.x == 0)
        for (int r = blockIdx.x; r < a.obs.nrep; r += gridDim.x) obs_finish(a.obs, r, true);
    // the rebuild of this step (if any) is complete: clear its summary,
    // fold this step into each replica's displacement bound (zero after a
    // rebuild) and clear the displacement slots for the next kick
    // (the diagnostic dump pass leaves the rebuild state alone)
    if (kList && !kDebug) {
        if (b == 0 && a.any && threadIdx.x == 0) *a.any = 0;
        // replica r by workgroup r (the first wavefront): bound, then slots
        if (threadIdx.x < 64) {
            for (int r = blockIdx.x; r < a.nrep; r += gridDim.x) {
                const float d = list_disp(a.dmax, a.ctl_w, r, a.reps[r]);
                if (threadIdx.x == 0) a.ctl_w[r].disp = a.ctl_w[r].flag ? 0.0f : d;
                if (threadIdx.x < kDispSlots)
                    a.dmax[((size_t)r * kDispSlots + threadIdx.x) * kDispStride] = 0u;
            }
        }
    }
    // element and bond workgroups first (a whole number of rounds over the
    // 8 XCDs, so each XCD's tiles still form one contiguous band), then the
    // tiles: the kernel's tail is made of the uniform tile work
    {
        const int ne = a.nespans + a.nbspans;
        const int nfirst = (ne + 7) & ~7;
        if (b < nfirst) {
            if (b < a.nespans) elem_block(a, b, s_red);
            else if (b < ne) bond_block(a, b - a.nespans, s_red);
            return;
        }
        b -= nfirst;
    }
    if (kList) {
        if (b < a.ntiles) {
            PairLds L;
            L.ent = nullptr;
            L.typ = nullptr;
            L.sp = nullptr;
            L.cstart = s_cstart;
            L.off = s_off;
            L.pt = s_pt;
            L.red = s_red;
            L.redf = s_redf;
            float2 *lp = reinterpret_cast<float2 *>(dyn);
            float *ls = reinterpret_cast<float *>(lp + a.lds_cap + 1);
            int *lid = reinterpret_cast<int *>(ls + (kSigma ? a.lds_cap + 1 : 0));
            int *lt = lid + (kDebug ? a.lds_cap + 1 : 0);
            list_tile_pass<kSigma, kTypes, kEnergy, kDebug, kKick>(a, xcd_remap(b, a.ntiles), L, lp,
                                                                   ls, lid, lt);
        }
    } else if (b < a.ntiles) {
        PairLds L;
        L.ent = dyn;
        int *tail = reinterpret_cast<int *>(dyn + a.lds_cap);
        L.typ = tail;
        L.sp = tail + (kTypes ? a.lds_cap : 0);
        L.cstart = s_cstart;
        L.off = s_off;
        L.pt = s_pt;
        L.red = s_red;
        L.redf = s_redf;
        pair_tile<kSigma, kTypes, kDebug, kEnergy, kKick>(a, xcd_remap(b, a.ntiles), L);
    }
}

// Workgroups of one force pass: element and bond spans (rounded up to a
// multiple of 8), then the tiles.
__host__ __device__ inline int forces_grid(const ForceArgs &a) {
    return ((a.nespans + a.nbspans + 7) & ~7) + a.ntiles;
}

// Total force per particle (pair + gathered corners), internal order.
__global__ __launch_bounds__(256) void k_total_force(int n, const float *fx,
                                                     const float *fy,
                                                     const float2 *cf,
                                                     const int32_t *goff,
                                                     const int32_t *gent,
                                                     float *ox, float *oy) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float x = fx[p], y = fy[p];
    if (goff) {
        // gent null: cf is held in gather order
        for (int e = goff[p]; e < goff[p + 1]; ++e) {
            const float2 c = cf[gent ? gent[e] : e];
            x += c.x;
            y += c.y;
        }
    }
    ox[p] = x;
    oy[p] = y;
}

// Observables of the current step for every replica, with the same chunks
// and order as the per-step path: level 1 (one workgroup per chunk), then
// level 2 (one thread per replica); accumulate: add them to the running sums.
__global__ __launch_bounds__(256) void k_observe_l1(ObsArgs o) {
    __shared__ double red[kObsRed];
    obs_level1(o, red);
}
__global__ __launch_bounds__(64) void k_observe_l2(ObsArgs o, int accumulate) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < o.nrep) obs_finish(o, r, accumulate != 0);
}

hipError_t launch_observe(const ObsArgs &o, bool accumulate, hipStream_t s) {
    if (o.nrep <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_observe_l1, dim3(std::min(o.nrep * o.nchunk, 4096)), dim3(256), 0, s,
                       o);
    hipLaunchKernelGGL(k_observe_l2, dim3((o.nrep + 63) / 64), dim3(64), 0, s, o,
                       accumulate ? 1 : 0);
    return hipGetLastError();
}

// Start a new list epoch (after a forced rebuild), which is counted.
// (nslots displacement slots are cleared: every set of the pair)
__global__ void k_ctl_reset(Ctl *ctl, int nrep, uint32_t *dmax, int nslots) {
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) {
        ctl[r].flag = 0;
        ctl[r].disp = 0.0f;
        ctl[r].nbuild += 1u;
    }
    for (int k = threadIdx.x; k < nslots; k += blockDim.x) dmax[(size_t)k * kDispStride] = 0u;
}

hipError_t launch_ctl_reset(Ctl *ctl, int nrep, uint32_t *dmax, int nslots, hipStream_t s) {
    if (nrep <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(256), 0, s, ctl, nrep, dmax, nslots);
    return hipGetLastError();
}

// Zero an int array (graph-safe replacement of a memset node).
__global__ void k_fill_int(int32_t *p, int32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ===========================================================================
// Adjoint pass (no reference counterpart).  With f_i = sum_j g(s) d_ij,
// s = |d_ij|^2, the directional derivative of the pair forces is
//   (J w)_i = sum_j [ g (w_i - w_j) + 2 g'(s) d_ij (d_ij . (w_i - w_j)) ]
// (the full-list loop of the forward pass, with w staged beside x).  For
// LJ, with u = sigma^2 / s, s6 = u^3, s12 = s6^2:
//   g = 48 eps (s12 - s6/2) / s,  g' = 48 eps (2 s6 - 7 s12) / s^2,
//   dg/dsigma = 48 eps (12 s12 - 3 s6) / (s sigma),  dg/deps = g / eps.
// ===========================================================================
template <bool kWrap, bool kSigma>
__device__ void adj_pair_body(const AdjArgs &A, const RepDesc &R, int W, int tile, PairLds &L,
                              const float2 *lw) {
    const ForceArgs &a = A.f;
    const int row = W + 2;
    const int c_lo = L.off[row + 1], c_hi = L.off[row + W + 1];
    const int nc = c_hi - c_lo;
    const int sbase = L.cstart[row + 1];
    const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
    const PairEntry pe = L.pt[0];
    const unsigned rc2b_u = __float_as_uint(R.rc2_uniform);
    double peps = 0.0, psig = 0.0;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
        const int me = c_lo + k;
        int lo = row + 1, hi = row + W + 1;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (L.off[mid] <= me) lo = mid; else hi = mid;
        }
        const int col = lo - row;
        const float4 pi = L.ent[me];
        const float2 wi = lw[me];
        const float xi = pi.x, yi = pi.y, xim = xi - Lx, yim = yi - Ly;
        const int gi = __float_as_int(pi.z);
        const int gix = gi >= 0 ? gi : -2;
        const float si = pi.w;
        const int a0 = L.off[col - 1], b0 = L.off[col + 2];
        const int a1 = L.off[row + col - 1], b1 = L.off[row + col + 2];
        const int a2 = L.off[2 * row + col - 1], b2 = L.off[2 * row + col + 2];
        const int total = (b0 - a0) + (b1 - a1) + (b2 - a2);
        int j = a0;
        if (j == b0) j = a1;
        if (j == b1) j = a2;
        float jx = 0.f, jy = 0.f, fx = 0.f, fy = 0.f, gx = 0.f, gy = 0.f;
        // the next candidate's LDS entry is read one iteration ahead
        float4 pn = L.ent[total > 0 ? j : 0];
        for (int t = 0; t < total; ++t) {
            const float4 pj = pn;
            int jn = j + 1;
            if (jn == b0) jn = a1;
            if (jn == b1) jn = a2;
            pn = L.ent[(t + 1 < total) ? jn : j];
            const int gj = __float_as_int(pj.z);
            float dx = xi - pj.x;
            float dy = yi - pj.y;
            int sx = 0, sy = 0;
            if (kWrap) {
                if (dx >= hLx) { dx -= Lx; sx = 1; } else if (dx < -hLx) { dx += Lx; sx = -1; }
                if (dy >= hLy) { dy -= Ly; sy = 1; } else if (dy < -hLy) { dy += Ly; sy = -1; }
            }
            const float dr2 = dx * dx + dy * dy;
            float s2ij, sij;
            unsigned rc2b;
            if (kSigma) {
                sij = (si + pj.w) * 0.5f;
                rc2b = __float_as_uint(pe.rc * pe.rc * sij * sij);
                s2ij = sij * sij;
            } else {
                sij = R.s_uniform;
                rc2b = rc2b_u;
                s2ij = R.s2_uniform;
            }
            const bool hit = (__float_as_uint(dr2) < rc2b) & (j != me) & (gj != gix);
            if (hit) {
                float dxa = dx, dya = dy, r2a = dr2;
                if (kWrap) {
                    if (sx > 0) dxa = xim - pj.x; else if (sx < 0) dxa = xi - (pj.x - Lx);
                    if (sy > 0) dya = yim - pj.y; else if (sy < 0) dya = yi - (pj.y - Ly);
                    r2a = dxa * dxa + dya * dya;
                }
                const float inv = __builtin_amdgcn_rcpf(r2a);
                const float u = s2ij * inv;
                const float s6 = u * u * u;
                const float s12 = s6 * s6;
                const float g = __builtin_fmaf(-0.5f, s6, s12) * inv;
                const float gp = __builtin_fmaf(2.0f, s6, -7.0f * s12) * inv * inv;
                const float gs = __builtin_fmaf(12.0f, s12, -3.0f * s6) * inv / sij;
                const float2 wj = lw[j];
                const float dwx = wi.x - wj.x, dwy = wi.y - wj.y;
                const float ddw = __builtin_fmaf(dxa, dwx, dya * dwy);
                const float g2 = 2.0f * gp * ddw;
                jx = __builtin_fmaf(g, dwx, __builtin_fmaf(g2, dxa, jx));
                jy = __builtin_fmaf(g, dwy, __builtin_fmaf(g2, dya, jy));
                fx = __builtin_fmaf(g, dxa, fx);
                fy = __builtin_fmaf(g, dya, fy);
                gx = __builtin_fmaf(gs, dxa, gx);
                gy = __builtin_fmaf(gs, dya, gy);
            }
            j = jn;
        }
        const int id = a.perm[sbase + k];
        A.jx[id] = pe.e48 * jx;
        A.jy[id] = pe.e48 * jy;
        // w . f_lj / eps = 48 w . (f / 48 eps); w . df/dsigma
        peps += 48.0 * ((double)wi.x * fx + (double)wi.y * fy);
        psig += (double)pe.e48 * ((double)wi.x * gx + (double)wi.y * gy);
    }
    block_sum2(peps, psig, L.red);
    if (threadIdx.x == 0) {
        A.pacc[2 * tile] += peps;
        A.pacc[2 * tile + 1] += psig;
    }
}

template <bool kSigma>
__device__ void adj_pair_tile(const AdjArgs &A, int tile, PairLds &L, float2 *lw) {
    const ForceArgs &a = A.f;
    const Tile T = a.tiles[tile];
    const RepDesc R = a.reps[T.rep];
    const int W = T.w_flags & 0xffff;
    const bool wrapT = (T.w_flags >> 16) & 1;
    if (stage_tile<false, false>(a, T, R, W, L, A.w, lw) < 0) {
        if (threadIdx.x == 0) atomicOr(a.err, 2u);
        return;
    }
    if (wrapT)
        adj_pair_body<true, kSigma>(A, R, W, tile, L, lw);
    else
        adj_pair_body<false, kSigma>(A, R, W, tile, L, lw);
}

// J w and the parameter terms over the neighbour lists (the backward sweep
// in list mode): the tile's positions at the configuration being
// differentiated and w are staged in the list build's layout (gathered
// through perm), each lane takes a list slot, as the forward list pass.
template <bool kWrap, bool kSigma>
__device__ void adj_list_body(const AdjArgs &A, const RepDesc &R, int W, int tile,
                              const PairLds &L, const float2 *lp, const float2 *lw,
                              const float *ls) {
    const ForceArgs &a = A.f;
    const int row = W + 2;
    const int c_lo = L.off[row + 1], c_hi = L.off[row + W + 1];
    const int nc = c_hi - c_lo;
    const int sbase = L.cstart[row + 1];
    const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
    const PairEntry pe = L.pt[0];
    const unsigned rc2b_u = __float_as_uint(R.rc2_uniform);
    const uint32_t Nu = (uint32_t)a.n;
    const int capg = a.nbr_cap >> 2;
    const uint64_t *nb = reinterpret_cast<const uint64_t *>(a.nbr);
    const char *lpb = reinterpret_cast<const char *>(lp);
    const char *lwb = reinterpret_cast<const char *>(lw);
    double peps = 0.0, psig = 0.0;
    for (int l = threadIdx.x; l < nc; l += blockDim.x) {
        const int q = sbase + l;
        const int nq = a.nnum[q];
        const int k = nq >> 16;
        const int me = c_lo + k, p = sbase + k;
        const float2 pi = lp[me];
        const float2 wi = lw[me];
        const float xi = pi.x, yi = pi.y, xim = xi - Lx, yim = yi - Ly;
        const float si = kSigma ? ls[me] : 0.0f;
        const int n4 = ((nq & kListLen) + 3) >> 2;
        float jx = 0.f, jy = 0.f, fx = 0.f, fy = 0.f, gx = 0.f, gy = 0.f;
        auto term = [&](int off) {
            const float2 pj = *reinterpret_cast<const float2 *>(lpb + off);
            const float ex = xi - pj.x, ey = yi - pj.y;
            float ox = 0.0f, oy = 0.0f;
            if (kWrap) {
                ox = (ex >= hLx) ? Lx : 0.0f;
                ox = (ex < -hLx) ? -Lx : ox;
                oy = (ey >= hLy) ? Ly : 0.0f;
                oy = (ey < -hLy) ? -Ly : oy;
            }
            const float dx = ex - ox, dy = ey - oy;
            const float dr2 = dx * dx + dy * dy;
            float s2ij, sij;
            unsigned rc2b;
            if (kSigma) {
                sij = (si + ls[off >> 3]) * 0.5f;
                rc2b = __float_as_uint(pe.rc * pe.rc * sij * sij);
                s2ij = sij * sij;
            } else {
                sij = R.s_uniform;
                rc2b = rc2b_u;
                s2ij = R.s2_uniform;
            }
            if (__float_as_uint(dr2) < rc2b) {
                float dxa = dx, dya = dy, r2a = dr2;
                if (kWrap) {
                    dxa = (ox > 0.0f) ? (xim - pj.x) : ((ox < 0.0f) ? (xi - (pj.x - Lx)) : dx);
                    dya = (oy > 0.0f) ? (yim - pj.y) : ((oy < 0.0f) ? (yi - (pj.y - Ly)) : dy);
                    r2a = dxa * dxa + dya * dya;
                }
                const float inv = __builtin_amdgcn_rcpf(r2a);
                const float u = s2ij * inv;
                const float s6 = u * u * u;
                const float s12 = s6 * s6;
                const float g = __builtin_fmaf(-0.5f, s6, s12) * inv;
                const float gp = __builtin_fmaf(2.0f, s6, -7.0f * s12) * inv * inv;
                const float gs = __builtin_fmaf(12.0f, s12, -3.0f * s6) * inv / sij;
                const float2 wj = *reinterpret_cast<const float2 *>(lwb + off);
                const float dwx = wi.x - wj.x, dwy = wi.y - wj.y;
                const float ddw = __builtin_fmaf(dxa, dwx, dya * dwy);
                const float g2 = 2.0f * gp * ddw;
                jx = __builtin_fmaf(g, dwx, __builtin_fmaf(g2, dxa, jx));
                jy = __builtin_fmaf(g, dwy, __builtin_fmaf(g2, dya, jy));
                fx = __builtin_fmaf(g, dxa, fx);
                fy = __builtin_fmaf(g, dya, fy);
                gx = __builtin_fmaf(gs, dxa, gx);
                gy = __builtin_fmaf(gs, dya, gy);
            }
        };
        uint64_t gv = nb[(uint32_t)q];
        for (int g = 0; g < n4; ++g) {
            const uint64_t cur = gv;
            gv = nb[(uint32_t)min(g + 1, capg - 1) * Nu + (uint32_t)q];
            __asm__ volatile("" ::: "memory");
            term((int)(cur & 0xffffu));
            term((int)((cur >> 16) & 0xffffu));
            term((int)((cur >> 32) & 0xffffu));
            term((int)(cur >> 48));
        }
        const int id = a.perm[p];
        A.jx[id] = pe.e48 * jx;
        A.jy[id] = pe.e48 * jy;
        peps += 48.0 * ((double)wi.x * fx + (double)wi.y * fy);
        psig += (double)pe.e48 * ((double)wi.x * gx + (double)wi.y * gy);
    }
    block_sum2(peps, psig, L.red);
    if (threadIdx.x == 0) {
        A.pacc[2 * tile] += peps;
        A.pacc[2 * tile + 1] += psig;
    }
}

template <bool kSigma>
__device__ void adj_list_tile(const AdjArgs &A, int tile, PairLds &L, float2 *lp, float2 *lw,
                              float *ls) {
    const ForceArgs &a = A.f;
    const Tile T = a.tiles[tile];
    const RepDesc R = a.reps[T.rep];
    const int W = T.w_flags & 0xffff;
    if (stage_layout(a, T, R, W, L) < 0) {
        if (threadIdx.x == 0) atomicOr(a.err, 2u);  // FM_DEV_TILE_OVERFLOW
        return;
    }
    struct Ent {
        float2 x, w;
        float s;
    };
    stage_copy(
        T, R, W, L,
        [&](int p) {
            const int i = a.perm[p];
            Ent u;
            u.x = make_float2(a.x[i], a.y[i]);
            u.w = A.w[i];
            u.s = kSigma ? a.sigs[p] : 0.0f;
            return u;
        },
        [&](int e, const Ent &u) {
            lp[e] = u.x;
            lw[e] = u.w;
            if (kSigma) ls[e] = u.s;
        });
    if (threadIdx.x == 0) {  // the padding entries' slot
        lp[a.lds_cap] = make_float2(kSentinel, kSentinel);
        lw[a.lds_cap] = make_float2(0.0f, 0.0f);
        if (kSigma) ls[a.lds_cap] = 1.0f;
    }
    __syncthreads();
    if ((T.w_flags >> 16) & 1)
        adj_list_body<true, kSigma>(A, R, W, tile, L, lp, lw, ls);
    else
        adj_list_body<false, kSigma>(A, R, W, tile, L, lp, lw, ls);
}

// Element part of J w: dP (the directional derivative of the first
// Piola-Kirchhoff stress along dF = dDs Bm) pushed through H = -A0 P Bm^T;
// and dL/dE_e += w . f_e / E_e (P is linear in (lambda, mu), both
// proportional to E at fixed nu).
__device__ void adj_elem_block(const AdjArgs &A, int b, double *red) {
    const ForceArgs &a = A.f;
    const Span sp = a.espans[b];
    const RepDesc R = a.reps[sp.rep];
    for (int e = sp.begin + (int)threadIdx.x; e < sp.end; e += blockDim.x) {
        const int n0 = a.t0[e], n1 = a.t1[e], n2 = a.t2[e];
        const float x0 = a.x[n0], y0 = a.y[n0];
        const float x1 = a.x[n1], y1 = a.y[n1];
        const float x2 = a.x[n2], y2 = a.y[n2];
        const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
        const float dx01 = x0 - x1, dx02 = x0 - x2, dy01 = y0 - y1, dy02 = y0 - y2;
        const int s1x = (dx01 > hLx) - (dx01 < -hLx);
        const int s2x = (dx02 > hLx) - (dx02 < -hLx);
        const int s1y = (dy01 > hLy) - (dy01 < -hLy);
        const int s2y = (dy02 > hLy) - (dy02 < -hLy);
        const float Ds00 = shifted_diff(x0, x2, -s2x, Lx);
        const float Ds10 = shifted_diff(y0, y2, -s2y, Ly);
        const float Ds01 = shifted_diff(x1, x2, s1x - s2x, Lx);
        const float Ds11 = shifted_diff(y1, y2, s1y - s2y, Ly);
        const float4 Dm = a.dmbm[2 * e];
        const float4 B = a.dmbm[2 * e + 1];
        const float2 alm = a.alm[e];
        const float al = alm.x, am = alm.y;
        const float G00 = (Ds00 - Dm.x) * B.x + (Ds01 - Dm.y) * B.z;
        const float G10 = (Ds10 - Dm.z) * B.x + (Ds11 - Dm.w) * B.z;
        const float G01 = (Ds00 - Dm.x) * B.y + (Ds01 - Dm.y) * B.w;
        const float G11 = (Ds10 - Dm.z) * B.y + (Ds11 - Dm.w) * B.w;
        const float F00 = 1.0f + G00, F11 = 1.0f + G11, F01 = G01, F10 = G10;
        const float2 w0 = A.w[n0], w1 = A.w[n1], w2 = A.w[n2];
        const float d00 = w0.x - w2.x, d10 = w0.y - w2.y, d01 = w1.x - w2.x, d11 = w1.y - w2.y;
        const float dF00 = d00 * B.x + d01 * B.z;
        const float dF10 = d10 * B.x + d11 * B.z;
        const float dF01 = d00 * B.y + d01 * B.w;
        const float dF11 = d10 * B.y + d11 * B.w;
        float P00, P01, P10, P11, Q00, Q01, Q10, Q11;
        if (a.model == 0) {
            const float tJ = (G00 + G11) + (G00 * G11 - G01 * G10);
            const float lJ = log1pf(tJ);
            const float iJ = 1.0f / (1.0f + tJ);
            const float A00 = (G00 - G11) + tJ * F00;
            const float A01 = G01 + tJ * G01 + G10;
            const float A10 = G10 + tJ * G10 + G01;
            const float A11 = (G11 - G00) + tJ * F11;
            const float ll = al * lJ;
            P00 = (am * A00 + ll * F11) * iJ;
            P01 = (am * A01 - ll * G10) * iJ;
            P10 = (am * A10 - ll * G01) * iJ;
            P11 = (am * A11 + ll * F00) * iJ;
            // F^-T = cof(F) / J
            const float c00 = F11 * iJ, c01 = -F10 * iJ, c10 = -F01 * iJ, c11 = F00 * iJ;
            const float tr = (F11 * dF00 - F01 * dF10 - F10 * dF01 + F00 * dF11) * iJ;
            // M = F^-T dF^T F^-T
            const float q00 = dF00 * c00 + dF10 * c10, q01 = dF00 * c01 + dF10 * c11;
            const float q10 = dF01 * c00 + dF11 * c10, q11 = dF01 * c01 + dF11 * c11;
            const float M00 = c00 * q00 + c01 * q10, M01 = c00 * q01 + c01 * q11;
            const float M10 = c10 * q00 + c11 * q10, M11 = c10 * q01 + c11 * q11;
            const float k1 = am - ll, k2 = al * tr;
            Q00 = am * dF00 + k1 * M00 + k2 * c00;
            Q01 = am * dF01 + k1 * M01 + k2 * c01;
            Q10 = am * dF10 + k1 * M10 + k2 * c10;
            Q11 = am * dF11 + k1 * M11 + k2 * c11;
        } else {
            const float E00 = G00 + 0.5f * (G00 * G00 + G10 * G10);
            const float E11 = G11 + 0.5f * (G01 * G01 + G11 * G11);
            const float E01 = 0.5f * (G01 + G10) + 0.5f * (G00 * G01 + G10 * G11);
            const float trE = E00 + E11;
            const float S00 = 2.0f * am * E00 + al * trE;
            const float S11 = 2.0f * am * E11 + al * trE;
            const float S01 = 2.0f * am * E01;
            P00 = F00 * S00 + F01 * S01;
            P01 = F00 * S01 + F01 * S11;
            P10 = F10 * S00 + F11 * S01;
            P11 = F10 * S01 + F11 * S11;
            // dE = (dF^T F + F^T dF) / 2, dS = 2 mu dE + lambda tr(dE) I
            const float dE00 = dF00 * F00 + dF10 * F10;
            const float dE11 = dF01 * F01 + dF11 * F11;
            const float dE01 = 0.5f * ((dF00 * F01 + dF10 * F11) + (F00 * dF01 + F10 * dF11));
            const float dtr = dE00 + dE11;
            const float dS00 = 2.0f * am * dE00 + al * dtr;
            const float dS11 = 2.0f * am * dE11 + al * dtr;
            const float dS01 = 2.0f * am * dE01;
            Q00 = dF00 * S00 + dF01 * S01 + F00 * dS00 + F01 * dS01;
            Q01 = dF00 * S01 + dF01 * S11 + F00 * dS01 + F01 * dS11;
            Q10 = dF10 * S00 + dF11 * S01 + F10 * dS00 + F11 * dS01;
            Q11 = dF10 * S01 + dF11 * S11 + F10 * dS01 + F11 * dS11;
        }
        const float f0x = -(P00 * B.x + P01 * B.y), f0y = -(P10 * B.x + P11 * B.y);
        const float f1x = -(P00 * B.z + P01 * B.w), f1y = -(P10 * B.z + P11 * B.w);
        const float g0x = -(Q00 * B.x + Q01 * B.y), g0y = -(Q10 * B.x + Q11 * B.y);
        const float g1x = -(Q00 * B.z + Q01 * B.w), g1y = -(Q10 * B.z + Q11 * B.w);
        A.cfa[3 * e] = make_float2(g0x, g0y);
        A.cfa[3 * e + 1] = make_float2(g1x, g1y);
        A.cfa[3 * e + 2] = make_float2(-g0x - g1x, -g0y - g1y);
        const double wf = (double)(w0.x * f0x + w0.y * f0y) + (double)(w1.x * f1x + w1.y * f1y) -
                          (double)(w2.x * (f0x + f1x) + w2.y * (f0y + f1y));
        A.dE[e] += wf / (double)A.eE[e];
    }
    (void)red;
}

// Bond part of J w: g = -k (r - r0) / r, g' = dg/ds = -k r0 / (2 r^3).
__device__ void adj_bond_block(const AdjArgs &A, int b) {
    const ForceArgs &a = A.f;
    const Span sp = a.bspans[b];
    const RepDesc R = a.reps[sp.rep];
    for (int l = sp.begin + (int)threadIdx.x; l < sp.end; l += blockDim.x) {
        const int i = a.bi[l], j = a.bj[l];
        const float xi = a.x[i], yi = a.y[i], xj = a.x[j], yj = a.y[j];
        float dx = xi - xj, dy = yi - yj;
        const int sx = (dx >= R.hLx) - (dx < -R.hLx);
        const int sy = (dy >= R.hLy) - (dy < -R.hLy);
        dx = shifted_diff(xi, xj, -sx, R.Lx);
        dy = shifted_diff(yi, yj, -sy, R.Ly);
        const float r2 = dx * dx + dy * dy;
        const float r = sqrtf(r2);
        const float k = a.bk[l], r0 = a.br0[l];
        const float g = -k * (r - r0) / r;
        const float gp = -k * r0 / (2.0f * r2 * r);
        const float2 wi = A.w[i], wj = A.w[j];
        const float dwx = wi.x - wj.x, dwy = wi.y - wj.y;
        const float g2 = 2.0f * gp * (dx * dwx + dy * dwy);
        const float jx = g * dwx + g2 * dx, jy = g * dwy + g2 * dy;
        A.cfa[a.cf_bond0 + 2 * l] = make_float2(jx, jy);
        A.cfa[a.cf_bond0 + 2 * l + 1] = make_float2(-jx, -jy);
    }
}

template <bool kSigma, bool kList>
__global__ __launch_bounds__(256) void k_adj_forces(AdjArgs A) {
    extern __shared__ float4 dyn[];
    __shared__ int s_cstart[kMaxSlots];
    __shared__ int s_off[kMaxSlots + 1];
    __shared__ PairEntry s_pt[kMaxPairTypes * kMaxPairTypes];
    __shared__ double s_red[2 * 4];
    const ForceArgs &a = A.f;
    int b = blockIdx.x;
    if (b < a.ntiles) {
        PairLds L;
        L.ent = dyn;
        L.typ = nullptr;
        L.sp = nullptr;
        L.cstart = s_cstart;
        L.off = s_off;
        L.pt = s_pt;
        L.red = s_red;
        if (kList) {
            float2 *lp = reinterpret_cast<float2 *>(dyn);
            float2 *lw = lp + a.lds_cap + 1;
            float *ls = reinterpret_cast<float *>(lw + a.lds_cap + 1);
            adj_list_tile<kSigma>(A, xcd_remap(b, a.ntiles), L, lp, lw, ls);
        } else {
            float2 *lw = reinterpret_cast<float2 *>(dyn + a.lds_cap);
            adj_pair_tile<kSigma>(A, xcd_remap(b, a.ntiles), L, lw);
        }
        return;
    }
    b -= a.ntiles;
    if (b < a.nespans) {
        adj_elem_block(A, b, s_red);
        return;
    }
    b -= a.nespans;
    if (b < a.nbspans) adj_bond_block(A, b);
}

__global__ __launch_bounds__(kKickThreads) void k_adj_kick(AdjKickArgs a) {
    const Span sp = a.spans[blockIdx.x];
    const RepDesc R = a.reps[sp.rep];
    const int p = sp.begin + (int)threadIdx.x;
    if (p >= sp.end) return;
    float xbx = a.xbx[p], xby = a.xby[p];
    if (a.accumulate) {
        float jx = a.jx[p], jy = a.jy[p];
        if (a.goff) {
            for (int e = a.goff[p]; e < a.goff[p + 1]; ++e) {
                const float2 c = a.cfa[a.gent[e]];
                jx += c.x;
                jy += c.y;
            }
        }
        xbx += jx;
        xby += jy;
        a.xbx[p] = xbx;
        a.xby[p] = xby;
    }
    float vbx = a.vbx[p], vby = a.vby[p];
    if (a.drift) {
        vbx = __builtin_fmaf(R.dt, xbx, vbx);
        vby = __builtin_fmaf(R.dt, xby, vby);
        a.vbx[p] = vbx;
        a.vby[p] = vby;
    }
    if (a.make_w) {
        const float s = a.c * a.inv_m[p];
        a.w[p] = make_float2(s * vbx, s * vby);
    }
}

__global__ __launch_bounds__(256) void k_adj_loss(LossArgs a) {
    __shared__ double red[8];
    const int k = blockIdx.x * blockDim.x + (int)threadIdx.x;
    double l = 0.0, z = 0.0;
    if (k < a.n) {
        const RepDesc R = a.reps[0];
        const int p = a.ids[k];
        const float2 t = a.target[k];
        float dx = a.x[p] - t.x, dy = a.y[p] - t.y;
        if (dx >= R.hLx) dx -= R.Lx; else if (dx < -R.hLx) dx += R.Lx;
        if (dy >= R.hLy) dy -= R.Ly; else if (dy < -R.hLy) dy += R.Ly;
        a.xbx[p] = a.scale * dx;
        a.xby[p] = a.scale * dy;
        l = (double)dx * dx + (double)dy * dy;
    }
    block_sum2(l, z, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = l;
}

// ---------------------------------------------------------------------------
// Elastomer cluster observables (femmd2D/src/feaElements.cpp:86-196,
// re-designed: one workgroup per cluster, FP64 sums in a fixed order).
// ---------------------------------------------------------------------------

// Sums K doubles over the workgroup in a fixed order; every thread gets the
// totals.  red holds K * (blockDim / 64) doubles.
template <int K>
__device__ void block_sums_all(double (&v)[K], double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[w * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int q = 0; q < nw; ++q) t += red[q * K + k];
        v[k] = t;
    }
    __syncthreads();
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

// Per cluster: the unfolded node positions (femmd2D/src/feaElements.cpp:
// 86-133: a cluster wider than half the box has its nodes beyond L/2 moved
// down by L, and all of it moved up by L if the mean then falls below 0),
// mass-weighted centre of mass and its velocity (:135-149), and the
// centroid as the circular mean of the wrapped positions (:161-193; here
// with the y coordinate's own sums).
__global__ __launch_bounds__(256) void k_cluster_props(ClusterArgs a) {
    __shared__ double red[9 * 4];
    __shared__ float mm[4 * 4];
    const int c = blockIdx.x;
    const RepDesc R = a.reps[a.crep[c]];
    const int b = a.off[c], n = a.off[c + 1] - b;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
    float xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (int k = tid; k < n; k += blockDim.x) {
        const int i = a.nodes[b + k];
        const float x = a.x[i], y = a.y[i];
        xmin = fminf(xmin, x);
        xmax = fmaxf(xmax, x);
        ymin = fminf(ymin, y);
        ymax = fmaxf(ymax, y);
    }
    xmin = wave_min(xmin);
    xmax = wave_max(xmax);
    ymin = wave_min(ymin);
    ymax = wave_max(ymax);
    if (lane == 0) {
        mm[4 * w] = xmin;
        mm[4 * w + 1] = xmax;
        mm[4 * w + 2] = ymin;
        mm[4 * w + 3] = ymax;
    }
    __syncthreads();
    for (int q = 0; q < nw; ++q) {
        xmin = fminf(xmin, mm[4 * q]);
        xmax = fmaxf(xmax, mm[4 * q + 1]);
        ymin = fminf(ymin, mm[4 * q + 2]);
        ymax = fmaxf(ymax, mm[4 * q + 3]);
    }
    const bool fx = (xmax - xmin) > R.hLx, fy = (ymax - ymin) > R.hLy;
    double m1[2] = {0.0, 0.0};
    for (int k = tid; k < n; k += blockDim.x) {
        const int i = a.nodes[b + k];
        const float x = a.x[i], y = a.y[i];
        m1[0] += (double)((fx && x > R.hLx) ? x - R.Lx : x);
        m1[1] += (double)((fy && y > R.hLy) ? y - R.Ly : y);
    }
    block_sums_all<2>(m1, red);
    const float addx = (fx && m1[0] / n < 0.0) ? R.Lx : 0.0f;
    const float addy = (fy && m1[1] / n < 0.0) ? R.Ly : 0.0f;
    const double twopi = 6.283185307179586;
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = tid; k < n; k += blockDim.x) {
        const int i = a.nodes[b + k];
        const float x = a.x[i], y = a.y[i];
        const float ux = ((fx && x > R.hLx) ? x - R.Lx : x) + addx;
        const float uy = ((fy && y > R.hLy) ? y - R.Ly : y) + addy;
        if (a.ux) {
            a.ux[i] = ux;
            a.uy[i] = uy;
        }
        const double m = a.mass[i];
        v[0] += m;
        v[1] += m * ux;
        v[2] += m * uy;
        v[3] += m * a.vx[i];
        v[4] += m * a.vy[i];
        const double tx = twopi * (double)x / (double)R.Lx, ty = twopi * (double)y / (double)R.Ly;
        v[5] += cos(tx);
        v[6] += sin(tx);
        v[7] += cos(ty);
        v[8] += sin(ty);
    }
    block_sums_all<9>(v, red);
    if (tid == 0) {
        const double M = v[0];
        const double pi = 3.141592653589793;
        double* o = a.out + (size_t)c * kClusterOut;
        o[0] = M;
        o[1] = v[1] / M;
        o[2] = v[2] / M;
        o[3] = v[3] / M;
        o[4] = v[4] / M;
        o[5] = (double)R.Lx * (atan2(-v[6], -v[5]) + pi) / twopi;
        o[6] = (double)R.Ly * (atan2(-v[8], -v[7]) + pi) / twopi;
        a.cm[c] = make_float2((float)o[1], (float)o[2]);
    }
}

// Cluster virial (femmd2D/src/interactions.cpp:370-382): over the pairs the
// force pass takes (cell walk, same predicate and exclusions), for pairs
// whose particles both belong to clusters, 1/2 f_ij . (R_ci - R_cj) with the
// minimum image, summed per tile in FP64.
template <bool kSigma, bool kTypes>
__global__ __launch_bounds__(256) void k_cluster_virial(ClusterVirialArgs A) {
    extern __shared__ float4 dyn[];
    __shared__ int s_cstart[kMaxSlots];
    __shared__ int s_off[kMaxSlots + 1];
    __shared__ PairEntry s_pt[kMaxPairTypes * kMaxPairTypes];
    __shared__ double s_red[2 * 4];
    const ForceArgs &a = A.f;
    const int tile = blockIdx.x;
    PairLds L;
    L.ent = dyn;
    int *tail = reinterpret_cast<int *>(dyn + a.lds_cap);
    L.typ = tail;
    L.sp = tail + (kTypes ? a.lds_cap : 0);
    L.cstart = s_cstart;
    L.off = s_off;
    L.pt = s_pt;
    L.red = s_red;
    const Tile T = a.tiles[tile];
    const RepDesc R = a.reps[T.rep];
    const int W = T.w_flags & 0xffff;
    if (stage_tile<kTypes, false, true>(a, T, R, W, L, nullptr, nullptr) < 0) {
        if (threadIdx.x == 0) {
            atomicOr(a.err, 2u);  // FM_DEV_TILE_OVERFLOW
            A.out[tile] = 0.0;
        }
        return;
    }
    const bool wrapT = (T.w_flags >> 16) & 1;
    const int row = W + 2;
    const int c_lo = L.off[row + 1], c_hi = L.off[row + W + 1];
    const int nc = c_hi - c_lo;
    const float Lx = R.Lx, Ly = R.Ly, hLx = R.hLx, hLy = R.hLy;
    const int nt = R.ntypes;
    double acc = 0.0, dummy = 0.0;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
        const int me = c_lo + k;
        const int ci = A.cid[L.sp[me]];
        if (ci < 0) continue;
        int lo = row + 1, hi = row + W + 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (L.off[mid] <= me) lo = mid; else hi = mid;
        }
        const int col = lo - row;
        const float4 pi = L.ent[me];
        const float xi = pi.x, yi = pi.y, xim = xi - Lx, yim = yi - Ly;
        const int gi = __float_as_int(pi.z);
        const int gix = gi >= 0 ? gi : -2;
        const float si = pi.w;
        const int ti = kTypes ? L.typ[me] : 0;
        const float2 Ri = A.cm[ci];
        for (int r = 0; r < 3; ++r) {
            const int j0 = L.off[r * row + col - 1], j1 = L.off[r * row + col + 2];
            for (int j = j0; j < j1; ++j) {
                const int cj = A.cid[L.sp[j]];
                if (cj < 0) continue;
                const float4 pj = L.ent[j];
                float dx = xi - pj.x, dy = yi - pj.y;
                int sx = 0, sy = 0;
                if (wrapT) {
                    if (dx >= hLx) { dx -= Lx; sx = 1; } else if (dx < -hLx) { dx += Lx; sx = -1; }
                    if (dy >= hLy) { dy -= Ly; sy = 1; } else if (dy < -hLy) { dy += Ly; sy = -1; }
                }
                const float dr2 = dx * dx + dy * dy;
                PairEntry pe = L.pt[0];
                if (kTypes) pe = L.pt[ti * nt + L.typ[j]];
                float s2ij;
                unsigned rc2b;
                if (kSigma || kTypes) {
                    const float sij = kSigma ? (si + pj.w) * 0.5f : R.s_uniform;  // (no sigma staged)
                    rc2b = __float_as_uint(pe.rc * pe.rc * sij * sij);
                    s2ij = sij * sij;
                } else {
                    rc2b = __float_as_uint(R.rc2_uniform);
                    s2ij = R.s2_uniform;
                }
                if (!(__float_as_uint(dr2) < rc2b) || j == me || __float_as_int(pj.z) == gix) continue;
                float dxa = dx, dya = dy, r2a = dr2;
                if (sx > 0) dxa = xim - pj.x; else if (sx < 0) dxa = xi - (pj.x - Lx);
                if (sy > 0) dya = yim - pj.y; else if (sy < 0) dya = yi - (pj.y - Ly);
                if (sx != 0 || sy != 0) r2a = dxa * dxa + dya * dya;
                const float inv = __builtin_amdgcn_rcpf(r2a);
                const float sr2 = s2ij * inv;
                const float sr6 = sr2 * sr2 * sr2;
                const float tv = __builtin_fmaf(-0.5f, sr6, sr6 * sr6);
                const float fc = pe.e48 * tv * inv;
                const float2 Rj = A.cm[cj];
                float ddx = Ri.x - Rj.x, ddy = Ri.y - Rj.y;
                ddx = (ddx >= hLx) ? ddx - Lx : ((ddx < -hLx) ? ddx + Lx : ddx);
                ddy = (ddy >= hLy) ? ddy - Ly : ((ddy < -hLy) ? ddy + Ly : ddy);
                acc += 0.5 * ((double)(fc * dxa) * ddx + (double)(fc * dya) * ddy);
            }
        }
    }
    block_sum2(acc, dummy, s_red);
    if (threadIdx.x == 0) A.out[tile] = acc;
}

// Kinetic energy sum (sum m v^2 / 2) of a set of particles (ids, or the
// first n when null), per workgroup; and the velocity scaling of the same set.
__global__ __launch_bounds__(256) void k_vel_ke(const int32_t *ids, int n, const float *vx,
                                                const float *vy, const float *mass,
                                                double *part) {
    __shared__ double red[2 * 4];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, z = 0.0;
    if (k < n) {
        const int i = ids ? ids[k] : k;
        const float v2 = vx[i] * vx[i] + vy[i] * vy[i];
        e = 0.5 * (double)mass[i] * (double)v2;
    }
    block_sum2(e, z, red);
    if (threadIdx.x == 0) part[blockIdx.x] = e;
}

__global__ __launch_bounds__(256) void k_vel_scale(const int32_t *ids, int n, float *vx, float *vy,
                                                   float f) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = ids ? ids[k] : k;
    vx[i] *= f;
    vy[i] *= f;
}

// ---------------------------------------------------------------------------
// launchers (host side)
// ---------------------------------------------------------------------------
hipError_t launch_adj_forces(bool sigma, bool list, int block, const AdjArgs &A, hipStream_t s) {
    const ForceArgs &a = A.f;
    const int nblk = a.ntiles + a.nespans + a.nbspans;
    if (nblk <= 0) return hipSuccess;
    if (list) {
        const size_t lds = (size_t)(a.lds_cap + 1) * (2 * sizeof(float2) + (sigma ? sizeof(float) : 0));
        if (sigma)
            hipLaunchKernelGGL((k_adj_forces<true, true>), dim3(nblk), dim3(block), lds, s, A);
        else
            hipLaunchKernelGGL((k_adj_forces<false, true>), dim3(nblk), dim3(block), lds, s, A);
        return hipGetLastError();
    }
    const size_t lds = (size_t)a.lds_cap * (sizeof(float4) + sizeof(float2));
    if (sigma)
        hipLaunchKernelGGL((k_adj_forces<true, false>), dim3(nblk), dim3(block), lds, s, A);
    else
        hipLaunchKernelGGL((k_adj_forces<false, false>), dim3(nblk), dim3(block), lds, s, A);
    return hipGetLastError();
}


hipError_t launch_adj_kick(int nblk, const AdjKickArgs &a, hipStream_t s) {
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_kick, dim3(nblk), dim3(kKickThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adj_loss(const LossArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_loss, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}
// Kernel timer of the profile pass: an armed pair of events takes the start
// and end timestamps of the next timed dispatch (hipExtLaunchKernel), the
// interval a kernel trace reports, without the gaps that event records
// between launches add.  Per host thread; armed only around eager launches.
static thread_local hipEvent_t t_start = nullptr, t_stop = nullptr;

void arm_kernel_timer(hipEvent_t start, hipEvent_t stop) {
    t_start = start;
    t_stop = stop;
}

// a timed kernel's launch: by the armed events if any (which it disarms)
#define FM_GO(KER, GRID, BLK, SHM, STR, ...)                                              \
    do {                                                                                  \
        if (t_start) {                                                                    \
            hipExtLaunchKernelGGL(KER, GRID, BLK, (uint32_t)(SHM), STR, t_start, t_stop, 0u, \
                                  __VA_ARGS__);                                           \
            t_start = t_stop = nullptr;                                                   \
        } else {                                                                          \
            hipLaunchKernelGGL(KER, GRID, BLK, SHM, STR, __VA_ARGS__);                    \
        }                                                                                 \
    } while (0)

hipError_t launch_kick(int mode, bool gather, bool energy, int nblk, const KickArgs &a,
                       hipStream_t s) {
    // mode: 0 = bin only, 1 = kick + drift, 2 = final kick, 3 = both,
    // 4 = kinetic-energy partials only; a.dmax set = list mode (the drift
    // measures the displacement since the list build instead of binning);
    // energy: the final-kick part writes the kinetic-energy partials
    if (nblk <= 0) return hipSuccess;
    dim3 g(nblk), b(kKickThreads);
    const bool list = a.dmax != nullptr;
#define FM_KICK1(K2, D, G, E) \
    do { \
        if (list) FM_GO((k_kick<K2, D, G, true, E>), g, b, 0, s, a); \
        else FM_GO((k_kick<K2, D, G, false, E>), g, b, 0, s, a); \
    } while (0)
#define FM_KICK(K2, D, G) \
    do { \
        if (K2 && energy) FM_KICK1(K2, D, G, true); else FM_KICK1(K2, D, G, false); \
    } while (0)
    switch (mode) {
        case 0:
            // gated list-mode binning: a smaller grid-stride grid keeps the
            // skipped steps cheap
            if (a.ctl) g = dim3(std::min(nblk, kGatedBinGrid));
            FM_GO(k_bin, g, b, 0, s, a);
            break;
        case 4: FM_GO(k_ke, g, b, 0, s, a); break;
        case 1:
            if (gather) FM_KICK1(false, true, true, false); else FM_KICK1(false, true, false, false);
            break;
        case 2:
            if (gather) FM_KICK(true, false, true); else FM_KICK(true, false, false);
            break;
        default:
            if (gather) FM_KICK(true, true, true); else FM_KICK(true, true, false);
            break;
    }
#undef FM_KICK
#undef FM_KICK1
    return hipGetLastError();
}

hipError_t launch_sort(const ScanArgs &a, hipStream_t s) {
    if (a.nsort <= 0) return hipSuccess;
    FM_GO(k_sort, dim3(a.nsort), dim3(kSortCells), 0, s, a);
    return hipGetLastError();
}

size_t forces_lds_bytes(int lds_cap, bool types, bool debug) {
    return (size_t)lds_cap * sizeof(float4) + (types ? lds_cap * sizeof(int) : 0) +
           (debug ? lds_cap * sizeof(int) : 0);
}

hipError_t launch_forces(bool sigma, bool types, bool debug, bool list, bool energy, int kick,
                         int block, const ForceArgs &a, hipStream_t s) {
    // kick: 0 forces only, 1 with the closing half-kick, 2 with the step's
    // kick and drift (the closing and step kicks always form the kinetic
    // energy when energy is set; the dump pass has no kick)
    const int nblk = a.ntiles + a.nespans + a.nbspans > 0 ? forces_grid(a) : 0;
    if (nblk <= 0) return hipSuccess;
    dim3 g(nblk), b(block);
    if (debug) kick = 0;
    if (kick == 1) energy = true;
    if (list) {
        // neighbour lists (types: the pair table per pair); debug: the
        // interacting-set dump over the lists (with energies)
        const size_t lds = (size_t)(a.lds_cap + 1) *
                           (sizeof(float2) + (sigma ? sizeof(float) : 0) + (debug ? sizeof(int) : 0) +
                            (types ? sizeof(int) : 0));
#define FM_LIST(S, T, D, E, K) FM_GO((k_forces<S, T, D, true, E, K>), g, b, lds, s, a)
#define FM_LIST_K(S, T, E)                    \
    do {                                      \
        if (kick == 2) FM_LIST(S, T, false, E, 2); \
        else if (kick == 1) FM_LIST(S, T, false, true, 1); \
        else FM_LIST(S, T, false, E, 0);      \
    } while (0)
#define FM_LIST_T(S, E)                       \
    do {                                      \
        if (types) FM_LIST_K(S, true, E); else FM_LIST_K(S, false, E); \
    } while (0)
        if (debug) {
            if (sigma) { if (types) FM_LIST(true, true, true, true, 0); else FM_LIST(true, false, true, true, 0); }
            else { if (types) FM_LIST(false, true, true, true, 0); else FM_LIST(false, false, true, true, 0); }
        }
        else if (sigma) { if (energy) FM_LIST_T(true, true); else FM_LIST_T(true, false); }
        else { if (energy) FM_LIST_T(false, true); else FM_LIST_T(false, false); }
#undef FM_LIST_T
#undef FM_LIST_K
#undef FM_LIST
        return hipGetLastError();
    }
    const size_t lds = forces_lds_bytes(a.lds_cap, types, debug);
#define FM_LAUNCH(S, T, D, E, K) FM_GO((k_forces<S, T, D, false, E, K>), g, b, lds, s, a)
#define FM_LAUNCH_K(S, T, E)                    \
    do {                                        \
        if (kick == 2) FM_LAUNCH(S, T, false, E, 2); \
        else if (kick == 1) FM_LAUNCH(S, T, false, true, 1); \
        else FM_LAUNCH(S, T, false, E, 0);      \
    } while (0)
    if (debug) {  // the dump pass always accumulates energies
        if (sigma) { if (types) FM_LAUNCH(true, true, true, true, 0); else FM_LAUNCH(true, false, true, true, 0); }
        else { if (types) FM_LAUNCH(false, true, true, true, 0); else FM_LAUNCH(false, false, true, true, 0); }
    } else if (energy) {
        if (sigma) { if (types) FM_LAUNCH_K(true, true, true); else FM_LAUNCH_K(true, false, true); }
        else { if (types) FM_LAUNCH_K(false, true, true); else FM_LAUNCH_K(false, false, true); }
    } else {
        if (sigma) { if (types) FM_LAUNCH_K(true, true, false); else FM_LAUNCH_K(true, false, false); }
        else { if (types) FM_LAUNCH_K(false, true, false); else FM_LAUNCH_K(false, false, false); }
    }
#undef FM_LAUNCH_K
#undef FM_LAUNCH
    return hipGetLastError();
}

size_t list_build_lds_bytes(int lds_cap, int cell_cap, int nbr_cap, int block, bool buffered) {
    // staging; list buffer and the ranking histogram (list_rank)
    return (size_t)(lds_cap + 3 * cell_cap) * sizeof(float4) +
           (buffered ? (size_t)nbr_cap * (block + 4) +
                           (size_t)2 * (nbr_cap / 4 + 1) * (block / 64) * sizeof(int)
                     : 0);
}

// Workgroups of the list build that fit on the device at once (the gated
// build is launched with this grid: skipped steps cost one resident wave of
// workgroups, and a rebuild loses nothing since the rest would only queue).
int list_build_resident(bool sigma, int block, const ForceArgs &a, int device) {
    const size_t lds = list_build_lds_bytes(a.lds_cap, a.cell_cap, a.nbr_cap, block, a.lbuf != 0);
    int per_cu = 0, ncu = 0;
    hipError_t e = sigma
        ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_list_build<true>, block, lds)
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_list_build<false>, block, lds);
    if (e != hipSuccess) per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        ncu = 0;
    if (per_cu <= 0 || ncu <= 0) return kGatedGrid;
    return per_cu * ncu;
}

hipError_t launch_list_build(bool sigma, int block, const ForceArgs &a, hipStream_t s) {
    if (a.ntiles <= 0) return hipSuccess;
    const size_t lds = list_build_lds_bytes(a.lds_cap, a.cell_cap, a.nbr_cap, block, a.lbuf != 0);
    const int gated = a.gated_grid > 0 ? a.gated_grid : kGatedGrid;
    const dim3 g(a.ctl ? std::min(a.ntiles, gated) : a.ntiles);
    if (sigma)
        FM_GO((k_list_build<true>), g, dim3(block), lds, s, a);
    else
        FM_GO((k_list_build<false>), g, dim3(block), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_total_force(int n, const float *fx, const float *fy,
                              const float2 *cf, const int32_t *goff,
                              const int32_t *gent, float *ox, float *oy,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_total_force, dim3((n + 255) / 256), dim3(256), 0, s, n,
                       fx, fy, cf, goff, gent, ox, oy);
    return hipGetLastError();
}

__global__ void k_copy2(float *dst, const float *x, const float *y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        dst[i] = x[i];
        dst[n + i] = y[i];
    }
}

hipError_t launch_copy2(float *dst, const float *x, const float *y, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy2, dim3((n + 255) / 256), dim3(256), 0, s, dst, x, y, n);
    return hipGetLastError();
}

hipError_t launch_fill_int(int32_t *p, int32_t v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_cluster_props(const ClusterArgs &a, hipStream_t s) {
    if (a.ncl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cluster_props, dim3(a.ncl), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_cluster_virial(bool sigma, bool types, int block, const ClusterVirialArgs &A,
                                 hipStream_t s) {
    const ForceArgs &a = A.f;
    if (a.ntiles <= 0) return hipSuccess;
    const size_t lds = (size_t)a.lds_cap * (sizeof(float4) + sizeof(int) + (types ? sizeof(int) : 0));
    const dim3 g(a.ntiles), b(block);
    if (sigma) {
        if (types) hipLaunchKernelGGL((k_cluster_virial<true, true>), g, b, lds, s, A);
        else hipLaunchKernelGGL((k_cluster_virial<true, false>), g, b, lds, s, A);
    } else {
        if (types) hipLaunchKernelGGL((k_cluster_virial<false, true>), g, b, lds, s, A);
        else hipLaunchKernelGGL((k_cluster_virial<false, false>), g, b, lds, s, A);
    }
    return hipGetLastError();
}

hipError_t launch_vel_ke(const int32_t *ids, int n, const float *vx, const float *vy,
                         const float *mass, double *part, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_vel_ke, dim3((n + 255) / 256), dim3(256), 0, s, ids, n, vx, vy, mass, part);
    return hipGetLastError();
}

hipError_t launch_vel_scale(const int32_t *ids, int n, float *vx, float *vy, float f, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_vel_scale, dim3((n + 255) / 256), dim3(256), 0, s, ids, n, vx, vy, f);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Re-ordering of the internal order.  The internal order is the cell order
// of the particles when the system was committed; as particles diffuse it
// decorrelates from the cells, and the kick's mirror into sorted order, the
// sort's gathers and the pair pass's force writes turn into scattered
// accesses.  A re-ordering gives the movable particles (no element, bond or
// cluster refers to them) new internal ids in the order of the last
// counting sort, on the same set of internal slots, so no topology changes:
// the k-th movable particle in sorted order takes the k-th movable slot.
// perm and spos are rewritten for the new ids (the sorted order itself, and
// so the neighbour lists, stay valid); the per-particle arrays are gathered
// through oldof.
//   k_ro_count   movable count per chunk of sorted slots
//   k_ro_scan    exclusive prefix of the chunk counts (one workgroup)
//   k_ro_map     rank of each movable particle -> new id; oldof, perm, spos
//   k_ro_gather  tmp[k][q] = arr[k][oldof[q]]
//   k_ro_copy    arr[k][q] = tmp[k][q]
// ---------------------------------------------------------------------------
__device__ __forceinline__ int ro_flags(const ReorderArgs &a, int b, int4 &ids) {
    const int s0 = b * kReorderChunk + 4 * (int)threadIdx.x;
    int v[4] = {-1, -1, -1, -1};
    if (s0 + 3 < a.n) {
        const int4 q = *reinterpret_cast<const int4 *>(a.perm + s0);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
        for (int u = 0; u < 4; ++u)
            if (s0 + u < a.n) v[u] = a.perm[s0 + u];
    }
    ids = make_int4(v[0], v[1], v[2], v[3]);
    int m = 0;
    for (int u = 0; u < 4; ++u) m |= (v[u] >= 0 && a.mov[v[u]]) ? (1 << u) : 0;
    return m;
}

__global__ __launch_bounds__(256) void k_ro_count(ReorderArgs a) {
    __shared__ int red[4];
    int4 ids;
    const int m = ro_flags(a, blockIdx.x, ids);
    const int c = wave_sum_int(__popc(m));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) a.bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(1024) void k_ro_scan(ReorderArgs a) {
    __shared__ int w[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c0 = 0; c0 < a.nblk; c0 += 1024) {
        const int i = c0 + (int)threadIdx.x;
        const int v = i < a.nblk ? a.bsum[i] : 0;
        const int incl = wave_incl_scan(v);
        if (lane == 63) w[wv] = incl;
        __syncthreads();
        int wb = carry;
        for (int k = 0; k < wv; ++k) wb += w[k];
        if (i < a.nblk) a.bsum[i] = wb + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = wb + incl;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_ro_map(ReorderArgs a) {
    __shared__ int w[4];
    int4 ids;
    const int m = ro_flags(a, blockIdx.x, ids);
    const int c = __popc(m);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int incl = wave_incl_scan(c);
    if (lane == 63) w[wv] = incl;
    __syncthreads();
    int rank = a.bsum[blockIdx.x];
    for (int k = 0; k < wv; ++k) rank += w[k];
    rank += incl - c;
    const int s0 = blockIdx.x * kReorderChunk + 4 * (int)threadIdx.x;
    const int v[4] = {ids.x, ids.y, ids.z, ids.w};
    for (int u = 0; u < 4; ++u) {
        const int id = v[u];
        if (id < 0) continue;
        // (a sort that dropped particles, reported as a cell overflow,
        // can repeat an id: the rank stays inside the movable slots)
        const bool mv = ((m >> u) & 1) && rank < a.nmov;
        const int nid = mv ? a.movids[rank++] : id;
        a.oldof[nid] = id;
        a.perm[s0 + u] = nid;
        if (a.spos) a.spos[nid] = s0 + u;
    }
}

__global__ __launch_bounds__(256) void k_ro_gather(ReorderArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    int o = a.oldof[q];
    if ((unsigned)o >= (unsigned)a.n) o = q;  // only after a dropped particle
    for (int k = 0; k < a.narr; ++k) a.tmp[(size_t)k * a.n + q] = a.arr[k][o];
}

__global__ __launch_bounds__(256) void k_ro_copy(ReorderArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    for (int k = 0; k < a.narr; ++k) a.arr[k][q] = a.tmp[(size_t)k * a.n + q];
}

hipError_t launch_decide(const DecideArgs &d, hipStream_t s) {
    hipLaunchKernelGGL(k_flags, dim3((d.nrep + 3) / 4), dim3(256), 0, s, d);
    hipLaunchKernelGGL(k_decide, dim3(kDecideGrid), dim3(1024), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_reorder(const ReorderArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    const int g = (a.n + 255) / 256;
    hipLaunchKernelGGL(k_ro_count, dim3(a.nblk), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_ro_scan, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(k_ro_map, dim3(a.nblk), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_ro_gather, dim3(g), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_ro_copy, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fm
