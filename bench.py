"""Throughput of the femmd2D step on one MI355X.

Workload (the headline configuration): 1,048,576 LJ particles at density
0.8 (rc = 2.5, T = 1) plus 1,024 CST elastomer disks of radius 5 (about
93k nodes, 154k triangles), periodic square box, FP32 state with FP64
energy sums, dt = 0.005, synthetic seeded inputs (femmd_old_amd.configs.c2).

A step is one pass of the hot path over the whole system: half-kick + drift
+ wrap, the neighbour-list update (binning, counting sort and list build
on the steps where some particle has moved more than skin / 2 since the
last build, the reference's rule; skin 0.6, the engine default), LJ / CST
forces, and the second
half-kick with its kinetic-energy reduction (fused into the next step's
first kernel).  --skin 0 walks the cells every step instead.  Inputs are
resident in HBM when the timed region starts.

Prints one JSON line: value = particle-steps/s; roofline for the dominant
kernel (its dispatches' start and end timestamps, taken live by HIP events
armed on each launch, and its algorithmic bytes; valu_frac = the kernel's
pure vector-issue time, from the committed SQ_INSTS_VALU count, over its
launch time),
with the step's fused lower bound (bmin) and a measured 1 GiB device copy
beside the 8 TB/s peak; cpu_baseline = the reference CPU path restated in
C++ (oracle/), timed on a bounded sample of the same workload on this host
with 8 threads (the reference's OpenMP pair loop), cpu_baseline_1core as
shipped (one thread, a shorter sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
# --no-graph: every leg launches the step kernels one by one (profiling runs)
NO_GRAPH = False


def _eager(obj):
    if NO_GRAPH:
        obj.set_tuning(use_graph=0)
    return obj


def algorithmic_bytes(n, ncells, ntri, nbond, list_len=None, uniform_mass=False, n_corner=0):
    """Bytes each kernel moves per executed launch in this layout (each
    input array read once, each output written once; FP32 values, int32
    indices).  The force pass also integrates the n - n_corner particles
    without element or bond corners; the node kick integrates the n_corner
    others.  uniform_mass: the kick takes the mass from the replica's
    descriptor (no inv_m read).

    Per-step cell walk (list_len None):
    forces   r sorted (16), perm: 20/particle; w fx,fy: 8/corner particle;
             kicked particles r vx,vy, w x,y,vx,vy, slot: +28 (+4 inv_m);
             r start 4/cell, counter read-modify-write 8/cell;
             CST: r 3 ids, Dm, Bm, A0*lambda, A0*mu, 3 node x,y; w 3 corners:
             100/element; bond: r ids, k, r0, 2 node x,y; w 2 ends: 48/bond
    kick     (node kick) r fx,fy,x,y,vx,vy,goff; w x,y,vx,vy,slot: 48 (+4
             inv_m)/corner particle; gather 12/corner
    sort     r count, w count (zeroed), w start: 12/cell; r own slot entry,
             x,y,group; w perm, sorted (16): 36/particle
    Neighbour lists (list_len = mean list length L, padded to groups of 4:
    about L + 1.5 slots of 2 bytes):
    forces   r sorted x,y (8), length, slots, internal id (4): 16 + 2(L+1.5)
             /particle; w fx,fy: 8/corner particle; kicked particles r vx,vy,
             w x,y,vx,vy and the sorted-order copy of x,y: +32 (+4 inv_m);
             r start: 4/cell; elements and bonds as above
    kick     (node kick) r fx,fy,x,y,vx,vy,spos,goff; w x,y,vx,vy and the
             sorted-order x,y: 56 (+4 inv_m)/corner particle; gather 12/corner
    bin      r x,y; w slot: 12/particle; counters 8/cell     (rebuild steps)
    sort     as above plus spos and the sorted x,y: 48/particle (rebuild steps)
    list     r sorted (16); w slots, length: 20 + 2(L+1.5)/particle;
             r start: 4/cell                                   (rebuild steps)
    """
    corners = 36 * ntri + 24 * nbond
    elems = 100 * ntri + 48 * nbond
    im = 0 if uniform_mass else 4
    nk = n - n_corner  # integrated by the force pass
    if list_len is None:
        b = {
            "kick": (48 + im) * n_corner + corners,
            "bin": 0,
            "sort": 36 * n + 12 * ncells,
            "list_build": 0,
            "forces": 20 * n + 8 * n_corner + (28 + im) * nk + 12 * ncells + elems,
        }
    else:
        slots = 2.0 * (float(list_len) + 1.5)
        b = {
            "kick": (56 + im) * n_corner + corners,
            "bin": 12 * n + 8 * ncells,
            "sort": 48 * n + 12 * ncells,
            "list_build": int((20 + slots) * n) + 4 * ncells,
            "forces": int((16 + slots) * n) + 8 * n_corner + (32 + im) * nk + 4 * ncells + elems,
        }
    return b


REBUILD_KERNELS = ("bin", "sort", "list_build")


def compulsory_bytes(n, ncells, ntri, nbond, n_fea_nodes, n_bonded_nodes, excl=True,
                     n_corner=0, uniform_mass=True):
    """The survey's compulsory bytes (each input array read once, each output
    written once, in its own layout-independent units: FP32 values, int32
    indices, no neighbour list), per force pass and per step.
    force pass: LJ r sorted x,y, perm (+ exclusion id), w fx,fy: 20 (+4)/particle,
                r start 4/cell; CST r 3 ids, Bm, A0, E, 3 node x,y, w 3 corners:
                84/triangle; bond r ids, k, r0, 2 node x,y, w 2 ends: 40/bond
    fused force pass (the kernel launched): the same, except that for each
    of the n - n_corner particles without element or bond corners it also
                kicks and drifts: instead of w fx,fy it r vx,vy (+ inv_m when
                masses differ), w x,y,vx,vy: +16 (+4)/particle
    step:       128/particle (124 without exclusions) + 12/cell + 120/triangle
                + 20/element node + 64/bond + 20/bonded node"""
    per_p = 24 if excl else 20
    forces = per_p * n + 4 * ncells + 84 * ntri + 40 * nbond
    fused = forces + (n - n_corner) * (16 if uniform_mass else 20)
    step = ((128 if excl else 124) * n + 12 * ncells + 120 * ntri + 20 * n_fea_nodes
            + 64 * nbond + 20 * n_bonded_nodes)
    return {"forces": forces, "forces_fused": fused, "step": step}


def fused_corner_particles(w) -> int:
    """Particles with element or bond corners (integrated by the node kick,
    not by the force pass)."""
    import numpy as np
    ga = w.global_arrays()
    u = [np.asarray(ga["tri"], np.int64).ravel()]
    if w.n_bonds:
        u += [np.asarray(w.bonds[0], np.int64), np.asarray(w.bonds[1], np.int64)]
    return int(np.unique(np.concatenate(u)).size) if sum(a.size for a in u) else 0


def measured_traffic(kernel: str, mode: str):
    """HBM bytes per launch of `kernel` on the full c2 workload in `mode`
    ("list" or "cells"), from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (counters scaled by a plain copy of known size);
    None when absent."""
    p = ROOT / "profiles" / "k_forces_c2_hbm_traffic.md"
    try:
        for line in p.read_text().splitlines():
            if line.startswith(f"traffic_bytes_per_launch {kernel}[{mode}]:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_valu_us(kernel: str, mode: str):
    """Pure VALU issue time (us) of one launch of `kernel` on the full c2
    workload in `mode`, from the committed SQ_INSTS_VALU summary
    (instructions x 4 cycles / (1024 SIMDs x 2.4 GHz)); None when absent."""
    p = ROOT / "profiles" / "k_forces_c2_sq_counters.md"
    try:
        for line in p.read_text().splitlines():
            if line.startswith(f"valu_us_per_launch {kernel}[{mode}]:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def batched_leg(steps: int, loop_steps: int, tune_cache, looped_leg: bool = True):
    """C4: 512 replicas, N ~ logU[256, 32768]; one batched, graph-captured
    step for all of them against looping the plain step over the replicas
    (each replica stepping loop_steps with its own captured graph)."""
    from femmd_old_amd import Batch, configs
    t0 = time.time()
    ws = configs.c4()
    ntot = sum(w.n_total for w in ws)
    systems = [configs.build_system(w, commit=False) for w in ws]
    b = Batch(systems)
    t_tune = time.time()
    b.tune(tune_cache)
    t_tune = time.time() - t_tune
    _eager(b)
    b.run(20)
    tb0 = time.perf_counter()
    b.run(steps)
    tb1 = time.perf_counter()
    batched = ntot * steps / (tb1 - tb0)
    looped, tl0, tl1 = None, 0.0, 0.0
    if looped_leg:
        for s in systems:  # warm-up also captures each replica's step graph
            s.allocate_cuda()
            _eager(s)
            s.run(loop_steps)
        tl0 = time.perf_counter()
        for s in systems:
            s.run(loop_steps)
        tl1 = time.perf_counter()
        looped = ntot * loop_steps / (tl1 - tl0)
    errs = b.device_errors()
    print(f"[bench] c4: {len(ws)} replicas, {ntot} particles, setup {t_tune and time.time() - t0:.1f}s, "
          f"tune {t_tune:.1f}s", file=sys.stderr, flush=True)
    return {"replicas": len(ws), "particles": ntot, "steps": steps,
            "batched_particle_steps_per_s": batched,
            "batched_ms_per_step": 1e3 * (tb1 - tb0) / steps,
            "looped_particle_steps_per_s": looped,
            "looped_ms_per_step": 1e3 * (tl1 - tl0) / loop_steps if looped else None,
            "speedup": batched / looped if looped else None, "device_errors": errs}


def c3_leg(steps: int, warmup: int, skin: float):
    """C3: 4,096 disks (r log-uniform in [4, 8], about 400k FEA nodes and
    700k CST elements, per-element E and nu) in 600k LJ solvent at density
    0.75, each disk's rim bonded to 8 anchor solvent particles (k = 100,
    r0 = 1), dt = 0.002: the element- and bond-gather-heavy workload.
    Timed like the headline (warm graphs, observables every step), with the
    per-kernel device times, the step's compulsory bytes and a run-twice
    bitwise check of the gathers (two fresh systems, same steps)."""
    import numpy as np
    from femmd_old_amd import configs
    t_gen = time.time()
    w = configs.c3()
    s = _eager(configs.build_system(w, skin=skin))
    print(f"[bench] workload c3: N={w.n_total} nodes={w.n_nodes} tri={w.n_tri} bonds={w.n_bonds} "
          f"(setup {time.time() - t_gen:.1f}s)", file=sys.stderr, flush=True)
    ncx, ncy, _, _, _ = s.cells()
    s.run(max(warmup, 1))
    ls0 = s.list_stats()
    t0 = time.perf_counter()
    s.run(steps)
    t1 = time.perf_counter()
    ls1 = s.list_stats()
    secs = t1 - t0
    prof = s.profile(20)
    ga = w.global_arrays()
    bonded = len(set(np.asarray(w.bonds[0], int).tolist()) |
                 set(np.asarray(w.bonds[1], int).tolist())) if w.n_bonds else 0
    cb = compulsory_bytes(w.n_total, ncx * ncy, w.n_tri, w.n_bonds, int(w.n_nodes), bonded,
                          excl=True, n_corner=fused_corner_particles(w))
    ms_step = 1e3 * secs / steps
    # run twice: two fresh systems, the same steps, bitwise the same state
    a = _eager(configs.build_system(w, skin=skin))
    b = _eager(configs.build_system(w, skin=skin))
    a.run(50)
    b.run(50)
    same = all(np.array_equal(u, v) for u, v in zip(a.get_state(), b.get_state()))
    ea, eb = a.energies(), b.energies()
    same = same and all(ea[k] == eb[k] for k in ("u_lj", "u_fea", "u_bond", "w_fea", "w_bond"))
    errs = s.device_errors()
    list_mode = ls1["rebuilds"] > 0
    print(f"[bench] c3: {steps} steps in {secs:.3f}s ({ms_step:.4f} ms/step); run twice bitwise "
          f"{same}", file=sys.stderr, flush=True)
    return {"workload": "c3: 600,000 LJ (rho 0.75) + 4,096 CST disks r in [4, 8], rim bonded to "
                        "8 anchors each (k 100, r0 1), dt 0.002",
            "particles": w.n_total, "triangles": w.n_tri, "bonds": w.n_bonds,
            "steps": steps, "ms_per_step": ms_step,
            "value": w.n_total * steps / secs, "unit": "particle-steps/s",
            "pair_path": "list" if list_mode else "cells",
            "steps_per_rebuild": round(steps / max(ls1["rebuilds"] - ls0["rebuilds"], 1), 2)
            if list_mode else None,
            "step_kernels_ms": {k: round(v, 5) for k, v in prof.items()},
            "forces_compulsory_bytes": cb["forces_fused"],
            "forces_frac": round(cb["forces_fused"] / (prof["forces"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                 4),
            "step_compulsory_bytes": cb["step"],
            "step_frac": round(cb["step"] / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "run_twice_bitwise": same, "device_errors": errs}


def adjoint_leg(steps: int, ckpt: int):
    """C5: elastic sheet (128 x 256 nodes, E a smooth random field in
    [20, 120], bottom row pinned) in 200,000 WCA solvent.  The loss is the
    MSE between the final free-node positions and a target rollout from the
    same initial state with E perturbed by 10 % (a second smooth field).
    Discrete adjoint of `steps` steps (forward, backward sweep) with a
    checkpoint every `ckpt` steps and recomputation, and with the whole
    trajectory stored (it fits in HBM), against the same steps forward."""
    import numpy as np
    from femmd_old_amd import configs
    w = configs.c5()
    st = _eager(configs.build_system(configs.with_params(w, E=configs.c5_target_E(w))))
    st.run(steps)
    ga = w.global_arrays()
    nodes = np.arange(w.n_solvent, w.n_total)
    ids = nodes[np.isfinite(ga["mass"][nodes])].astype(np.int32)
    xt, yt, _, _ = st.get_state()
    tgt = np.stack([xt[ids], yt[ids]], 1)
    st.close()
    s = _eager(configs.build_system(w))
    s0 = s.get_state()
    s.run(steps)  # warm-up: graph capture
    s.set_state(*s0)
    s.adjoint(steps, ids, tgt, ckpt=ckpt)  # warm-up: adjoint buffers and graphs
    s.set_state(*s0)
    s.adjoint(steps, ids, tgt)
    s.set_state(*s0)
    t0 = time.perf_counter()
    s.run(steps)
    t1 = time.perf_counter()
    s.set_state(*s0)
    t2 = time.perf_counter()
    g = s.adjoint(steps, ids, tgt, ckpt=ckpt)
    t3 = time.perf_counter()
    s.set_state(*s0)
    t4 = time.perf_counter()
    gs = s.adjoint(steps, ids, tgt)  # whole trajectory stored
    t5 = time.perf_counter()
    fwd, adj_ck, adj = t1 - t0, t3 - t2, t5 - t4
    errs = s.device_errors()
    print(f"[bench] c5 adjoint: N={w.n_total} tri={w.n_tri} {steps} steps: forward {fwd:.3f}s, "
          f"checkpoint every {ckpt} {adj_ck:.3f}s ({adj_ck / fwd:.2f}x), whole trajectory "
          f"stored {adj:.3f}s ({adj / fwd:.2f}x), loss {g['loss']:.6e}", file=sys.stderr, flush=True)
    return {"workload": "c5: 128x256-node CST sheet (64,770 tri) in 200,000 WCA solvent; "
                        "MSE to a target rollout with E x (1 + 0.1 field2)",
            "particles": w.n_total, "steps": steps, "checkpoint_every": ckpt,
            "adjoint_particle_steps_per_s": w.n_total * steps / adj_ck,
            "adjoint_ms_per_step": 1e3 * adj_ck / steps,
            "forward_ms_per_step": 1e3 * fwd / steps,
            "adjoint_over_forward": adj_ck / fwd,
            "trajectory_stored": {"adjoint_ms_per_step": 1e3 * adj / steps,
                                  "adjoint_over_forward": adj / fwd},
            "loss": g["loss"], "grad_norm_E": float(np.linalg.norm(g["dE"])),
            "deps": g["deps"], "dsigma": g["dsigma"],
            "stored_vs_checkpointed_max_rel_dE": float(np.abs(gs["dE"] - g["dE"]).max() /
                                                       max(np.abs(g["dE"]).max(), 1e-300)),
            "device_errors": errs}


def copy_probe(nbytes: int):
    """Reads nbytes and writes nbytes with one elementwise kernel (5 launches):
    what FETCH_SIZE / WRITE_SIZE report for it scales the counters of the
    engine's kernels to HBM bytes."""
    import torch
    n = nbytes // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(f"[bench] copy probe: {5} launches of {n * 4} bytes read + written", file=sys.stderr)


def hbm_copy_probe(nbytes: int = 1 << 30, reps: int = 5):
    """Measured device copy bandwidth (read + write of nbytes, larger than the
    256 MB Infinity Cache, timed with HIP events) and the device's name and
    compute-unit count, reported beside the 8 TB/s peak.  Runs in its own
    process (bench.py --copy-bw): torch's HIP runtime and the engine's are
    not loaded into one process."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    p = torch.cuda.get_device_properties(0)
    del src, dst
    torch.cuda.empty_cache()
    return round(gbs, 1), {"name": p.name, "compute_units": p.multi_processor_count,
                           "memory_GB": round(p.total_memory / 1e9, 1)}


def copy_bw_child():
    import subprocess
    try:
        r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--copy-bw"],
                           capture_output=True, text=True, timeout=300)
        d = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
        return d["gbps"], d["device"]
    except Exception as e:  # reported, not fatal: the step timing stands alone
        return None, {"error": str(e)[:200]}


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpus": os.cpu_count(), "model": model}


def cpu_baseline(w, steps, threads):
    import numpy as np
    import oracle as O
    from femmd_old_amd import configs
    osys = O.from_workload(w)
    ga = w.global_arrays()
    vx, vy = configs.node_velocities(w)
    b = O.Baseline(osys, vx, vy, ga["mass"], w.dt, threads=threads)
    b.run(2)  # warm-up
    secs = b.run(steps)
    return {"value": w.n_total * steps / secs, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} steps of the full workload ({w.n_total} particles), reference "
                      f"DEV=CPU path restated in FP32 (Verlet list, skin 0.4, g++ -O2), "
                      f"{threads} thread(s), {secs:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--long-steps", type=int, default=2000,
                    help="steps of the steady-state run reported beside the timed steps")
    ap.add_argument("--cpu-steps", type=int, default=160)
    ap.add_argument("--cpu-threads", type=int, default=8,
                    help="threads of the CPU baseline (the reference's pair loop runs "
                         "num_threads(8), femmd2D/src/interactions.cpp:354)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--no-looped", action="store_true",
                    help="skip the looped single-replica comparison of the batched leg")
    ap.add_argument("--tune-cache", default=None)
    ap.add_argument("--skin", type=float, default=0.6,
                    help="neighbour-list skin (0: walk the cells every step)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--only", choices=["c3", "adjoint", "batched"], default=None,
                    help="run one secondary leg alone and print its JSON (profiling runs)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel of every leg one by one (no graph replay; "
                         "profiling runs)")
    ap.add_argument("--c3-steps", type=int, default=500)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (5 launches), the "
                         "calibration of the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--copy-bw", action="store_true",
                    help="only measure the device copy bandwidth (1 GiB, JSON line)")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-ckpt", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the workload (tests)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design "
                                   "note only (bench.py --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.copy_bw:
        gbs, dev = hbm_copy_probe()
        print(json.dumps({"gbps": gbs, "device": dev}), flush=True)
        return

    global NO_GRAPH
    NO_GRAPH = args.no_graph
    if args.only:
        # one secondary leg alone (its kernels profiled apart from the headline's)
        leg = {"c3": lambda: c3_leg(args.c3_steps, max(args.warmup, 1), args.skin),
               "adjoint": lambda: adjoint_leg(args.adjoint_steps, args.adjoint_ckpt),
               "batched": lambda: batched_leg(args.batched_steps, max(4, args.batched_steps // 10),
                                              args.tune_cache, not args.no_looped)}[args.only]()
        print(json.dumps({args.only: leg}), flush=True)
        return

    from femmd_old_amd import configs

    t_gen = time.time()
    w = configs.c2(scale=args.scale)
    s = _eager(configs.build_system(w, skin=args.skin))
    print(f"[bench] workload c2: N={w.n_total} nodes={w.n_nodes} tri={w.n_tri} L={w.Lx:.2f} "
          f"(setup {time.time() - t_gen:.1f}s)", file=sys.stderr, flush=True)
    ncx, ncy, _, _, _ = s.cells()

    # the warm-up's first run captures every kind of step body into a graph,
    # so the timed run replays graphs only, whatever the warm-up length
    s.run(max(args.warmup, 1))
    e0 = s.energies()
    ls0 = s.list_stats()
    s.reset_averages()
    t0 = time.perf_counter()
    s.run(args.steps)  # returns after a stream synchronise
    t1 = time.perf_counter()
    e1 = s.energies()
    ls1 = s.list_stats()
    avg = s.averages()
    secs = t1 - t0
    value = w.n_total * args.steps / secs
    # the same loop over a long run (steady state), and with the observables
    # formed only at the end of the run (the energy and virial sums skipped on
    # the other steps: not the reference's per-step work)
    t2 = time.perf_counter()
    s.run(args.long_steps)
    t3 = time.perf_counter()
    e2 = s.energies()
    s.set_observe_every(0)
    s.run(max(args.warmup, 1))
    t4 = time.perf_counter()
    s.run(args.steps)
    t5 = time.perf_counter()
    s.run(args.long_steps)
    t6 = time.perf_counter()
    s.set_observe_every(1)
    list_mode = ls1["rebuilds"] > 0
    rebuild_frac = (ls1["rebuilds"] - ls0["rebuilds"]) / args.steps if list_mode else 1.0
    print(f"[bench] timed {args.steps} steps in {secs:.3f}s; lists {ls1}", file=sys.stderr,
          flush=True)

    prof = s.profile(args.profile_steps)
    import numpy as np
    m = w.global_arrays()["mass"]
    ab = algorithmic_bytes(w.n_total, ncx * ncy, w.n_tri, w.n_bonds,
                           ls1["mean"] if list_mode else None,
                           uniform_mass=bool(np.isfinite(m[0]) and m[0] > 0 and np.all(m == m[0])),
                           n_corner=fused_corner_particles(w))
    kern = "forces"
    ga = w.global_arrays()
    n_nodes = int(w.n_nodes)
    bonded = len(set(np.asarray(w.bonds[0], int).tolist()) |
                 set(np.asarray(w.bonds[1], int).tolist())) if w.n_bonds else 0
    uniform = bool(np.isfinite(m[0]) and m[0] > 0 and np.all(m == m[0]))
    cb = compulsory_bytes(w.n_total, ncx * ncy, w.n_tri, w.n_bonds, n_nodes, bonded,
                          excl=bool(w.elastomers), n_corner=fused_corner_particles(w),
                          uniform_mass=uniform)
    achieved = cb["forces_fused"] / (prof[kern] * 1e-3) / 1e9
    model_achieved = ab[kern] / (prof[kern] * 1e-3) / 1e9
    step_bytes = ab["kick"] + ab["forces"] + rebuild_frac * sum(ab[k] for k in REBUILD_KERNELS)
    mode = "list" if list_mode else "cells"
    ms_step = 1e3 * secs / args.steps
    roof = {"bound": "hbm", "kernel": f"k_{kern}", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(f"k_{kern}", mode) if args.scale == 1.0 else None,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE (profiles/)",
            "bytes_per_launch": cb["forces_fused"],
            "bytes_rule": "compulsory bytes of the fused force pass (each input read once, "
                          "each output written once, no neighbour list): 24/particle + "
                          "4/cell + 84/triangle + 40/bond, + 16 per particle it kicks and "
                          "drifts (r vx,vy, w x,y,vx,vy in place of w fx,fy)",
            "launch_ms": round(prof[kern], 5),
            "layout_bytes_per_launch": ab[kern],
            "layout_achieved": round(model_achieved, 1),
            "layout_frac": round(model_achieved / HBM_PEAK_GBS, 4),
            "step_kernels_ms": {k: round(v, 5) for k, v in prof.items()},
            "step_compulsory_bytes": cb["step"],
            "step_frac": round(cb["step"] / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "step_layout_bytes": int(step_bytes),
            "step_layout_frac": round(step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the fused lower bound of a step: each particle's state, element and
    # bond record read and written once (36 B/particle, 36 B/element,
    # 16 B/bond); its fraction of the roofline bounds any fusion of the passes
    bmin = 36 * w.n_total + 36 * w.n_tri + 16 * w.n_bonds
    valu = measured_valu_us(f"k_{kern}", mode) if args.scale == 1.0 else None
    if valu is not None:
        # the kernel's other bound: vector issue (its HBM fraction is low
        # because it computes ~17 instructions per 2-byte list entry)
        roof["valu_us"] = valu
        roof["valu_frac"] = round(valu / (prof[kern] * 1e3), 4)
    roof["bmin_bytes"] = bmin
    roof["bmin_frac"] = round(bmin / (secs / args.steps) / 1e9 / HBM_PEAK_GBS, 4)
    roof["measured_copy_GBps"], roof["device"] = copy_bw_child()

    E0 = e0["ke"] + e0["u_lj"] + e0["u_fea"] + e0["u_bond"]
    E1 = e1["ke"] + e1["u_lj"] + e1["u_fea"] + e1["u_bond"]
    out = {
        "metric": "particle-steps/s @1M 2D LJ+FEA; batched vs serial replicas; adjoint steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator femmd_old_amd.configs.c2)",
        "config": {"workload": "c2: 1,048,576 LJ + 1,024 CST disks r=5, rho 0.8, rc 2.5, "
                               "dt 0.005, periodic", "n_particles": w.n_total,
                   "n_triangles": w.n_tri, "cells": [ncx, ncy], "box": w.Lx,
                   "reductions": "f64", "pair_path": mode,
                   "skin": args.skin if list_mode else 0.0,
                   "list_mean_len": round(ls1["mean"], 2),
                   "steps_per_rebuild": round(1.0 / rebuild_frac, 2) if rebuild_frac else None,
                   "observables": "every step: KE, max v^2, U and W of LJ, CST and bonds, "
                                  "pressure, formed on the device and accumulated into running "
                                  "averages (Integrator::doStep -> Properties::compute)"},
        "energy_drift": (E1 - E0) / abs(E0),
        "energy_drift_long": {"steps": args.steps + args.long_steps,
                              "relative": (e2["ke"] + e2["u_lj"] + e2["u_fea"] + e2["u_bond"] - E0)
                              / abs(E0)},
        "averages": {"steps": avg["count"], "temperature": avg["uKinetic"]["avg"],
                     "pressure": avg["pressure"]["avg"], "uPotential": avg["uPotential"]["avg"]},
        "steady_state": {"steps": args.long_steps,
                         "ms_per_step": 1e3 * (t3 - t2) / args.long_steps},
        "observables_at_run_end_only": {
            "ms_per_step": 1e3 * (t5 - t4) / args.steps,
            "value": w.n_total * args.steps / (t5 - t4),
            "steady_state_ms_per_step": 1e3 * (t6 - t5) / args.long_steps},
        "roofline": roof,
    }
    if not args.no_batched:
        del s
        out["batched"] = batched_leg(args.batched_steps, max(4, args.batched_steps // 10),
                                     args.tune_cache, not args.no_looped)
    if not args.no_c3:
        out["c3"] = c3_leg(args.c3_steps, max(args.warmup, 1), args.skin)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(args.adjoint_steps, args.adjoint_ckpt)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_steps, args.cpu_threads)
        out["cpu_baseline"]["host"] = host_info()
        if args.cpu_threads != 1:
            # the reference as shipped (no OpenMP): one core, a shorter sample
            out["cpu_baseline_1core"] = cpu_baseline(w, max(4, args.cpu_steps // 8), 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
