#!/usr/bin/env python3
"""Benchmark of the gravis force-and-integration hot path on one MI355X.

Headline workload (the reference-style KDK run the metric is quoted on):
Plummer sphere N = 262,144 (seed 42, sim-core's ChaCha20 ICs), eps = 0.01,
dt = 1e-4, leapfrog KDK with mixed-precision brute-force gravity. One "step"
is one full KDK step (kick+drift, Morton sort, force, kick) with the state
resident in HBM. value = N(N-1) directed interactions per step / step time.

Also reported: the plain FP64 GPU path on the same state, the mixed-vs-FP64
RMS/max relative force error, the live roofline of the dominant kernel
(HIP events on the stream it runs on) and a CPU baseline (the oracle's
restatement of sim-core's BruteForce, 1 thread, on a bounded sample).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G-interactions/s at N=256K (mixed vs FP64, err<=1e-6); BH steps/s at N=4M"
FLOP_PER_INTERACTION = 20.0  # one softened pair force (GPU Gems 3 ch. 31 convention)
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector/matrix dense peak (datasheet)
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_BPS = 8.0e12         # MI355X HBM3E (datasheet)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--fp64-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--bh-n-each", type=int, default=2097152)
    ap.add_argument("--bh-steps", type=int, default=20)
    ap.add_argument("--no-bh", action="store_true")
    ap.add_argument("--bh-group", type=int, default=0,
                    help="Barnes-Hut target group size (1..64) instead of autotuning "
                         "(profiling runs)")
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--ens-adj-steps", type=int, default=64)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--sph-n", type=int, default=1048576)
    ap.add_argument("--sph-steps", type=int, default=20)
    ap.add_argument("--sph-warmup", type=int, default=3)
    ap.add_argument("--sph-cpu-sample", type=int, default=65536)
    ap.add_argument("--no-sph", action="store_true")
    ap.add_argument("--c4-n", type=int, default=1048576)
    ap.add_argument("--c4-steps", type=int, default=200)
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--c1-steps", type=int, default=10000)
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--bf-shape", default=None,
                    help="ipt,tile,waves_per_cu (ipt 1 or 2, or m / b for the matrix-core "
                         "kernel with FP32 / split-bf16 operands): pin the mixed brute-force "
                         "launch shape instead of autotuning (profiling runs)")
    ap.add_argument("--probe-shapes", default=None,
                    help="only run three mixed brute-force calls per shape, shapes as "
                         "'ipt,tile,waves_per_cu' joined by ';' (per-shape PMC traffic runs)")
    ap.add_argument("--probe-workload", choices=("bh", "bh_fp64", "sph", "c4", "ens"), default=None,
                    help="only run one secondary workload and print its JSON (per-kernel PMC "
                         "traffic runs; combine with its --*-steps option and "
                         "--no-cpu-baseline)")
    ap.add_argument("--probe-copy", action="store_true",
                    help="only run a 1 GiB device copy (PMC byte-counter calibration)")
    return ap.parse_args()


def traffic_from_profiles(name, key):
    """HBM bytes per launch of launch shape `key` from the committed rocprofv3
    --pmc summary; None when that shape was not profiled (the partial-sum
    writes depend on the shape's j split)."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return float(d["shapes"][key]["hbm_bytes_per_launch"])
    except Exception:
        return None


def walk_pmc_traffic(precision, group):
    """HBM bytes per Barnes-Hut walk on the two-sphere workload from the
    committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE scaled by a
    copy of known size); None when that walk was not profiled."""
    path = os.path.join(ROOT, "profiles", "bh_c3_n4194304_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))["walks"][f"{precision},{group}"]
        return {"fetch_bytes": d["fetch_bytes"], "write_bytes": d["write_bytes"],
                "bytes_per_walk": d["fetch_bytes"] + d["write_bytes"],
                "source": "profiles/bh_c3_n4194304_traffic.json"}
    except Exception:
        return None


def phase(bytes_, seconds, note):
    """Compulsory bytes of one phase against its measured time."""
    if not seconds:
        return None
    return {"ms": round(seconds * 1e3, 4), "compulsory_bytes": int(bytes_),
            "GBps": round(bytes_ / seconds / 1e9, 1),
            "frac_hbm_peak": round(bytes_ / seconds / PEAK_HBM_BPS, 4), "bytes": note}


def tree_phase_rooflines(n, nodes, t):
    """The octree build's phases (HIP events around each, launch gaps
    included) against their compulsory bytes: Morton keys and a 7-pass
    9-bit radix sort of the 63-bit keys as 12-byte (key, index) pairs; the
    gather into Morton order, one read of the records for the sums and the
    node arrays."""
    passes = 7
    return {
        "morton_sort": phase(n * (24 + 12) + passes * n * 24, t.get("morton_sort"),
                             f"N*(24 pos in + 12 key/index out) + {passes} passes * N*24 (12 B "
                             "read + 12 B written per key)"),
        "tree_build": phase(n * 68 + n * 32 + nodes * 112, t.get("tree_build"),
                            "N*68 gather (32 B in, 4 B index, 32 B record out) + N*32 record "
                            "read for the sums + nodes*112 (16 B meta, 32 B COM/mass, 64 B "
                            "bounds)"),
    }


def sph_phase_rooflines(n, stats, nodes, kms):
    """One SPH step's phases (milliseconds from the library's event timers)
    against their compulsory bytes: the query reads each target's position
    once and writes its list; a density pass reads position, mass and h and
    writes rho, h, omega, P, c once per particle; the force pass reads the
    112-byte record and writes 56 bytes per particle."""
    sec = {k: (v / 1e3 if v else None) for k, v in kms.items()}
    out = tree_phase_rooflines(n, nodes, sec)
    out["sph_query"] = phase(n * 40 + stats["pool_entries"] * 4, sec.get("sph_query"),
                             "N*32 positions + N*8 list heads + pool entries*4 written")
    out["sph_density"] = phase(n * 80, sec.get("sph_density"),
                               "N*(40 in: position, mass, h; 40 out: rho, h, omega, P, c)")
    out["sph_force"] = phase(n * 168, sec.get("sph_force"),
                             "N*(112 in: the force record; 56 out: a, du/dt, div v, max mu, "
                             "dalpha/dt)")
    return out


def host_info():
    """The host the CPU baselines ran on: logical CPUs the process sees
    (nproc) and the CPU model from /proc/cpuinfo."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"nproc": nproc, "cpu_model": model}


def with_host(cpu):
    """A cpu_baseline record with the host's nproc and CPU model added."""
    if cpu is not None:
        cpu.update(host_info())
    return cpu


def cpu_baseline(n_sample, eps):
    """Oracle restatement of BruteForce::compute_accelerations (gravity.rs:
    47-80: i<j pairs, Newton's third law, 1 thread) on the first n_sample
    particles of the same Plummer ICs; one force evaluation."""
    import oracle as O
    s = O.plummer(262144, 42)
    sub = O.State(n_sample)
    for f in ("x", "y", "z", "m"):
        setattr(sub, f, getattr(s, f)[:n_sample].copy())
    t0 = time.perf_counter()
    O.bf_accel(sub, eps)
    dt = time.perf_counter() - t0
    inter = n_sample * (n_sample - 1)
    return {"value": inter / dt / 1e9, "unit": "G-interactions/s", "cores": 1, "kind": "port",
            "sample": f"one brute-force evaluation (sim-core i<j algorithm, f64, 1 thread) on the "
                      f"first {n_sample} particles of the N=262144 Plummer ICs, eps={eps}: "
                      f"{dt:.2f} s"}


def probe_copy():
    """A plain device-to-device copy of known size (1 GiB read + 1 GiB
    written per launch) for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "copy", "bytes_read": a.numel() * 8, "bytes_written": b.numel() * 8}))


def shape_key(t):
    """Mixed brute-force launch shape as 'ipt,tile,waves_per_cu' ('m' and 'b'
    for the matrix-core kernel with FP32 and split-bf16 operands), the key of
    the traffic profile."""
    kern = {0: str(t["bf_ipt"]), 1: "m", 2: "b"}[t["bf_matrix"]]
    return f"{kern},{t['bf_tile']},{t['bf_waves_per_cu']}"


def set_shape(ctx, key):
    ipt, tile, wpc = key.split(",")
    ctx.set_tuning("bf_matrix", {"m": 1, "b": 2}.get(ipt, 0))
    if ipt not in ("m", "b"):
        ctx.set_tuning("bf_ipt", int(ipt))
    ctx.set_tuning("bf_tile", int(tile))
    ctx.set_tuning("bf_waves_per_cu", int(wpc))


def probe_shapes(args):
    """Three mixed brute-force force calls per listed launch shape on the
    bench's Plummer ICs (rocprofv3 --pmc runs of the per-shape HBM traffic)."""
    import gravis_amd as g
    p = g.scenarios.plummer(args.n, 42)
    ctx = g.Context(0)
    for key in args.probe_shapes.split(";"):
        set_shape(ctx, key)
        bf = g.BruteForce(args.eps, "mixed", ctx=ctx)
        for _ in range(3):
            p.clear_accelerations()
            bf.compute_accelerations(p)
        ctx.synchronize()
        print(json.dumps({"shape": key, "tuning": ctx.tuning()}), flush=True)
    ctx.close()


def run_c1(args, g, ctx):
    """Plummer N=1000 (seed 42), eps 0.05, dt 1e-3, FP64 brute force, 10,000
    fused KDK steps (host arrays in and out); the oracle runs the same loop
    on one core for a bounded sample of steps."""
    import oracle as O
    n, eps, dt, T = 1000, 0.05, 1e-3, args.c1_steps
    log(f"bench: C1 Plummer N={n}, {T} FP64 KDK steps")
    p = g.scenarios.plummer(n, 42)
    gr = g.BruteForce(eps, ctx=ctx)
    gr.compute_accelerations(p)
    e0 = g.diagnostics.compute(p, eps, ctx=ctx).total_energy
    g.LeapfrogKDK().run(p, gr, dt, 10)  # warm-up
    t0 = time.perf_counter()
    g.LeapfrogKDK().run(p, gr, dt, T)
    t_gpu = time.perf_counter() - t0
    d = g.diagnostics.compute(p, eps, ctx=ctx)
    s = O.plummer(n, 42)
    O.bf_accel(s, eps)
    cpu_steps = min(T, 2000)
    t0 = time.perf_counter()
    O.kdk_steps(s, dt, cpu_steps, eps, nthreads=1)
    t_cpu = (time.perf_counter() - t0) / cpu_steps
    return {"n": n, "steps": T, "steps_per_s": round(T / t_gpu, 1),
            "ms_per_step": round(t_gpu / T * 1e3, 4),
            "energy_drift": (d.total_energy - e0) / abs(e0), "virial_ratio": d.virial_ratio,
            "cpu_baseline": with_host({"value": round(1.0 / t_cpu, 1), "unit": "steps/s",
                                       "cores": 1, "kind": "port",
                                       "sample": f"{cpu_steps} steps of the oracle's sim-core "
                                                 f"loop (i<j brute force, f64, 1 thread)"}),
            "workload": "Plummer N=1000, seed 42, eps 0.05, dt 1e-3, FP64 brute force, "
                        "leapfrog KDK fused on the device"}


def cpu_threads():
    """Host threads for the multi-threaded CPU baselines: at most 16, the
    host share of one GPU's process."""
    return max(1, min(16, os.cpu_count() or 1))


def bh_cpu_baseline(p, eps, theta):
    """The oracle's restatement of the reference Barnes-Hut
    (barnes_hut.rs:44-192: serial insertion octree, octree.rs:145-239, and a
    per-particle stack walk run in parallel over particles, as rayon does)
    on the same ICs: one force evaluation, the cost of one KDK step."""
    import oracle as O
    n = p.count
    s = O.State(n)
    s.x, s.y, s.z, s.m = [np.ascontiguousarray(getattr(p, f)) for f in ("x", "y", "z", "mass")]
    th = cpu_threads()
    t0 = time.perf_counter()
    O.bh_accel(s, eps, theta, th)
    dt = time.perf_counter() - t0
    return {"value": round(1.0 / dt, 4), "unit": "steps/s (one force evaluation per KDK step)",
            "cores": th, "kind": "port",
            "sample": f"one reference Barnes-Hut force evaluation (insertion octree + parallel "
                      f"walk, f64, theta={theta}) on the same {n} particles: {dt:.2f} s"}


def run_bh(args, g, torch, ctx, precision="mixed", steps=None):
    """Barnes-Hut KDK steps on two Plummer spheres of n_each particles (seeds
    42/43, a = 1, M = 1 each) 10 a apart on a parabolic approach, eps 0.01,
    dt 5e-4, theta 0.5, leaf 16, target groups of the size the autotuner
    picks for this precision (at most 64). precision 'mixed':
    FP32 pair forces on FP64 group-relative separations (FP64 tree, state and
    opening test); 'fp64': the reference arithmetic."""
    steps = steps or args.bh_steps
    import ctypes as C
    ne = args.bh_n_each
    n = 2 * ne
    dev = torch.device("cuda:0")
    log(f"bench: BH two Plummer spheres N={n}; generating ICs")
    p = g.scenarios.plummer_pair(ne, 42, 43, 10.0, 0.0)
    T = {f: torch.from_numpy(getattr(p, f)).to(dev) for f in ("x", "y", "z", "vx", "vy", "vz",
                                                              "mass")}
    for f in ("ax", "ay", "az"):
        T[f] = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    # group size 0: the context's, chosen by the autotuner (or pinned)
    bh = g.BarnesHut(0.01, 0.5, 16, 0, precision=precision, ctx=ctx)
    if args.bh_group:
        ctx.set_tuning("bh_group_size", args.bh_group)
        log(f"bench: BH group size pinned to {args.bh_group}")
    else:
        log(f"bench: autotuned BH ({precision}): {ctx.autotune(bh, p)}")
    group = ctx.tuning()["bh_group_size"]
    L = ctx._L
    keys = ("x", "y", "z", "vx", "vy", "vz", "mass", "ax", "ay", "az")
    d = bh.desc()

    def kdk(steps):
        ctx.check(L.gravis_kdk_device(ctx.handle, C.byref(d), n,
                                      *[C.c_void_p(T[k].data_ptr()) for k in keys], 5e-4, steps))

    ctx.check(L.gravis_accel_device(ctx.handle, C.byref(d), n,
                                    *[C.c_void_p(T[k].data_ptr()) for k in
                                      ("x", "y", "z", "mass", "ax", "ay", "az")]))

    def measure(nsteps):
        """Two warm-up steps, nsteps timed KDK steps, then three steps with
        the walk's HIP-event timer (a walk's time varies call to call with
        the order its queue hands out groups): step time, mean walk time,
        walk counters of the last step and the walk's algorithmic bytes."""
        kdk(2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            kdk(1)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / nsteps
        ctx.set_profiling(True)
        kdk(3)
        ctx.set_profiling(False)
        ms, launches = ctx.kernel_time("bh_walk")
        ws = ms / 1e3 / max(1, launches)
        st = ctx.bh_stats()
        # Algorithmic bytes of one walk (per group, every node it touches
        # once and every leaf particle it takes as a source once, plus the
        # targets in and out): node_tests * B_node + leaf_particles * B_part
        # + N * 48, with B_node = 20 B (FP32 COM and mass, node size) or
        # 40 B (FP64) and B_part = 16 B or 32 B. A larger group shares one
        # list among more targets, so it moves fewer bytes per step.
        b_node, b_part = (20, 16) if precision == "mixed" else (40, 32)
        wb = st["node_tests"] * b_node + st["leaf_particles"] * b_part + n * 48
        phases = {}
        for k in ("morton_sort", "tree_build"):
            ms, launches = ctx.kernel_time(k)
            phases[k] = ms / 1e3 / max(1, launches)
        return t, ws, st, wb, phases

    log(f"bench: timing {steps} BH KDK steps ({precision})")
    t_step, walk_s, st, walk_bytes, phases = measure(steps)
    # the other group sizes on the same run, for comparison (the autotuner's
    # choice stays the reported one)
    others = []
    if precision == "mixed" and not args.bh_group:
        for gs in (32, 64):
            if gs == group:
                continue
            ctx.set_tuning("bh_group_size", gs)
            log(f"bench: BH group size {gs} for comparison")
            t2, w2, _, b2, _ = measure(max(5, steps // 2))
            others.append({"group_size": gs, "ms_per_step": round(t2 * 1e3, 3),
                           "walk_kernel_ms": round(w2 * 1e3, 3),
                           "walk_algorithmic_GBps": round(b2 / w2 / 1e9, 1),
                           "walk_frac_hbm_peak": round(b2 / w2 / PEAK_HBM_BPS, 4)})
        ctx.set_tuning("bh_group_size", group)
    cpu = None
    if precision == "mixed" and not args.no_cpu_baseline:
        log(f"bench: CPU baseline, reference Barnes-Hut on {n} particles")
        cpu = with_host(bh_cpu_baseline(p, 0.01, 0.5))
    pmc = walk_pmc_traffic(precision, group)
    if pmc:
        pmc["GBps"] = round(pmc["bytes_per_walk"] / walk_s / 1e9, 1)
        pmc["frac_hbm_peak"] = round(pmc["bytes_per_walk"] / walk_s / PEAK_HBM_BPS, 4)
    return {"n": n, "precision": precision, "steps_per_s": round(1.0 / t_step, 3),
            "cpu_baseline": cpu,
            "ms_per_step": round(t_step * 1e3, 3), "steps": steps, "theta": 0.5, "leaf_size": 16, "group_size": group,
            "softening": 0.01, "dt": 5e-4,
            "walk_kernel_ms": round(walk_s * 1e3, 3),
            "walk_stats": st,
            "walk_pairs_per_s": st["pairs"] / walk_s,
            "walk_achieved_tflops": round(FLOP_PER_INTERACTION * st["pairs"] / walk_s / 1e12, 3),
            "walk_algorithmic_bytes": walk_bytes,
            "walk_algorithmic_GBps": round(walk_bytes / walk_s / 1e9, 1),
            "walk_frac_hbm_peak": round(walk_bytes / walk_s / PEAK_HBM_BPS, 4),
            "walk_bytes_definition": "node_tests*%d + leaf_particles*%d + N*48 (each node and "
                                     "leaf particle counted once per target group)"
                                     % ((20, 16) if precision == "mixed" else (40, 32)),
            "walk_pmc_traffic": pmc,
            "tree_phases": tree_phase_rooflines(n, st["nodes"], phases),
            "other_group_sizes": others,
            "workload": "two Plummer spheres (seeds 42/43, a=1, M=1 each) 10a apart, parabolic "
                        "approach, KDK with Barnes-Hut theta=0.5, leaf 16"}


def ens_cpu_baseline(cnts, steps=10, systems=4):
    """The oracle's FP64 ensemble forward (gather-form brute force, KDK) and
    checkpoint-free adjoint on the first `systems` systems of the same
    ensemble, `steps` steps each; per-step times and pair rates."""
    import oracle as O
    pairs = 0.0
    tf = ta = 0.0
    for k in range(systems):
        ns = int(cnts[k])
        st = O.plummer(ns, 42 + k)
        pairs += float(ns) * (ns - 1)
        t0 = time.perf_counter()
        O.ens_forward(st.copy(), 0.05, 1e-3, steps)
        tf += time.perf_counter() - t0
        gx, gy, gz = (np.ones(ns) for _ in range(3))
        t0 = time.perf_counter()
        O.ens_adjoint(st.copy(), 0.05, 1e-3, steps, gx, gy, gz)
        ta += time.perf_counter() - t0
    ns_list = [int(c) for c in cnts[:systems]]
    return {"value": round(pairs * steps / tf / 1e9, 4), "unit": "G-interactions/s", "cores": 1,
            "kind": "port",
            "adjoint_G_pairs_per_s": round(pairs * steps / ta / 1e9, 4),
            "sample": f"oracle FP64 forward and adjoint (f64, 1 thread) of systems N = {ns_list}, "
                      f"{steps} steps each: forward {tf:.2f} s, adjoint {ta:.2f} s"}


def run_ensemble(args, g, torch, ctx):
    """256 Plummer spheres (seed 42 + s), N_s drawn in [512, 8192] (seed 42),
    eps 0.05, dt 1e-3, KDK with mixed brute force; one force launch covers
    every system and each step replays a captured hipGraph. Then the
    checkpointed adjoint (every 32 steps) of L = sum |x(T) - target|^2
    (targets seeded 10000 + s)."""
    import ctypes as C
    import gravis_amd.ensemble as E
    cnts = E.counts(42, 256, 512, 8192)
    n = int(cnts.sum())
    inter = float((cnts.astype(np.float64) * (cnts.astype(np.float64) - 1)).sum())
    log(f"bench: ensemble of {len(cnts)} systems, N_total={n}")
    st = E.plummer_ensemble(cnts)
    tg = [np.concatenate(a) for a in zip(*[E.targets(10000 + s, int(c))
                                            for s, c in enumerate(cnts)])]
    dev = torch.device("cuda:0")
    T = {k: torch.from_numpy(st[k]).to(dev) for k in st}
    TT = [torch.from_numpy(a).to(dev) for a in tg]
    torch.cuda.synchronize()
    off = E._offsets(cnts)
    solver = E._solver(0.05, "mixed")
    L = ctx._L
    p = lambda t: C.c_void_p(t.data_ptr())
    ctx.check(L.gravis_ensemble_accel_device(ctx.handle, C.byref(solver), len(cnts), E._p(off),
                                             *[p(T[k]) for k in ("x", "y", "z", "m", "ax", "ay",
                                                                  "az")]))
    keys = ("x", "y", "z", "vx", "vy", "vz", "m", "ax", "ay", "az")

    def kdk(steps):
        ctx.check(L.gravis_ensemble_kdk_device(ctx.handle, C.byref(solver), len(cnts), E._p(off),
                                               *[p(T[k]) for k in keys], 1e-3, steps, 1))
    kdk(2)
    torch.cuda.synchronize()
    log(f"bench: timing {args.ens_steps} ensemble KDK steps (hipGraph)")
    t0 = time.perf_counter()
    kdk(args.ens_steps)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.ens_steps
    out = torch.zeros(7, n, dtype=torch.float64, device=dev)
    loss = torch.zeros(len(cnts), dtype=torch.float64, device=dev)
    log(f"bench: timing the adjoint over {args.ens_adj_steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.check(L.gravis_ensemble_adjoint_device(
        ctx.handle, C.byref(solver), len(cnts), E._p(off),
        *[p(T[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")], 1e-3, args.ens_adj_steps, 32,
        *[p(a) for a in TT], None, None, None, None, None, None, p(loss), None,
        *[p(out[r]) for r in range(7)]))
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline, oracle ensemble forward and adjoint on 4 systems")
        cpu = with_host(ens_cpu_baseline(cnts))
    return {"systems": len(cnts), "n_total": n, "interactions_per_step": inter,
            "steps": args.ens_steps, "ms_per_step": round(t_step * 1e3, 4),
            "G_interactions_per_s": round(inter / t_step / 1e9, 3),
            "cpu_baseline": cpu,
            "adjoint": {"steps": args.ens_adj_steps, "checkpoint_every": 32,
                        "seconds": round(t_adj, 4),
                        "ms_per_step": round(t_adj / args.ens_adj_steps * 1e3, 4),
                        "G_pairs_per_s": round(inter * args.ens_adj_steps / t_adj / 1e9, 3),
                        "note": "forward with checkpoints + segment recompute + FP64 pair-VJP"},
            "workload": "256 Plummer spheres, N_s in [512, 8192] (seeded), eps 0.05, dt 1e-3, "
                        "mixed brute force, hipGraph step"}


def run_sph(args, g, torch, ctx):
    """Evrard collapse (EvrardCollapse::generate, seed 42: n gas particles
    uniform in R = 1, M = 1, u0 = 0.05) with sim-core's SPH (Wendland C2,
    fixed-point h, Monaghan viscosity with the Morris-Monaghan switch, Price
    conductivity) and FP64 Barnes-Hut self-gravity (theta 0.5, leaf 16,
    eps = suggested_softening), stepped with headless's rule
    dt = min(dt_next, 1.5 dt) from dt = 1e-3. State stays in HBM."""
    import ctypes as C
    from gravis_amd import _lib
    from gravis_amd.sph import SphSolver
    n = args.sph_n
    dev = torch.device("cuda:0")
    log(f"bench: SPH Evrard N={n}; generating ICs")
    p = g.scenarios.evrard(n)
    eps = g.scenarios.evrard_softening(n)
    fields = dict(x="x", y="y", z="z", vx="vx", vy="vy", vz="vz", ax="ax", ay="ay", az="az",
                  m="mass", rho="density", p="pressure", u="internal_energy", dudt="du_dt",
                  h="smoothing_length", c="sound_speed", alpha="alpha_visc")
    T = {k: torch.from_numpy(getattr(p, v)).to(dev) for k, v in fields.items()}
    T["type"] = torch.from_numpy(p.particle_type).to(dev)
    inter = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    st = _lib.SphState()
    st.n = n
    for k in _lib.SPH_FIELDS:
        setattr(st, k, T[k].data_ptr())
    st.type = T["type"].data_ptr()
    solver = SphSolver(ctx)
    prm = solver.params()
    bh = g.BarnesHut(eps, 0.5, 16, 64, precision="fp64", ctx=ctx)
    d = bh.desc()
    L = ctx._L
    ctx.check(L.gravis_accel_device(ctx.handle, C.byref(d), n,
                                    *[C.c_void_p(T[k].data_ptr()) for k in
                                      ("x", "y", "z", "m", "ax", "ay", "az")]))
    ctx.check(L.gravis_sph_compute_device(ctx.handle, C.byref(prm), C.byref(st),
                                          *[C.c_void_p(t.data_ptr()) for t in inter],
                                          None, None))
    dt = C.c_double(1e-3)
    tsim = C.c_double(0.0)

    def run(steps):
        ctx.check(L.gravis_sph_run_device(ctx.handle, C.byref(prm), C.byref(d), C.byref(st),
                                          steps, C.byref(dt), C.byref(tsim)))

    log(f"bench: {args.sph_warmup} SPH warmup steps")
    run(args.sph_warmup)
    torch.cuda.synchronize()
    log(f"bench: timing {args.sph_steps} SPH steps")
    t_start = tsim.value
    t0 = time.perf_counter()
    run(args.sph_steps)
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.sph_steps
    sim_span = tsim.value - t_start
    ctx.set_profiling(True)
    run(1)
    ctx.set_profiling(False)
    kms = {}
    for k in ("sph_query", "sph_density", "sph_force", "bh_walk", "morton_sort", "tree_build"):
        ms, launches = ctx.kernel_time(k)
        kms[k] = round(ms, 3)
    stats = solver.stats()
    roofs = sph_phase_rooflines(n, stats, ctx.bh_stats()["nodes"], kms)
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        ns = args.sph_cpu_sample
        s = O.evrard(ns)
        O.sph_compute(s)          # converge h first (untimed)
        O.sph_timestep(s, {"div_v": np.zeros(ns), "max_mu": np.zeros(ns)})
        t0 = time.perf_counter()
        O.sph_step(s, 1e-4, gravity=0)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(ns / t_cpu, 1), "unit": "particle-steps/s (SPH only)",
               "cores": 1, "kind": "port",
               "sample": f"one step_with_sph without gravity on a {ns}-particle Evrard sphere "
                         f"(sim-core algorithm, f64, 1 thread, lists by the reference octree "
                         f"ball query): {t_cpu:.2f} s"}
    return {"n": n, "steps_per_s": round(1.0 / t_step, 3), "ms_per_step": round(t_step * 1e3, 3),
            "particle_steps_per_s": round(n / t_step, 1),
            "steps": args.sph_steps, "warmup": args.sph_warmup, "sim_time_span": sim_span,
            "dt_last": dt.value, "softening": eps, "theta": 0.5,
            "kernel_ms_one_step": kms, "last_compute_stats": stats,
            "phase_rooflines": roofs, "cpu_baseline": with_host(cpu),
            "workload": "Evrard collapse (uniform sphere R=1, M=1, u0=0.05, seed 42), sim-core "
                        "SPH (Wendland C2, fixed-point h, Monaghan AV + switch, conductivity), "
                        "FP64 Barnes-Hut gravity theta=0.5, headless dt rule"}


def run_sph_c4(args, g, torch, ctx):
    """The grad-h SPH configuration with Barnes-Hut self-gravity: n gas
    particles in R = 1, M = 1 with rho ~ 1/r (seed 42), u = 0.05, at rest,
    gamma 5/3, cubic-spline kernel, eta 1.2, Newton h to 1e-4, Monaghan
    viscosity alpha 1 / beta 2 (fixed), no conduction, gravity softening
    eps = h, FP32 pair loops with FP64 sums (density, forces, gravity walk),
    FP64 state; c4_steps KDK steps with the global Courant dt (CFL 0.3,
    headless's dt rule from dt = 1e-4). State stays in HBM."""
    import ctypes as C
    from gravis_amd import _lib
    from gravis_amd.sph import SphSolver
    n = args.c4_n
    dev = torch.device("cuda:0")
    log(f"bench: SPH (cubic spline, Newton h, eps = h) rho~1/r N={n}; generating ICs")
    p = g.scenarios.evrard_1r(n)
    fields = dict(x="x", y="y", z="z", vx="vx", vy="vy", vz="vz", ax="ax", ay="ay", az="az",
                  m="mass", rho="density", p="pressure", u="internal_energy", dudt="du_dt",
                  h="smoothing_length", c="sound_speed", alpha="alpha_visc")
    T = {k: torch.from_numpy(getattr(p, v)).to(dev) for k, v in fields.items()}
    T["type"] = torch.from_numpy(p.particle_type).to(dev)
    inter = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    torch.cuda.synchronize()
    st = _lib.SphState()
    st.n = n
    for k in _lib.SPH_FIELDS:
        setattr(st, k, T[k].data_ptr())
    st.type = T["type"].data_ptr()
    solver = SphSolver(ctx, variant=True)
    prm = solver.params()
    bh = g.BarnesHut(0.0, 0.5, 16, 64, precision="mixed", ctx=ctx)
    d = bh.desc()
    L = ctx._L
    ctx.check(L.gravis_sph_accel_device(ctx.handle, C.byref(prm), C.byref(d), C.byref(st)))
    ctx.check(L.gravis_sph_compute_device(ctx.handle, C.byref(prm), C.byref(st),
                                          *[C.c_void_p(t.data_ptr()) for t in inter],
                                          None, None))
    dt = C.c_double(1e-4)
    tsim = C.c_double(0.0)
    torch.cuda.synchronize()
    log(f"bench: timing {args.c4_steps} SPH steps (variant)")
    t0 = time.perf_counter()
    ctx.check(L.gravis_sph_run_device(ctx.handle, C.byref(prm), C.byref(d), C.byref(st),
                                      args.c4_steps, C.byref(dt), C.byref(tsim)))
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / args.c4_steps
    ctx.set_profiling(True)
    ctx.check(L.gravis_sph_run_device(ctx.handle, C.byref(prm), C.byref(d), C.byref(st), 1,
                                      C.byref(dt), C.byref(tsim)))
    ctx.set_profiling(False)
    kms = {}
    for k in ("sph_query", "sph_density", "sph_force", "bh_walk", "morton_sort", "tree_build"):
        ms, _ = ctx.kernel_time(k)
        kms[k] = round(ms, 3)
    stats = solver.stats()
    roofs = sph_phase_rooflines(n, stats, ctx.bh_stats()["nodes"], kms)
    ke = float(0.5 * (T["m"] * (T["vx"] ** 2 + T["vy"] ** 2 + T["vz"] ** 2)).sum())
    eth = float((T["m"] * T["u"]).sum())
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        ns = args.sph_cpu_sample
        s = O.evrard_1r(ns)
        vp = O.SphParams.variant()
        O.sph_compute(s, vp)  # converge h first (untimed)
        t0 = time.perf_counter()
        O.sph_step(s, 1e-4, gravity=0, params=vp)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(ns / t_cpu, 1), "unit": "particle-steps/s (SPH only)",
               "cores": 1, "kind": "port",
               "sample": f"one SPH step of the same physics without gravity on a {ns}-particle "
                         f"rho~1/r sphere (oracle, f64, 1 thread): {t_cpu:.2f} s"}
    return {"n": n, "steps": args.c4_steps, "steps_per_s": round(1.0 / t_step, 3),
            "ms_per_step": round(t_step * 1e3, 3), "particle_steps_per_s": round(n / t_step, 1),
            "sim_time": tsim.value, "dt_last": dt.value, "kinetic_energy": ke,
            "thermal_energy": eth, "kernel_ms_one_step": kms, "last_compute_stats": stats,
            "phase_rooflines": roofs, "cpu_baseline": with_host(cpu),
            "workload": "rho~1/r gas sphere (R=1, M=1, u=0.05, seed 42), cubic spline, "
                        "eta=1.2, Newton h to 1e-4, alpha=1 beta=2, eps=h Barnes-Hut "
                        "theta=0.5 (mixed walk), FP32 pairs / FP64 sums, CFL 0.3"}


def main():
    args = parse()
    if args.probe_copy:
        probe_copy()
        return
    if args.probe_shapes:
        probe_shapes(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(one process drives one GPU; sharding across GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    import gravis_amd as g

    n, eps, dt = args.n, args.eps, args.dt
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(dev)  # default priority; every op below runs on it
    torch.cuda.set_stream(stream)
    if args.probe_workload:
        # one workload alone (rocprofv3 --pmc traffic runs of its kernels)
        ctx = g.Context(0)
        ctx.set_stream(stream.cuda_stream)
        run = {"bh": lambda: run_bh(args, g, torch, ctx, "mixed"),
               "bh_fp64": lambda: run_bh(args, g, torch, ctx, "fp64"),
               "sph": lambda: run_sph(args, g, torch, ctx),
               "c4": lambda: run_sph_c4(args, g, torch, ctx),
               "ens": lambda: run_ensemble(args, g, torch, ctx)}[args.probe_workload]
        print(json.dumps({"probe": args.probe_workload, "result": run()}), flush=True)
        ctx.close()
        return
    log(f"bench: Plummer N={n} eps={eps} dt={dt}; generating ICs")
    p = g.scenarios.plummer(n, 42)
    T = {f: torch.from_numpy(getattr(p, f)).to(dev) for f in ("x", "y", "z", "vx", "vy", "vz",
                                                              "mass")}
    for f in ("ax", "ay", "az"):
        T[f] = torch.zeros(n, dtype=torch.float64, device=dev)

    torch.cuda.synchronize()
    ctx = g.Context(0)
    ctx.set_stream(stream.cuda_stream)
    mixed = g.BruteForce(eps, "mixed", ctx=ctx)
    fp64 = g.BruteForce(eps, "fp64", ctx=ctx)
    import ctypes as C
    L = ctx._L

    def ptrs(keys):
        return [C.c_void_p(T[k].data_ptr()) for k in keys]

    state_keys = ("x", "y", "z", "vx", "vy", "vz", "mass", "ax", "ay", "az")

    def kdk(solver, steps):
        d = solver.desc()
        ctx.check(L.gravis_kdk_device(ctx.handle, C.byref(d), n, *ptrs(state_keys), dt, steps))

    def accel(solver, out):
        d = solver.desc()
        ctx.check(L.gravis_accel_device(ctx.handle, C.byref(d), n,
                                        *[C.c_void_p(T[k].data_ptr()) for k in
                                          ("x", "y", "z", "mass")],
                                        *[C.c_void_p(o.data_ptr()) for o in out]))

    # Launch shapes of the mixed kernel picked by the library's autotuner on
    # these particles (outside the timed region).
    if args.bf_shape:
        set_shape(ctx, args.bf_shape)
        tuning = dict(ctx.tuning(), pinned=True)
    else:
        tuning = ctx.autotune(mixed, p)
    log(f"bench: mixed brute-force launch shape: {tuning}")
    # the timer and kernel of the chosen shape
    kname = "bf_matrix" if tuning["bf_matrix"] else "bf_mixed"
    # a(t0), as headless does before the loop (main.rs:151-152)
    accel(mixed, (T["ax"], T["ay"], T["az"]))
    log(f"bench: warmup {args.warmup} mixed KDK steps")
    for _ in range(args.warmup):
        kdk(mixed, 1)
    torch.cuda.synchronize()

    log(f"bench: timing {args.steps} mixed KDK steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kdk(mixed, 1)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    inter = float(n) * float(n - 1)
    value = inter / t_mixed / 1e9

    # Live kernel timing of the dominant kernel (events on the same stream).
    prof_steps = max(3, min(10, args.steps))
    ctx.set_profiling(True)
    for _ in range(prof_steps):
        kdk(mixed, 1)
    ctx.set_profiling(False)
    torch.cuda.synchronize()
    ms_total, launches = ctx.kernel_time(kname)
    k_mixed_s = ms_total / 1e3 / max(1, launches)
    achieved = FLOP_PER_INTERACTION * inter / k_mixed_s / 1e12

    # Accuracy on the current state: mixed vs FP64 GPU path.
    am = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    a6 = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    accel(mixed, am)
    accel(fp64, a6)
    torch.cuda.synchronize()
    A6 = torch.stack(a6, 1)
    AM = torch.stack(am, 1)
    m2 = (A6 * A6).sum(1)
    e2 = ((AM - A6) ** 2).sum(1)
    ok = m2 >= 1e-30
    rel2 = e2[ok] / m2[ok]
    err_rms = float(torch.sqrt(rel2.sum() / n))
    err_max = float(torch.sqrt(rel2.max()))

    # Plain FP64 GPU path on the same state (its own KDK steps).
    log(f"bench: timing {args.fp64_steps} FP64 KDK steps")
    kdk(fp64, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        kdk(fp64, 1)
    torch.cuda.synchronize()
    t_fp64 = (time.perf_counter() - t0) / args.fp64_steps
    ctx.set_profiling(True)
    kdk(fp64, 1)
    ctx.set_profiling(False)
    ms64, l64 = ctx.kernel_time("bf_fp64")
    k_fp64_s = ms64 / 1e3 / max(1, l64)

    # The vector-ALU kernel (no matrix cores) at its usual fastest shapes on
    # the same state, beside the matrix-core kernel the autotuner picked.
    valu = None
    if tuning["bf_matrix"]:
        log("bench: VALU-only mixed kernel shapes")
        best = None
        for key in ("2,128,256", "2,128,512", "1,128,256"):
            set_shape(ctx, key)
            accel(mixed, am)
            torch.cuda.synchronize()
            ctx.set_profiling(True)
            for _ in range(3):
                accel(mixed, am)
            ctx.set_profiling(False)
            ms, launches = ctx.kernel_time("bf_mixed")
            k_s = ms / 1e3 / max(1, launches)
            if best is None or k_s < best[1]:
                best = (key, k_s)
        set_shape(ctx, shape_key(tuning))
        valu = {"kernel": "bf_mixed_kernel", "shape": best[0],
                "kernel_ms": round(best[1] * 1e3, 4),
                "G_interactions_per_s": round(inter / best[1] / 1e9, 3),
                "frac_fp32_peak": round(FLOP_PER_INTERACTION * inter / best[1] / 1e12
                                        / PEAK_FP32_TFLOPS, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline on {args.cpu_sample} particles")
        cpu = with_host(cpu_baseline(args.cpu_sample, eps))

    # C1: the reference's own CPU case (Plummer N=1000, eps 0.05, dt 1e-3,
    # 10,000 FP64 KDK steps) on the GPU, the oracle beside it
    c1 = None
    if not args.no_c1:
        c1 = run_c1(args, g, ctx)

    bh = bh64 = None
    if not args.no_bh:
        bh = run_bh(args, g, torch, ctx, "mixed")
        bh64 = run_bh(args, g, torch, ctx, "fp64", max(2, args.bh_steps // 4))

    ens = None
    if not args.no_ensemble:
        ens = run_ensemble(args, g, torch, ctx)

    sph = None
    if not args.no_sph:
        sph = run_sph(args, g, torch, ctx)
    c4 = None
    if not args.no_c4:
        c4 = run_sph_c4(args, g, torch, ctx)

    traffic = traffic_from_profiles(f"bf_mixed_n{n}", shape_key(tuning))
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "G-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair forces / f64 state",
        "data": "synthetic: Plummer sphere, sim-core ChaCha20 ICs, seed 42",
        "config": {"workload": f"Plummer N={n}, eps={eps}, dt={dt}, leapfrog KDK, mixed-precision "
                               "brute-force gravity (FP32 pair forces on coordinates formed "
                               "in FP64 relative to each wavefront's target box, r^2 on the "
                               "matrix cores or the VALU as autotuned, compensated "
                               "accumulation)",
                   "n": n, "softening": eps, "dt": dt, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic, "kernel": f"{kname}_kernel",
                     "kernel_ms": round(k_mixed_s * 1e3, 4),
                     "flop_per_interaction": FLOP_PER_INTERACTION},
        "cpu_baseline": cpu,
        "tuning": tuning,
        "fp64": {"value": round(inter / t_fp64 / 1e9, 3), "unit": "G-interactions/s",
                 "ms_per_step": round(t_fp64 * 1e3, 4), "kernel_ms": round(k_fp64_s * 1e3, 4),
                 "achieved_tflops": round(FLOP_PER_INTERACTION * inter / k_fp64_s / 1e12, 3),
                 "frac_fp64_peak": round(FLOP_PER_INTERACTION * inter / k_fp64_s / 1e12
                                         / PEAK_FP64_TFLOPS, 4)},
        "mixed_over_fp64": round(t_fp64 / t_mixed, 3),
        "valu_kernel": valu,
        "force_error_vs_fp64": {"rms": err_rms, "max": err_max},
        "c1": c1,
        "barnes_hut": bh,
        "barnes_hut_fp64": bh64,
        "ensemble": ens,
        "sph": sph,
        "sph_c4": c4,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
