#!/usr/bin/env python3
"""Benchmark of the MI355X H-apply path (one GPU).

Default workload (the configuration the headline metric is quoted on): complex FP64
H = -1/2 Laplacian + v_loc + V_nl on the sphere R = 12 bohr, h = 0.25 (462,781 points),
order-4 star stencil, 64 states in two packed batches of 32, 16 Gaussian wells,
16 atoms x 4 Kleinman-Bylander projectors.  One step = H applied to both batches.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hamiltonian|laplacian|taylor|adjoint|batched]

Prints one JSON line.  `value` is state-points/s with the states already resident in
HBM; `roofline` is the dominant kernel (k_zmarch, the fused stencil + local + ket pass)
timed with HIP events on its launch stream over the timed region; `cpu_baseline` is the
CPU oracle (C/OpenMP restatement of the reference CPU path) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "octopus-debian-package_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "FP64 H-apply state-points/s & %HBM peak; batched vs plain; adjoint step time"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def tuning(args) -> dict:
    """Tuning knobs given on the command line (unset ones keep the library defaults)."""
    t = {}
    for k in ("kernel", "xcd_chunk", "gather_tile", "zseg", "lane_chunk", "lanes_per_thread"):
        v = getattr(args, k)
        if v is not None:
            t[k] = v
    return t


PROFILE = ROOT / "profiles" / "k_zmarch_hamiltonian_R12_h025_pack32.md"


def profiled_traffic():
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a 1 GiB copy), or None."""
    try:
        for line in PROFILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1].split()[0])
    except OSError:
        pass
    return None


def hamiltonian_bytes(s, nst: int, complex_: bool = True) -> int:
    """Compulsory HBM bytes of one H-apply of a batch: psi read over np_part rows, H psi
    written over np rows, v, the projector matrices, maps and their points, weights."""
    es = 16 if complex_ else 8
    b = es * nst * (s.np_part + s.np) + 8 * s.np + 8 * 25
    if s.proj is not None:
        o = s.proj.offsets
        b += int(sum(8 * int(npt) * int(npj) + 4 * int(npt) + 8 * int(npt) for npt, npj in zip(o[:, 0], o[:, 1])))
    return b


def host_cpu() -> dict:
    """Model, sockets and cores of this host (lscpu)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "cpus_online": info.get("CPU(s)")}


def cpu_baseline_hamiltonian(s, nst: int, budget_s: float) -> dict:
    """The oracle's H-apply (its C/OpenMP restatement of the reference CPU path, lanes
    innermost as the reference leaf vectorises them) in a timing build for this host:
    median of >= 10 applies after 2 warm-ups, bounded by budget_s."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from octhip import workloads
    flags = O.use_timing_build()
    z = workloads.gaussian_envelope_states(s, nst)
    psi = workloads.pack(z, s.np_part)
    h = O.Hamiltonian(s.op, s.v, s.proj)
    for _ in range(2):
        h.apply(psi)
    times = []
    t_start = time.perf_counter()
    while len(times) < 10 or (time.perf_counter() - t_start < budget_s and len(times) < 50):
        t0 = time.perf_counter()
        h.apply(psi)
        times.append(time.perf_counter() - t0)
    per = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": s.np * nst / per, "unit": "state-points/s", "cores": threads, "kind": "port",
            "sample": f"median of {len(times)} H-applies (after 2 warm-ups) of one {nst}-state complex batch on the "
                      f"full R=12 sphere: {per * 1e3:.1f} ms each, OpenMP oracle on {threads} threads",
            "build": flags, "host": host_cpu()}


def bench_hamiltonian(args, torch, api, _lib, workloads):
    log("building the R=12 h=0.25 sphere workload")
    s = workloads.hamiltonian_system()
    nst_total, pack = 64, 32
    hm = api.HamiltonianDevice.from_system(s, lattice=not args.no_lattice)
    hm.tune(bra_mode=args.bra, **tuning(args))
    batches, outs = [], []
    for b in range(nst_total // pack):
        z = workloads.gaussian_envelope_states(s, pack, first=b * pack)
        batches.append(api.Batch.from_states(z, s.np_part))
        outs.append(api.Batch(pack, s.np, True))
    torch.cuda.synchronize()
    # batch b runs on stream b mod --streams, as the reference's propagation loop sets one
    # stream per batch (td/propagation_ops_elec.F90:236); every stream has its own work set
    main = torch.cuda.current_stream()
    streams = [main] if args.streams <= 1 else [torch.cuda.Stream() for _ in range(args.streams)]

    def step():
        if len(streams) == 1:
            for pb, hb in zip(batches, outs):
                api.hamiltonian_elec_apply_batch(hm, pb, hb)
            return
        for st in streams:
            st.wait_stream(main)
        for i, (pb, hb) in enumerate(zip(batches, outs)):
            with torch.cuda.stream(streams[i % len(streams)]):
                api.hamiltonian_elec_apply_batch(hm, pb, hb)
        for st in streams:
            main.wait_stream(st)

    log(f"np={s.np} np_part={s.np_part} warmup={args.warmup} steps={args.steps} streams={len(streams)}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # timed region: HIP events around the dominant kernel only (every event pair costs
    # device time between kernels)
    # the fused pass is k_zmarch on the lattice, k_apply (tiled gather) with --no-lattice
    kname = "k_apply" if args.no_lattice or args.kernel == 1 else "k_zmarch"
    _lib.profiling_select(kname)
    _lib.profiling_start()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.profiling_stop()
    ms_step = (t1 - t0) * 1e3 / args.steps
    tot, n = _lib.profiling_get(kname)
    if n == 0:
        raise RuntimeError(f"no {kname} launches were recorded")
    k_apply_us = tot * 1e3 / n
    # per-kernel breakdown from a separate pass with every kernel timed
    _lib.profiling_select(None)
    _lib.profiling_start()
    for _ in range(min(args.steps, 5)):
        step()
    torch.cuda.synchronize()
    _lib.profiling_stop()
    kern = {}
    for k in (kname, "k_bra", "k_bra_reduce", "k_ket", "k_ket_sum", "k_zero_rows"):
        tot, n = _lib.profiling_get(k)
        kern[k] = {"avg_us": tot * 1e3 / n if n else None, "launches": n}
    kern[kname]["avg_us_timed_region"] = k_apply_us
    bytes_batch = hamiltonian_bytes(s, pack)
    t_apply = k_apply_us * 1e-6
    achieved = bytes_batch / t_apply / 1e9
    value = s.np * nst_total / (ms_step * 1e-3)
    happly_gbs = bytes_batch * (nst_total // pack) / (ms_step * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": value, "unit": "state-points/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Gaussian-envelope states, Gaussian wells, synthetic KB projectors)",
        "config": {"workload": "complex H-apply, sphere R=12 h=0.25 (np=462781, np_part=564189), "
                               "order-4 star, 64 states as 2 packed batches of 32, 16 wells, "
                               "16 atoms x 4 KB projectors (rc=3)",
                   "states": nst_total, "pack_size": pack, "np": s.np, "np_part": s.np_part,
                   "streams": len(streams)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": args.traffic if args.traffic is not None else profiled_traffic(),
                     "traffic_source": "--traffic" if args.traffic is not None else str(PROFILE.relative_to(ROOT)),
                     "kernel": kname, "algorithmic_bytes_per_launch": bytes_batch,
                     "avg_launch_us": k_apply_us},
        "h_apply_hbm_frac": happly_gbs / HBM_PEAK_GBS,
        "kernels": kern,
    }
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on one batch")
        res["cpu_baseline"] = cpu_baseline_hamiltonian(s, pack, args.cpu_seconds)
    return res


def bench_laplacian(args, torch, api, _lib, workloads):
    from octhip.stencil import star_points
    log("building the 64^3 cube workload")
    s = workloads.laplacian_system()
    f = workloads.gaussian_functions(s, 8)
    fi = api.Batch.from_states(f, s.np_part, 8)
    fo = api.Batch(8, s.np, False, 8)
    op = api.NLOperatorDevice(s.op, star_points(4))
    op.tune(**tuning(args))
    flush = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB, evicts the MALL
    for _ in range(args.warmup):
        api.nl_operator_operate_batch(op, fi, fo)
    torch.cuda.synchronize()
    _lib.profiling_start()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        flush.zero_()
        api.nl_operator_operate_batch(op, fi, fo)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.profiling_stop()
    kname = "k_apply" if args.kernel == 1 else "k_zmarch"
    tot, n = _lib.profiling_get(kname)
    t_k = tot * 1e-3 / n
    b = 8 * 8 * (s.np_part + s.np) + 8 * 25
    achieved = b / t_k / 1e9
    return {"metric": "FP64 Laplacian state-points/s", "value": s.np * 8 / t_k, "unit": "state-points/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": (t1 - t0) * 1e3 / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "real Laplacian, 64^3 cube h=0.3, 8 functions, pack 8, "
                                                        "MALL flushed between reps"},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kname,
                         "algorithmic_bytes_per_launch": b, "avg_launch_us": t_k * 1e6}}


def bench_taylor(args, torch, api, _lib, workloads):
    """Taylor-4 time steps over 128 states (4 batches of 32) on the R=12 sphere, launched
    kernel by kernel (eager) and as one captured HIP graph per step."""
    log("building the R=12 h=0.25 sphere workload (Taylor)")
    s = workloads.hamiltonian_system()
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(bra_mode=args.bra, **tuning(args))
    batches = [api.Batch.from_states(workloads.gaussian_envelope_states(s, 32, first=32 * b), s.np_part)
               for b in range(4)]
    te = api.Exponential(4)
    dt = args.dt

    def eager():
        for pb in batches:
            te.apply_batch(hm, pb, dt)

    graph = api.TaylorGraph(te, hm, batches, dt)

    def timed(fn, prof):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        if prof:
            _lib.profiling_start()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if prof:
            _lib.profiling_stop()
        return (t1 - t0) * 1e3 / args.steps

    ms_eager = timed(eager, True)
    kern = {k: _lib.profiling_get(k) for k in ("k_zmarch", "k_apply", "k_bra", "k_bra_reduce", "k_ket",
                                               "k_ket_sum", "k_zero_rows")}
    ms_graph = timed(graph.launch, False)
    ms = min(ms_graph, ms_eager)
    return {"metric": "FP64 Taylor-4 step state-point H-applies/s (128 states)",
            "value": s.np * 128 * 4 / (ms_graph * 1e-3),
            "unit": "state-point H-applies/s", "ms_per_step": ms_graph, "steps": args.steps,
            "warmup": args.warmup, "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Taylor-4 dt={dt}, 128 states as 4 batches of 32, R=12 sphere h=0.25, "
                                   "one HIP graph per step"},
            "eager_ms_per_step": ms_eager, "graph_speedup": ms_eager / ms_graph, "best_ms_per_step": ms,
            "kernels_eager": {k: {"avg_us": v[0] * 1e3 / v[1] if v[1] else None, "launches": v[1]}
                              for k, v in kern.items()}}


def bench_adjoint(args, torch, api, _lib, workloads):
    """100 Taylor-4 steps (dt 0.02) on the R=10 h=0.3 sphere with 32 states and a dipole
    field, forward with checkpoints every 10 steps, then the adjoint (dJ/deps, dJ/dv)."""
    from octhip.autograd import overlap_objective
    log("building the R=10 h=0.3 sphere workload (adjoint)")
    s = workloads.adjoint_system()
    nsteps, dt, K = 100, 0.02, 10
    z, tg, eps = workloads.adjoint_problem(s, 32, nsteps, dt)
    vol = s.mesh.vol_pp
    hm = api.HamiltonianDevice.from_system(s)
    hm.tune(bra_mode=args.bra, **tuning(args))
    adj = api.Adjoint(hm, s.mesh.x, [1.0, 0.0, 0.0], 6, nsteps, dt, vol, checkpoint_every=K)
    psi0 = api.Batch.from_states(z, s.np_part)
    phib = api.Batch.from_states(tg, s.np_part)
    te = api.Exponential(4)

    def plain():  # the forward propagation alone, no checkpoints
        p = psi0.clone()
        for e in eps:
            hm.set_dipole_field([e, 0.0, 0.0])
            te.apply_batch(hm, p, dt)

    def fwd_bwd():
        p = psi0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        adj.forward(p, eps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J, chi = overlap_objective(phib, p, s.np, vol)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        adj.backward(chi)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        return t1 - t0, t3 - t2, float(J)

    warm = max(1, min(args.warmup, 2))  # whole 100-step propagations: two are warm enough
    for _ in range(warm):
        plain()
        fwd_bwd()
    torch.cuda.synchronize()
    reps = max(1, min(args.steps, 5))
    t0 = time.perf_counter()
    for _ in range(reps):
        plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    # timed without per-kernel events (as the plain forward), then one pass with events
    # around every kernel for the breakdown (the events cost device time between kernels)
    f = [fwd_bwd() for _ in range(reps)]
    t_f = min(x[0] for x in f)
    t_b = min(x[1] for x in f)
    _lib.profiling_select(None)
    _lib.profiling_start()
    fwd_bwd()
    _lib.profiling_stop()
    acc_ms, acc_n = _lib.profiling_get("k_adj_accum")
    breakdown = {}
    for k in ("k_zmarch", "k_apply", "k_bra", "k_bra_reduce", "k_ket", "k_ket_sum", "k_adj_accum", "k_zero_rows"):
        t, n = _lib.profiling_get(k)
        breakdown[k] = {"total_ms_per_rep": t, "launches_per_rep": n}
    step_ms = t_b * 1e3 / nsteps
    return {"metric": "FP64 adjoint step time (Taylor-4 backward + gradient accumulation)",
            "value": step_ms, "unit": "ms/step", "higher_is_better": False, "n_gpus": 1,
            "steps": reps, "warmup": warm, "ms_per_step": step_ms, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded states, targets, field)",
            "config": {"workload": "adjoint of 100 Taylor-4 steps dt=0.02, sphere R=10 h=0.3 "
                                   f"(np={s.np}), 32 states, dipole x*eps(t), checkpoints every 10"},
            "forward_ms": t_f * 1e3, "backward_ms": t_b * 1e3, "plain_forward_ms": t_plain * 1e3,
            "backward_over_forward_step": t_b / t_plain, "J": f[-1][2],
            "k_adj_accum_avg_us": acc_ms * 1e3 / acc_n if acc_n else None, "kernels_fwd_bwd": breakdown}


def bench_batched(args, torch, api, _lib, workloads):
    """8 uneven spheres (R = 6..14, h = 0.25, 8..256 states): one H-apply + 50 Taylor-4
    steps, as one batched launch sequence over all (system, batch) pairs and as plain
    per-batch launches system by system; block sizes and launch shapes from the library's
    tuner (octhip_hamiltonian_autotune_pack); both paths use the same blocks."""
    log("building the 8 uneven systems")
    systems = workloads.uneven_systems()
    nsteps, dt = 50, args.dt
    plan = []
    for s, nst in systems:
        hm = api.HamiltonianDevice.from_system(s)
        # block size by the time of what the workload runs most (Taylor steps), chosen by
        # the library's tuner, which also leaves that row length's launch shape configured
        pack, us = hm.autotune_pack(nst, reps=3)
        best = (us, pack, int(np.log2(2 * pack)), us / -(-nst // pack))
        batches = []
        for first in range(0, nst, best[1]):
            n = min(best[1], nst - first)
            z = workloads.gaussian_envelope_states(s, n, first=first)
            batches.append(api.Batch.from_states(z, s.np_part, pack_size=best[1]))
        plan.append((s, nst, hm, best[1], batches))
        log(f"system np={s.np} nst={nst}: pack {best[1]} ({best[3]:.1f} us per batch Taylor step)")
    te = api.Exponential(4)

    def run_plain():
        for s, nst, hm, pack, batches in plan:
            outs = [api.Batch(b.nst, s.np, True, b.pack_size) for b in batches]
            for b, o in zip(batches, outs):
                api.hamiltonian_elec_apply_batch(hm, b, o)
            for _ in range(nsteps):
                for b in batches:
                    te.apply_batch(hm, b, dt)

    pairs = [(hm, b) for s, nst, hm, pack, batches in plan for b in batches]
    bp = api.BatchedPropagator(pairs, dt)

    def run_batched():
        bp.apply()
        for _ in range(nsteps):
            bp.taylor_step()

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    timed(run_plain)  # warm-up (allocations of the per-handle work buffers)
    timed(run_batched)
    tp = min(timed(run_plain) for _ in range(max(1, min(args.steps, 3))))
    tb = min(timed(run_batched) for _ in range(max(1, min(args.steps, 3))))
    sp = sum(s.np * nst for s, nst, *_ in plan)
    return {"metric": "batched vs plain launches (1 H-apply + 50 Taylor-4 steps, 8 uneven systems)",
            "value": tp / tb, "unit": "speedup (plain time / batched time)", "higher_is_better": True,
            "n_gpus": 1, "steps": nsteps, "warmup": 1, "ms_per_step": tb * 1e3 / nsteps, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "8 spheres R=6,8,9,10,11,12,13,14 h=0.25 with 8..256 states, 1 H-apply + "
                                   "50 Taylor-4 steps", "packs": [p[3] for p in plan],
                       "pairs": len(pairs)},
            "plain_s": tp, "batched_s": tb,
            "batched_state_point_happlies_per_s": sp * (1 + 4 * nsteps) / tb,
            "plain_state_point_happlies_per_s": sp * (1 + 4 * nsteps) / tp}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="hamiltonian", choices=["hamiltonian", "laplacian", "taylor", "adjoint", "batched"])
    ap.add_argument("--kernel", type=int, default=None, help="0 auto, 1 gather, 2 z-march")
    ap.add_argument("--xcd-chunk", type=int, default=None)
    ap.add_argument("--gather-tile", type=int, default=None)
    ap.add_argument("--zseg", type=int, default=None)
    ap.add_argument("--lane-chunk", type=int, default=None)
    ap.add_argument("--lanes-per-thread", type=int, default=None)
    ap.add_argument("--bra", type=int, default=1)
    ap.add_argument("--streams", type=int, default=1, help="H-apply: batches spread over this many streams")
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="H-apply workload only (skip the batched-vs-plain and adjoint parts of the metric)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-lattice", action="store_true", help="index neighbours through ri, not the site map")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one device copy of this size first (reference point for HBM counters)")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per k_apply launch measured with rocprofv3 (FETCH+WRITE, scaled)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(the path shards by state batches/systems without communication; see the design notes)",
              file=sys.stderr)
        return 2
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    from octhip import _lib, api, workloads
    _lib.init(0)
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_calibration_mib} MiB read + written")
        del src, dst
    fn = {"hamiltonian": bench_hamiltonian, "laplacian": bench_laplacian, "taylor": bench_taylor,
          "adjoint": bench_adjoint, "batched": bench_batched}[args.workload]
    res = fn(args, torch, api, _lib, workloads)
    if args.workload == "hamiltonian" and not args.no_extras:
        # the other two parts of the metric, timed in the same run
        log("batched vs plain launches: 8 uneven spheres, 1 H-apply + 50 Taylor-4 steps")
        b = bench_batched(args, torch, api, _lib, workloads)
        res["batched_vs_plain"] = {"value": b["value"], "unit": b["unit"], "plain_s": b["plain_s"],
                                   "batched_s": b["batched_s"], "workload": b["config"]["workload"],
                                   "packs": b["config"]["packs"]}
        log("adjoint: 100 Taylor-4 steps on the R=10 h=0.3 sphere, 32 states, checkpoints every 10")
        a = bench_adjoint(args, torch, api, _lib, workloads)
        res["adjoint_over_forward"] = {"value": a["backward_over_forward_step"],
                                       "unit": "adjoint backward time / plain forward time",
                                       "adjoint_step_ms": a["value"], "backward_ms": a["backward_ms"],
                                       "forward_with_checkpoints_ms": a["forward_ms"],
                                       "plain_forward_ms": a["plain_forward_ms"], "workload": a["config"]["workload"]}
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
