"""States divided over ranks (octhip.sharding), on the CPU over gloo.

* divide_range reproduces multicomm_divide_range's ranges (src/basic/multicomm.F90:996-1024)
  on hand-worked cases, and covers every object exactly once;
* two and three ranks, each with its own block of states, finish J, dJ/deps, dJ/dv and the
  weighted density with one all-reduce; the results match one rank holding every state
  to 1e-12 relative (only the order of the sums over states differs).  The per-rank work
  is the oracle's forward/backward on that rank's states: on a GPU box each rank would
  run the device adjoint of its batches instead.
"""
import multiprocessing as mp
import os
import socket
import sys
from pathlib import Path

import numpy as np
import pytest

from octhip import sharding

HERE = Path(__file__).resolve().parent


def test_divide_range_matches_the_reference_rule():
    # 10 objects over 4 ranks: the first 10 mod 4 = 2 ranks take 3, the others 2
    assert sharding.divide_range(10, 4) == [(0, 3), (3, 3), (6, 2), (8, 2)]
    assert sharding.divide_range(8, 2) == [(0, 4), (4, 4)]
    assert sharding.divide_range(5, 5) == [(0, 1), (1, 1), (2, 1), (3, 1), (4, 1)]
    # fewer objects than ranks: one each, the rest empty
    assert sharding.divide_range(3, 5) == [(0, 1), (1, 1), (2, 1), (0, 0), (0, 0)]
    rng = np.random.default_rng(3)
    for _ in range(200):
        n, p = int(rng.integers(0, 300)), int(rng.integers(1, 40))
        parts = sharding.divide_range(n, p)
        covered = [i for f, c in parts for i in range(f, f + c)]
        assert covered == list(range(n)) if p <= n else sorted(covered) == list(range(n))
        counts = [c for _, c in parts]
        if p <= n:
            assert max(counts) - min(counts) <= 1


def free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def run_world(world, outdir):
    env_path = os.pathsep.join([str(HERE.parent / "octopus-debian-package_amd"), str(HERE.parent / "oracle"),
                                str(HERE)])
    os.environ["PYTHONPATH"] = env_path + os.pathsep + os.environ.get("PYTHONPATH", "")
    sys.path.insert(0, str(HERE))
    import _sharding_worker as w

    port = free_port()
    ctx = mp.get_context("spawn")
    procs = [ctx.Process(target=w.main, args=(r, world, port, str(outdir))) for r in range(world)]
    for p in procs:
        p.start()
    for p in procs:
        p.join(timeout=240)
    for p in procs:
        if p.is_alive():
            p.kill()
    assert all(p.exitcode == 0 for p in procs), [p.exitcode for p in procs]
    return dict(np.load(outdir / f"world{world}.npz"))


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_sharded_gradient_and_density_match_one_rank(tmp_path):
    sys.path.insert(0, str(HERE))
    import _sharding_worker as w

    s, vol, z, tg, eps, dt, weights = w.problem()
    J, deps, dv, rho = w.local(s, vol, z, tg, eps, dt, weights, 0, z.shape[0])
    for world in (2, 3):
        r = run_world(world, tmp_path)
        for key, ref in (("J", J), ("deps", deps), ("dv", dv), ("rho", rho)):
            assert rel(r[key], ref) <= 1e-12, (world, key)
        for key, ref in (("J2", J), ("deps2", deps), ("dv2", dv)):
            assert rel(r[key], ref) <= 1e-12, (world, key)
    # the split is not a no-op: a rank's own share is not the total
    J1, *_ = w.local(s, vol, z, tg, eps, dt, weights, 0, 2)
    assert abs(J1 - J) > 1e-3 * abs(J)
