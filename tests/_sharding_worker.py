"""Worker of tests/test_sharding.py: one rank of a gloo group on the CPU.  Each rank takes
its block of states, runs the oracle's forward/backward on them and finishes J, dJ/deps,
dJ/dv and the density with octhip.sharding's single all-reduce."""
import os
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "octopus-debian-package_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import numpy as np


def problem():
    from helpers import small_sphere
    from octhip import workloads

    s = small_sphere(radius=1.8, spacing=0.3, natoms=2, nwells=3, rc=0.9)
    vol = s.mesh.vol_pp
    nst = 5
    z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=1.5), vol)
    tg = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=1.0, first=100), vol)
    nsteps, dt = 5, 0.05
    t = (np.arange(nsteps) + 0.5) * dt
    eps = 0.3 * np.sin(0.9 * t + 0.2) + 0.1
    weights = np.linspace(2.0, 1.0, nst)
    return s, vol, z, tg, eps, dt, weights


def local(s, vol, z, tg, eps, dt, weights, first, count):
    """J, dJ/deps, dJ/dv and the weighted density of states first..first+count-1."""
    import oracle as O
    from octhip import workloads

    if count == 0:
        return 0.0, np.zeros(len(eps)), np.zeros(s.np), np.zeros(s.np)
    zz, tt = z[first:first + count], tg[first:first + count]
    h = O.Hamiltonian(s.op, s.v, s.proj, xyz=s.mesh.x[: s.np], field=np.zeros(3))
    adj = O.TaylorAdjoint(h, s.mesh.x[: s.np], [1.0, 0.0, 0.0], dt, 4, vol)
    psiN = adj.forward(workloads.pack(zz, s.np_part), eps)
    J, chiN = O.overlap_objective(workloads.pack(tt, s.np_part), psiN, s.np, vol)
    _, deps, dv = adj.backward(chiN)
    rho = np.zeros(s.np)
    w = np.zeros(psiN.shape[1] // 2)
    w[:count] = weights[first:first + count]
    O.density_accumulate(np.ascontiguousarray(psiN), s.np, len(w), True, w, rho)
    return J, deps, dv, rho


def main(rank, world, port, outdir):
    import torch.distributed as dist
    from octhip import sharding

    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist.init_process_group("gloo", rank=rank, world_size=world)
    try:
        s, vol, z, tg, eps, dt, weights = problem()
        nst = z.shape[0]
        first, count = sharding.local_states(nst, rank, world)
        J, deps, dv, rho = local(s, vol, z, tg, eps, dt, weights, first, count)
        J, deps, dv, rho = sharding.allreduce_sums([J, deps, dv, rho])
        # the same through the gradient helper (its own all-reduce)
        G = sharding.sharded_objective_gradient(
            lambda f, c: local(s, vol, z, tg, eps, dt, weights, f, c)[:3], nst)
        if rank == 0:
            np.savez(os.path.join(outdir, f"world{world}.npz"), J=J, deps=deps, dv=dv, rho=rho,
                     J2=G[0], deps2=G[1], dv2=G[2])
    finally:
        dist.destroy_process_group()
