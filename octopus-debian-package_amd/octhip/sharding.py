"""States divided over ranks, one GPU per rank: how the path shards.

The reference distributes states over the processes of its states group (distributed_t
over st%nst, src/basic/distributed.F90:75-148, with the ranges of multicomm_divide_range,
src/basic/multicomm.F90:960-1032).  Here each rank owns a contiguous block of states, packs
them into its own batches on its own GPU and propagates them with no exchange at all: the
Hamiltonian is the same on every rank and the states do not couple during a Taylor step.
What couples them are sums over states, finished by one all-reduce each time they are
needed:

* the density, sum_st w_st |psi_st|^2 (density_calc_accumulate, then the reduction over
  the states group in density_calc_end, src/states/density.F90);
* the optimal-control objective J = sum_st |<phi_st|psi_st>|^2 and its gradients dJ/deps
  (one value per step) and dJ/dv (one value per mesh point), which the device adjoint
  returns per batch.

``allreduce_sums`` packs every such quantity of a rank into one float64 buffer and reduces
it with a single collective: over RCCL (backend "nccl") for device tensors, over gloo for
host tensors.  Only a one-GPU box is used in this build; the split and the reduction are
exercised by multi-process CPU tests over gloo.
"""
from __future__ import annotations

import numpy as np


def divide_range(nobjs: int, nprocs: int) -> list[tuple[int, int]]:
    """(first, count) of each rank, 0-based: the non-ScaLAPACK branch of
    multicomm_divide_range (src/basic/multicomm.F90:996-1024).  With at least as many
    objects as ranks, the first nobjs mod nprocs ranks take one object more; with fewer,
    rank r < nobjs takes object r and the others nothing."""
    if nobjs < 0 or nprocs < 1:
        raise ValueError("need nobjs >= 0 and nprocs >= 1")
    out = []
    if nprocs <= nobjs:
        for rank in range(nprocs):
            jj = nobjs // nprocs
            ii = nobjs - jj * nprocs
            if ii > 0 and rank < ii:
                jj += 1
                start = rank * jj
            else:
                final = nobjs - (nprocs - rank - 1) * jj  # 1-based last object
                start = final - jj
            out.append((start, jj))
    else:
        for rank in range(nprocs):
            out.append((rank, 1) if rank < nobjs else (0, 0))
    return out


def local_states(nst: int, rank: int, world: int) -> tuple[int, int]:
    """(first state, number of states) owned by this rank."""
    return divide_range(nst, world)[rank]


def allreduce_sums(parts, group=None):
    """Sum every array (or scalar) of ``parts`` over the ranks of ``group`` with one
    collective; returns them in the same shapes.  Device tensors go over the process
    group's backend as they are; numpy arrays and scalars travel as host float64 (the
    form gloo takes)."""
    import torch
    import torch.distributed as dist

    shapes, flats, device = [], [], None
    for p in parts:
        if isinstance(p, torch.Tensor):
            device = p.device if device is None else device
            flats.append(p.detach().reshape(-1).to(torch.float64))
            shapes.append(("t", tuple(p.shape)))
        else:
            a = np.asarray(p, dtype=np.float64)
            flats.append(torch.from_numpy(a.reshape(-1).copy()))
            shapes.append(("n", a.shape))
    dev = device if device is not None else torch.device("cpu")
    buf = torch.cat([f.to(dev) for f in flats])
    dist.all_reduce(buf, op=dist.ReduceOp.SUM, group=group)
    out, k = [], 0
    for (kind, shape), f in zip(shapes, flats):
        n = f.numel()
        piece = buf[k:k + n].reshape(shape)
        k += n
        if kind == "t":
            out.append(piece)
        else:
            arr = piece.cpu().numpy().copy()
            out.append(float(arr) if arr.shape == () else arr)
    return out


def sharded_objective_gradient(local_fn, nst: int, group=None):
    """J, dJ/deps and dJ/dv of all nst states: this rank evaluates ``local_fn(first, count)``
    on its own block of states (returning (J, dJ/deps, dJ/dv) of those states, e.g. from
    the device adjoint of its batches) and one all-reduce sums the three over the ranks."""
    import torch.distributed as dist

    rank, world = dist.get_rank(group), dist.get_world_size(group)
    first, count = local_states(nst, rank, world)
    J, deps, dv = local_fn(first, count)
    return tuple(allreduce_sums([J, deps, dv], group))
