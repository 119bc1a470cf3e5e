"""The batched launch over uneven systems (octhip_batched_*) against per-pair calls.

Every kernel of the batched path runs the same code on the same data as the per-pair
entry points, so H-apply and Taylor steps must be bit-identical to them, for systems of
different sizes, pack sizes 8/16/32 and several batches of one system.  One pair is also
checked against the oracle (<= 1e-12 relative per state).
"""
import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err
from octhip import api, workloads

pytestmark = pytest.mark.gpu


def systems():
    out = []
    for i, (R, nst) in enumerate([(2.5, 5), (3.0, 12), (3.5, 40)]):
        s = workloads.hamiltonian_system(radius=R, spacing=0.3, seed=900 + i, nwells=3 + i, natoms=3, rc=1.2,
                                         name=f"uneven-small[{i}]")
        out.append((s, nst))
    return out


def make_pairs(sys_list):
    pairs, meta = [], []
    for s, nst in sys_list:
        hm = api.HamiltonianDevice.from_system(s)
        z = workloads.orthonormalise(workloads.gaussian_envelope_states(s, nst, width=2.0), s.mesh.vol_pp)
        for first in range(0, nst, 32):
            zz = z[first:first + 32]
            pairs.append((hm, api.Batch.from_states(zz, s.np_part)))
            meta.append((s, zz))
    return pairs, meta


def test_batched_matches_per_pair_calls(gpu):
    print("[batched vs per-pair]", flush=True)
    sl = systems()
    pairs, meta = make_pairs(sl)
    assert sorted({b.pack_size for _, b in pairs}) == [8, 16, 32]
    eager = [(hm, b.clone()) for hm, b in pairs]
    bp = api.BatchedPropagator(pairs, 0.01)
    bp.apply()
    for (hm, b), hp in zip(eager, bp.hpsi):
        ref = api.Batch(b.nst, hm.np, True, b.pack_size)
        api.hamiltonian_elec_apply_batch(hm, b, ref)
        gpu.cuda.synchronize()
        assert gpu.equal(hp.ff, ref.ff)
    te = api.Exponential(4)
    for step in range(3):
        bp.taylor_step()
        for hm, b in eager:
            te.apply_batch(hm, b, 0.01)
        gpu.cuda.synchronize()
        for (_, b), (_, e) in zip(pairs, eager):
            assert gpu.equal(b.ff[: b.nrows], e.ff[: e.nrows]), f"step {step}"
    # one pair against the oracle after the 3 steps
    s, zz = meta[-1]
    psi = workloads.pack(zz, s.np_part)
    oh = O.Hamiltonian(s.op, s.v, s.proj)
    for _ in range(3):
        oh.taylor(psi, 0.01)
    assert max_rel_err(pairs[-1][1].ff.cpu().numpy()[: s.np], psi[: s.np]) <= 1e-12


def test_batched_rejects_mixed_projector_sets(gpu):
    s0 = workloads.hamiltonian_system(radius=2.5, spacing=0.3, seed=910, nwells=2, natoms=2, rc=1.2)
    s1 = workloads.hamiltonian_system(radius=2.5, spacing=0.3, seed=911, nwells=2, natoms=0)
    p = []
    for s in (s0, s1):
        hm = api.HamiltonianDevice.from_system(s)
        p.append((hm, api.Batch.from_states(workloads.gaussian_envelope_states(s, 4), s.np_part)))
    with pytest.raises(api._lib.OcthipError):
        api.BatchedPropagator(p, 0.01)


def test_batched_graph_replays_bitwise(gpu):
    print("[batched graph]", flush=True)
    pairs, _ = make_pairs(systems())
    twin = [(hm, b.clone()) for hm, b in pairs]
    bp = api.BatchedPropagator(pairs, 0.01, with_apply=False)
    bq = api.BatchedPropagator(twin, 0.01, with_apply=False)
    g = bq.graph()
    other = twin[0][1].clone()
    hother = api.Batch(other.nst, twin[0][0].np, True, other.pack_size)
    for rep in range(3):
        bp.taylor_step()
        g.launch()
        api.hamiltonian_elec_apply_batch(twin[0][0], other, hother)  # unrelated work between replays
        gpu.cuda.synchronize()
        for (_, a), (_, b) in zip(pairs, twin):
            assert gpu.equal(a.ff[: a.nrows], b.ff[: b.nrows]), f"replay {rep}"


@pytest.mark.parametrize("order", [2, 3, 6])
def test_batched_other_orders_match_per_pair_calls(gpu, order):
    """Orders 2 and 3 (powers added in the last pass, fewer of them) and 6 (accumulated
    order by order) against per-pair Taylor steps of the same order."""
    print(f"[batched order {order}]", flush=True)
    pairs, _ = make_pairs(systems()[:2])
    eager = [(hm, b.clone()) for hm, b in pairs]
    bp = api.BatchedPropagator(pairs, 0.01, order=order, with_apply=False)
    te = api.Exponential(order)
    for step in range(2):
        bp.taylor_step()
        for hm, b in eager:
            te.apply_batch(hm, b, 0.01)
        gpu.cuda.synchronize()
        for (_, b), (_, e) in zip(pairs, eager):
            assert gpu.equal(b.ff[: b.nrows], e.ff[: e.nrows]), f"order {order} step {step}"

