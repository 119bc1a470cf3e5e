"""The autotuners (octhip_hamiltonian_autotune, octhip_hamiltonian_autotune_pack) and their
persisted table.

CPU: the table file round-trips and malformed or invalid lines are rejected.
GPU: tuning picks a configuration, the H-apply with it still matches the oracle
(<= 1e-12 per state), and a second Hamiltonian of the same shape is configured from the
table without timing.
"""
import numpy as np
import pytest

import oracle as O
from helpers import max_rel_err, small_sphere
from octhip import _lib, api, workloads


def test_table_roundtrip_and_validation(tmp_path):
    good = tmp_path / "t.txt"
    good.write_text("# comment\n1000 1200 3 5 2 16 2 16 4 1 12.500\n")
    api.tuning_load(good)
    out = tmp_path / "o.txt"
    api.tuning_save(out)
    lines = [ln for ln in out.read_text().splitlines() if not ln.startswith("#")]
    assert "1000 1200 3 5 2 16 2 16 4 1 12.500" in lines
    # block-size lines
    good.write_text("pack 1000 1200 3 40 16 55.000\n")
    api.tuning_load(good)
    api.tuning_save(out)
    assert "pack 1000 1200 3 40 16 55.000" in out.read_text().splitlines()
    bad = tmp_path / "bad.txt"
    bad.write_text("pack 1000 1200 3 40 24 55.0\n")  # not a power of two
    with pytest.raises(_lib.OcthipError):
        api.tuning_load(bad)
    bad.write_text("1 2 3\n")
    with pytest.raises(_lib.OcthipError):
        api.tuning_load(bad)
    bad.write_text("1000 1200 3 5 7 16 2 16 4 1 12.5\n")  # kernel 7 does not exist
    with pytest.raises(_lib.OcthipError):
        api.tuning_load(bad)


@pytest.mark.gpu
def test_autotune_keeps_parity_and_reuses_table(gpu, tmp_path):
    print("[autotune]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)
    hm = api.HamiltonianDevice.from_system(s)
    us = hm.autotune(ldf=4, reps=2, force=True)
    assert us > 0
    z = workloads.gaussian_envelope_states(s, 8, width=2.0)
    psib = api.Batch.from_states(z, s.np_part)
    hpsib = api.Batch(8, s.np, True)
    api.hamiltonian_elec_apply_batch(hm, psib, hpsib)
    gpu.cuda.synchronize()
    ref = O.Hamiltonian(s.op, s.v, s.proj).apply(workloads.pack(z, s.np_part))
    assert max_rel_err(hpsib.ff.cpu().numpy(), ref) <= 1e-12
    hm2 = api.HamiltonianDevice.from_system(s)
    assert hm2.autotune(ldf=4) == us          # from the table, no timing
    for k in ("kernel", "lane_chunk", "lanes_per_thread", "zseg"):
        assert hm2.kinetic.tuning(k) == hm.kinetic.tuning(k)
    api.tuning_save(tmp_path / "table.txt")
    assert len((tmp_path / "table.txt").read_text().splitlines()) >= 2


@pytest.mark.gpu
def test_autotune_pack_picks_a_block_size_and_reuses_it(gpu):
    print("[autotune_pack]", flush=True)
    s = small_sphere(radius=3.0, spacing=0.3, natoms=3, rc=1.2)
    hm = api.HamiltonianDevice.from_system(s)
    for nst, top in ((5, 8), (40, 64), (200, 64)):
        pack, us = hm.autotune_pack(nst, reps=2, force=True)
        assert pack in (8, 16, 32, 64) and pack <= top and us > 0
        # the same question again comes from the table
        assert hm.autotune_pack(nst) == (pack, us)
        # and the chosen block size runs a Taylor step that matches the oracle
        z = workloads.gaussian_envelope_states(s, min(nst, pack), width=2.0)
        psib = api.Batch.from_states(z, s.np_part, pack_size=pack)
        api.Exponential(4).apply_batch(hm, psib, 0.01)
        gpu.cuda.synchronize()
        ref = workloads.pack(z, s.np_part, pack)
        O.Hamiltonian(s.op, s.v, s.proj).taylor(ref, 0.01)
        assert max_rel_err(psib.ff.cpu().numpy()[: s.np], ref[: s.np]) <= 1e-12
