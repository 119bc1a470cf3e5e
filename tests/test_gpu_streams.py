"""Batches on concurrent streams (the reference's propagation loop sets one stream per
batch, td/propagation_ops_elec.F90:236, accel_set_stream(ib)).

Every stream gets its own projector and Taylor work buffers, so H-apply and Taylor steps
issued on two streams at once must be bit-identical to the same calls issued one after
another on one stream.  A Taylor graph keeps the buffers it was captured with: eager
steps on the capturing stream between replays must not disturb the replays.  Releasing a
stream's buffers (octhip_hamiltonian_release_stream) lets the next call on it allocate a
fresh set with the same results.
"""
import pytest

from octhip import api, workloads

pytestmark = pytest.mark.gpu


def system():
    return workloads.hamiltonian_system(radius=3.5, spacing=0.3, seed=940, nwells=4, natoms=4, rc=1.2,
                                        name="streams")


def batches(s, n, pack=16):
    out = []
    for b in range(n):
        z = workloads.gaussian_envelope_states(s, pack, first=b * pack, width=2.0)
        out.append(api.Batch.from_states(z, s.np_part))
    return out


def test_two_streams_match_one_stream(gpu):
    print("[two streams]", flush=True)
    torch = gpu
    s = system()
    hm = api.HamiltonianDevice.from_system(s)
    ins = batches(s, 4)
    ref = [api.Batch(b.nst, s.np, True, b.pack_size) for b in ins]
    out = [api.Batch(b.nst, s.np, True, b.pack_size) for b in ins]
    for b, r in zip(ins, ref):
        api.hamiltonian_elec_apply_batch(hm, b, r)
    te = api.Exponential(4)
    seq = [b.clone() for b in ins]
    for _ in range(2):
        for b in seq:
            te.apply_batch(hm, b, 0.01)
    par = [b.clone() for b in ins]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    main = torch.cuda.current_stream()
    for st in streams:
        st.wait_stream(main)
    for i, (b, o) in enumerate(zip(ins, out)):
        with torch.cuda.stream(streams[i % 2]):
            api.hamiltonian_elec_apply_batch(hm, b, o)
    for _ in range(2):
        for i, b in enumerate(par):
            with torch.cuda.stream(streams[i % 2]):
                te.apply_batch(hm, b, 0.01)
    for st in streams:
        main.wait_stream(st)
    torch.cuda.synchronize()
    for i in range(len(ins)):
        assert torch.equal(out[i].ff, ref[i].ff), f"apply batch {i}"
        assert torch.equal(par[i].ff[: par[i].nrows], seq[i].ff[: seq[i].nrows]), f"Taylor batch {i}"


def test_graph_keeps_its_buffers(gpu):
    print("[graph buffers]", flush=True)
    torch = gpu
    s = system()
    hm = api.HamiltonianDevice.from_system(s)
    a, other = batches(s, 2)
    twin = a.clone()
    te = api.Exponential(4)
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        g = api.TaylorGraph(te, hm, [twin], 0.01)
        for rep in range(3):
            g.launch()
            te.apply_batch(hm, other, 0.01)  # eager step on the capturing stream between replays
            te.apply_batch(hm, a, 0.01)
            torch.cuda.synchronize()
            assert torch.equal(twin.ff[: twin.nrows], a.ff[: a.nrows]), f"replay {rep}"


def test_release_stream_frees_and_reallocates(gpu):
    print("[release stream]", flush=True)
    torch = gpu
    s = system()
    hm = api.HamiltonianDevice.from_system(s)
    b, = batches(s, 1)
    ref = api.Batch(b.nst, s.np, True, b.pack_size)
    api.hamiltonian_elec_apply_batch(hm, b, ref)
    st = torch.cuda.Stream()
    st.wait_stream(torch.cuda.current_stream())
    te = api.Exponential(4)
    x = b.clone()
    with torch.cuda.stream(st):
        te.apply_batch(hm, x, 0.01)        # allocates the stream's Taylor buffers
        hm.release_stream()                # frees them (after synchronising the stream)
        y = b.clone()
        te.apply_batch(hm, y, 0.01)        # allocates a new set
        out = api.Batch(b.nst, s.np, True, b.pack_size)
        api.hamiltonian_elec_apply_batch(hm, b, out)
        hm.release_stream()
        hm.release_stream()                # nothing left to release: no error
    torch.cuda.synchronize()
    assert torch.equal(x.ff[: x.nrows], y.ff[: y.nrows])
    assert torch.equal(out.ff, ref.ff)

