"""X(nl_operator_operate_batch) with points = OP_INNER / OP_OUTER (the op%inner / op%outer
runs of nl_operator_build, src/grid/nl_operator.F90:457-514), on the device.

* The split itself, on the CPU from the operator's own ri / rimap_inv: an inner run's
  points have every neighbour among the inner points; every outer run has a point with a
  neighbour among the boundary rows; inner and outer runs cover every point once.
* On the device: OP_INNER writes exactly the inner runs' rows and OP_OUTER exactly the
  outer runs' rows (the other rows keep a NaN sentinel); each part is bit-identical to
  OP_ALL on its rows and to the oracle's leaf there, for real and complex rows, with and
  without a factor, on a sphere (z-march for OP_ALL) and on the gather kernel.
"""
import numpy as np
import pytest

import oracle as O
from helpers import small_sphere
from octhip import _lib, api, workloads
from octhip.stencil import star_points


def split(op):
    """Point masks of the inner and outer runs, by the reference's rule."""
    n = op.n
    ri = np.asarray(op.ri, np.int64).reshape(-1, n)
    rinv = np.asarray(op.rimap_inv, np.int64)
    inner = np.zeros(op.np, bool)
    for l in range(op.nri):
        if rinv[l + 1] + ri[l].max() <= op.np:
            inner[rinv[l]:rinv[l + 1]] = True
    return inner, ~inner


def test_split_rule_on_the_cpu():
    s = small_sphere(radius=3.0, spacing=0.3, natoms=0)
    op = s.op
    inner, outer = split(op)
    assert inner.any() and outer.any()
    ri = np.asarray(op.ri, np.int64).reshape(-1, op.n)
    rimap = np.repeat(np.arange(op.nri), np.diff(np.asarray(op.rimap_inv)))
    nb = np.arange(op.np)[:, None] + ri[rimap]          # neighbours of every point
    assert np.all(nb[inner] < op.np)                   # inner: only inner neighbours
    rinv = np.asarray(op.rimap_inv)
    for l in range(op.nri):
        pts = np.arange(rinv[l], rinv[l + 1])
        if outer[pts[0]]:
            assert np.any(nb[pts] >= op.np)            # outer: touches the boundary


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("cplx", [False, True])
def test_inner_outer_parts_match_all(gpu, kernel, cplx):
    print(f"[operate_batch inner/outer kernel={kernel} complex={cplx}]", flush=True)
    s = small_sphere(radius=3.5, spacing=0.3, natoms=0)
    inner, outer = split(s.op)
    opd = api.NLOperatorDevice(s.op, star_points(4))
    if kernel:
        opd.tune(kernel=kernel)
    nst = 8
    z = workloads.gaussian_envelope_states(s, nst, width=2.0)
    if not cplx:
        z = np.ascontiguousarray(z.real)
    fi = api.Batch.from_states(z, s.np_part)
    host_in = fi.ff.cpu().numpy()
    for factor in (None, -0.5):
        full = api.Batch(nst, s.np, cplx)
        api.nl_operator_operate_batch(opd, fi, full, factor)
        parts = {}
        for name, pts in (("inner", _lib.OP_INNER), ("outer", _lib.OP_OUTER)):
            b = api.Batch(nst, s.np, cplx)
            b.ff.fill_(float("nan"))
            api.nl_operator_operate_batch(opd, fi, b, factor, points=pts)
            parts[name] = b.ff.cpu().numpy()
        gpu.cuda.synchronize()
        f = full.ff.cpu().numpy()
        ref = O.operate_ri_vec(s.op, host_in, 1.0 if factor is None else factor)
        assert np.array_equal(f, ref)
        assert np.array_equal(parts["inner"][inner], f[inner])
        assert np.all(np.isnan(parts["inner"][outer]))
        assert np.array_equal(parts["outer"][outer], f[outer])
        assert np.all(np.isnan(parts["outer"][inner]))


@pytest.mark.gpu
def test_bad_points_selector_is_rejected(gpu):
    s = small_sphere(radius=2.0, spacing=0.3, natoms=0)
    opd = api.NLOperatorDevice(s.op, star_points(4))
    fi = api.Batch(4, s.np_part, True)
    fo = api.Batch(4, s.np, True)
    with pytest.raises(_lib.OcthipError):
        api.nl_operator_operate_batch(opd, fi, fo, points=7)
