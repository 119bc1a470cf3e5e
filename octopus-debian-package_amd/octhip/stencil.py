"""Star finite-difference Laplacian and the run-compressed operator index.

* stencil geometry: src/grid/stencil_star.F90:59-93 -- for each axis, offsets -order..order
  with the centre taken once (it is point index `order`, 0-based, i.e. 5th of 25);
* weights: src/grid/derivatives.F90:639-757 solves mat * w = rhs over the polynomials
  1, x^k, y^k, z^k (k=1..2*order, src/grid/stencil_star.F90:97-116) with rhs the Laplacian
  of each polynomial at 0; `weights_polyfit` restates that solve, `weights_closed_form`
  gives the textbook central-difference coefficients it must reproduce;
* run compression: src/grid/nl_operator.F90:308-448 -- consecutive points whose relative
  neighbour offsets are identical share one row of `ri`; a new run also starts where a
  point first sees a boundary neighbour (force_change); rimap_inv holds run starts and
  the final np (0-based half-open runs, as the C leaf src/grid/operate_inc.c:54-56 reads
  them).  NLOperatorCompactBoundaries defaults to no (src/grid/nl_operator.F90:210-221),
  so boundary neighbours keep their own indices.
"""
from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction

import numpy as np


def star_points(order: int = 4, dim: int = 3) -> np.ndarray:
    pts = []
    got_center = False
    for d in range(dim):
        for j in range(-order, order + 1):
            if j == 0:
                if got_center:
                    continue
                got_center = True
            p = [0, 0, 0]
            p[d] = j
            pts.append(p)
    return np.asarray(pts, dtype=np.int32)


def central_d2_coefficients(order: int) -> list[Fraction]:
    """Exact 1-D second-derivative coefficients c_0..c_order on offsets 0..order.

    Solved in rationals from sum_k c_k k^(2m) = delta_{m,1}*(2m)!/2 ... (symmetric
    stencil), i.e. Taylor matching of even moments."""
    n = order
    # unknowns c_0..c_n ; equations for even powers 0,2,..,2n: sum_j c_j j^p (both sides)
    A = [[Fraction(0)] * (n + 1) for _ in range(n + 1)]
    b = [Fraction(0)] * (n + 1)
    for m in range(n + 1):
        p = 2 * m
        for j in range(n + 1):
            mult = 1 if j == 0 else 2
            A[m][j] = Fraction(mult * (j ** p if p > 0 else 1))
        b[m] = Fraction(2) if p == 2 else Fraction(0)
    # Gaussian elimination in rationals
    for i in range(n + 1):
        piv = next(r for r in range(i, n + 1) if A[r][i] != 0)
        A[i], A[piv] = A[piv], A[i]
        b[i], b[piv] = b[piv], b[i]
        for r in range(n + 1):
            if r != i and A[r][i] != 0:
                f = A[r][i] / A[i][i]
                A[r] = [a - f * c for a, c in zip(A[r], A[i])]
                b[r] -= f * b[i]
    return [b[i] / A[i][i] for i in range(n + 1)]


def weights_closed_form(order: int, spacing: float, points: np.ndarray | None = None) -> np.ndarray:
    """Laplacian weights in stencil point order from the exact 1-D coefficients."""
    pts = star_points(order) if points is None else points
    c = central_d2_coefficients(order)
    w = np.empty(pts.shape[0])
    for i, p in enumerate(pts):
        k = int(np.max(np.abs(p)))
        if k == 0:
            w[i] = float(3 * c[0]) / spacing ** 2
        else:
            w[i] = float(c[k]) / spacing ** 2
    return w


def weights_polyfit(order: int, spacing: float) -> np.ndarray:
    """Restatement of derivatives_make_discretization for the star Laplacian
    (src/grid/derivatives.F90:664-718): square polynomial system solved by LU."""
    pts = star_points(order)
    npoly = pts.shape[0]
    pol = np.zeros((npoly, 3), dtype=int)   # pol(:, 1) = 0 constant
    n = 1
    for d in range(3):
        for j in range(1, 2 * order + 1):
            pol[n, d] = j
            n += 1
    x = pts.astype(np.float64) * spacing
    mat = np.ones((npoly, npoly))
    for j in range(1, npoly):
        mat[j, :] = np.prod(x ** pol[j][None, :], axis=1)
    rhs = np.zeros(npoly)
    for j in range(npoly):
        if sorted(pol[j]) == [0, 0, 2]:
            rhs[j] = 2.0
    return np.linalg.solve(mat, rhs)


@dataclass
class NLOperator:
    """Host-side image of nl_operator_t after nl_operator_build."""
    n: int                  # stencil size
    w: np.ndarray           # (n,) constant weights
    nri: int
    ri: np.ndarray          # (nri, n) int32 relative offsets, row-major: ri[l*n + j]
    rimap: np.ndarray       # (np,) int32 run of each point
    rimap_inv: np.ndarray   # (nri+1,) int32 run starts, last = np
    np: int
    np_part: int
    center: int             # index of the zero offset in the stencil


def neighbours(mesh, points: np.ndarray) -> np.ndarray:
    """(np, n) absolute neighbour index of every inner point."""
    c = mesh.coords[: mesh.np].astype(np.int64)
    nb = np.empty((mesh.np, points.shape[0]), dtype=np.int64)
    for j, off in enumerate(points):
        nb[:, j] = mesh.index_of(c + off)
    if np.any(nb < 0):
        raise ValueError("stencil reaches a point outside inner+boundary set")
    return nb


def nl_operator_build(mesh, order: int = 4, weights: np.ndarray | None = None) -> NLOperator:
    pts = star_points(order)
    n = pts.shape[0]
    w = weights_closed_form(order, mesh.spacing, pts) if weights is None else np.asarray(weights)
    nb = neighbours(mesh, pts)
    idx = np.arange(mesh.np, dtype=np.int64)
    rel = nb - idx[:, None]
    has_bnd = np.any(nb >= mesh.np, axis=1)
    new_run = np.empty(mesh.np, dtype=bool)
    new_run[0] = True
    if mesh.np > 1:
        change = np.any(rel[1:] != rel[:-1], axis=1)
        force = has_bnd[1:] & ~has_bnd[:-1]
        new_run[1:] = change | force
    rimap = (np.cumsum(new_run) - 1).astype(np.int32)
    starts = np.nonzero(new_run)[0]
    ri = rel[starts].astype(np.int32)
    rimap_inv = np.concatenate([starts, [mesh.np]]).astype(np.int32)
    center = int(np.nonzero(np.all(pts == 0, axis=1))[0][0])
    return NLOperator(n=n, w=w.astype(np.float64), nri=int(starts.size), ri=np.ascontiguousarray(ri),
                      rimap=rimap, rimap_inv=rimap_inv, np=mesh.np, np_part=mesh.np_part, center=center)
