"""Real-space mesh: inner points, boundary points and the blocked point order.

Restates the point-set rules of Octopus's mesh construction for finite (non-periodic)
boxes, so that the packed state layout handed to the device path is the one Octopus
itself would hand over:

* extent per axis: largest j with j*h <= bound + 1e-12
  (src/grid/mesh_init.F90:96-118, BOX_BOUNDARY_DELTA at src/boxes/box.F90:38);
* inner points: lattice points inside the box; sphere rule |r|^2 <= (R + 1e-12)^2
  (src/boxes/box_sphere.F90:100), scanned in cubic (x-fastest) order
  (src/grid/mesh_init.F90:262-279, src/grid/index.F90:313-330);
* boundary points: stencil neighbours of inner points that are neither inner nor
  inside the box, deduplicated (src/grid/mesh_init.F90:298-328);
* both sets sorted by the blocked index (src/grid/mesh_init.F90:786-1055,
  get_blocked_index at src/grid/index.F90:358-394) with block sizes derived from
  StatesBlockSize (src/grid/mesh_init.F90:883-891); inner points are 0..np-1 and
  boundary points np..np_part-1.

The even-sized cube (64^3 at half-integer positions) cannot be produced by Octopus's
symmetric mesh; it is built with the same rules from an explicit lattice range.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BOX_BOUNDARY_DELTA = 1e-12


def default_block_size(states_block_size: int) -> tuple[int, int, int]:
    """MeshBlockSize default (src/grid/mesh_init.F90:883-891)."""
    sbs = abs(int(states_block_size))
    if sbs < 16:
        return (80 * 4 // sbs, 4, 10)
    return (max(4 * 16 // sbs, 1), 15, 15)


def blocked_index(pts: np.ndarray, bsize, nblocks) -> np.ndarray:
    """get_blocked_index with x fastest (src/grid/index.F90:358-394); pts >= 0."""
    pts = pts.astype(np.int64)
    jl = np.zeros(pts.shape[0], dtype=np.int64)
    jb = np.zeros(pts.shape[0], dtype=np.int64)
    sl, sb = 1, 1
    for d in range(3):
        jl += (pts[:, d] % bsize[d]) * sl
        sl *= bsize[d]
        jb += (pts[:, d] // bsize[d]) * sb
        sb *= nblocks[d]
    return jl + jb * int(np.prod(bsize))


@dataclass
class Mesh:
    spacing: float
    coords: np.ndarray          # (np_part, 3) int32 lattice coordinates, inner then boundary
    np: int
    np_part: int
    origin_shift: np.ndarray    # position = (coords - origin_shift) * spacing
    lo: np.ndarray              # enlarged lattice range [lo, hi]
    hi: np.ndarray
    block: tuple
    lookup: np.ndarray = field(repr=False)  # enlarged cube -> point index, -1 if none
    kind: str = ""

    @property
    def x(self) -> np.ndarray:
        """Cartesian positions (np_part, 3)."""
        return (self.coords.astype(np.float64) - self.origin_shift) * self.spacing

    @property
    def vol_pp(self) -> float:
        return self.spacing ** 3

    def index_of(self, c: np.ndarray) -> np.ndarray:
        """Point index (0-based) of lattice coords c (n, 3); -1 outside the point set."""
        rel = c - self.lo
        ok = np.all((rel >= 0) & (rel < self.lookup.shape), axis=1)
        out = np.full(c.shape[0], -1, dtype=np.int64)
        r = rel[ok]
        out[ok] = self.lookup[r[:, 0], r[:, 1], r[:, 2]]
        return out


def _assemble(inside_fn, lo_in, hi_in, shift, spacing, stencil_points, block, kind) -> Mesh:
    enl = int(np.max(np.abs(stencil_points)))
    lo = np.asarray(lo_in, dtype=np.int64) - enl
    hi = np.asarray(hi_in, dtype=np.int64) + enl
    sizes = hi - lo + 1
    # candidate inner points, cubic order (x fastest)
    ax = [np.arange(lo_in[d], hi_in[d] + 1) for d in range(3)]
    zz, yy, xx = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    cand = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.int64)
    inner = cand[inside_fn(cand)]
    # boundary points: neighbours not inner and not in the box
    inner_mask = np.zeros(tuple(sizes), dtype=bool)
    r = inner - lo
    inner_mask[r[:, 0], r[:, 1], r[:, 2]] = True
    bnd_mask = np.zeros_like(inner_mask)
    for off in stencil_points:
        if not np.any(off):
            continue
        nb = inner + off
        rr = nb - lo
        not_inner = ~inner_mask[rr[:, 0], rr[:, 1], rr[:, 2]]
        nbc = nb[not_inner]
        outside = ~inside_fn(nbc)
        q = nbc[outside] - lo
        bnd_mask[q[:, 0], q[:, 1], q[:, 2]] = True
    bz, by, bx = np.nonzero(np.transpose(bnd_mask, (2, 1, 0)))
    bnd = np.stack([bx, by, bz], axis=1).astype(np.int64) + lo
    # blocked ordering; coordinates shifted to be non-negative like idx%offset
    nblocks = [int(sizes[d]) // block[d] + 1 for d in range(3)]
    off0 = -lo
    oi = np.argsort(blocked_index(inner + off0, block, nblocks), kind="stable")
    ob = np.argsort(blocked_index(bnd + off0, block, nblocks), kind="stable")
    coords = np.concatenate([inner[oi], bnd[ob]]).astype(np.int32)
    npt, npp = inner.shape[0], coords.shape[0]
    lookup = np.full(tuple(sizes), -1, dtype=np.int64)
    rc = coords.astype(np.int64) - lo
    lookup[rc[:, 0], rc[:, 1], rc[:, 2]] = np.arange(npp)
    return Mesh(spacing=spacing, coords=coords, np=npt, np_part=npp,
                origin_shift=np.asarray(shift, dtype=np.float64), lo=lo, hi=hi,
                block=tuple(block), lookup=lookup, kind=kind)


def sphere_mesh(radius: float, spacing: float, stencil_points: np.ndarray,
                states_block_size: int = 32, block=None) -> Mesh:
    """Spherical box centred at the origin."""
    n = 0
    while (n + 1) * spacing <= radius + BOX_BOUNDARY_DELTA:
        n += 1
    rr = (radius + BOX_BOUNDARY_DELTA) ** 2

    def inside(c):
        x = c.astype(np.float64) * spacing
        return np.einsum("ij,ij->i", x, x) <= rr

    block = block or default_block_size(states_block_size)
    return _assemble(inside, [-n] * 3, [n] * 3, [0.0] * 3, spacing, stencil_points, block,
                     kind=f"sphere(R={radius},h={spacing})")


def cube_mesh(npts: int, spacing: float, stencil_points: np.ndarray,
              states_block_size: int = 8, block=None) -> Mesh:
    """npts^3 cube at positions (i - (npts-1)/2) * h, zero Dirichlet outside."""
    def inside(c):
        return np.all((c >= 0) & (c < npts), axis=1)

    block = block or default_block_size(states_block_size)
    return _assemble(inside, [0] * 3, [npts - 1] * 3, [(npts - 1) / 2.0] * 3, spacing,
                     stencil_points, block, kind=f"cube({npts}^3,h={spacing})")
