"""Periodic cells and Bloch phases: the inputs Octopus builds for one k-point of a periodic
system, in the layouts the device path takes.

* cell: the n^3 inner points of a cubic cell (edge n*h) with the 4-deep boundary shell the
  order-4 star reaches; boundary point b takes the value of its periodic image
  (mesh_periodic_point, the periodic branch of boundaries_set,
  src/grid/boundaries_inc.F90:366-743): ``bnd_src[b - np]`` = image row (0-based);
* phase(p) = exp(-i k.x_p) over all np_part points and phase_corr(b) =
  exp(i k.x_image(b)) phase(b) over the boundary points
  (src/hamiltonian/hamiltonian_elec.F90:1217-1232);
* projector spheres may cross the cell boundary: every lattice point within rc of the
  atom is mapped to its image inside the cell, and its projector phase is
  exp(-i k.(x_actual - x_image)) (projector_init_phases, src/hamiltonian/projector.F90:
  195-237).  Sphere points keep increasing image index, as for finite systems.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import rng
from .mesh import cube_mesh
from .projectors import ProjectorSet, ylm_real
from .stencil import NLOperator, nl_operator_build, star_points


@dataclass
class PeriodicSystem:
    mesh: object
    op: NLOperator
    n: int
    v: np.ndarray
    proj: ProjectorSet | None
    pphase_of: object       # k -> projector phases (total_points complex)
    bnd_src: np.ndarray     # (np_part - np,) int32, image row of each boundary row
    seed: int

    @property
    def np(self) -> int:
        return self.mesh.np

    @property
    def np_part(self) -> int:
        return self.mesh.np_part

    @property
    def cell(self) -> float:
        return self.n * self.mesh.spacing

    def image(self, coords: np.ndarray) -> np.ndarray:
        """0-based inner row of the periodic image of lattice coords (m, 3)."""
        return self.mesh.index_of(np.mod(coords, self.n))

    def phases(self, k) -> tuple[np.ndarray, np.ndarray]:
        """phase (np_part complex) and phase_corr (np_part - np complex) of k-point k."""
        k = np.asarray(k, np.float64)
        x = self.mesh.x
        phase = np.exp(-1j * (x @ k))
        xi = self.mesh.x[self.bnd_src]
        corr = np.exp(1j * (xi @ k)) * phase[self.np:]
        return phase, corr


def periodic_cell(n: int, spacing: float, states_block_size: int = 32):
    mesh = cube_mesh(n, spacing, star_points(4), states_block_size=states_block_size)
    bc = mesh.coords[mesh.np:].astype(np.int64)
    src = mesh.index_of(np.mod(bc, n))
    assert np.all((src >= 0) & (src < mesh.np)), "every boundary point has an inner image"
    return mesh, src.astype(np.int32)


def periodic_system(n: int = 24, spacing: float = 0.3, natoms: int = 3, nwells: int = 4, rc: float = 1.5,
                    seed: int = 4242, states_block_size: int = 32, edge_atom: bool = True) -> PeriodicSystem:
    """A cubic cell with Gaussian wells and KB atoms (l <= 1), one atom placed across the
    cell boundary (edge_atom) so that its sphere wraps around."""
    mesh, bnd_src = periodic_cell(n, spacing, states_block_size)
    op = nl_operator_build(mesh, 4)
    L = n * spacing
    half = L / 2.0
    xin = mesh.x[: mesh.np]
    # wells at seeded sites, summed over the 27 nearest images (smooth and periodic)
    sites = (rng.uniform(seed, 1, 3 * nwells, -1.0, 1.0).reshape(nwells, 3)) * (half - 1.0)
    v = np.zeros(mesh.np)
    for c in sites:
        for sh in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T:
            d = xin - (c + sh * L)
            v += -2.0 * np.exp(-np.einsum("ij,ij->i", d, d) / 0.8 ** 2)
    centers = (rng.uniform(seed, 2, 3 * natoms, -1.0, 1.0).reshape(natoms, 3)) * (half - rc - 0.5)
    if edge_atom and natoms > 0:
        centers[0] = [half - 0.4 * rc, -half + 0.3 * rc, 0.1]  # its sphere crosses two faces
    alphas = rng.uniform(seed, 3, 2 * natoms, 0.6, 1.2).reshape(natoms, 2)
    energies = rng.uniform(seed, 4, 2 * natoms, -1.0, 1.0).reshape(natoms, 2)
    sysobj = PeriodicSystem(mesh=mesh, op=op, n=n, v=v, proj=None, pphase_of=None, bnd_src=bnd_src, seed=seed)
    if natoms > 0:
        proj, diffs = periodic_kb_projectors(sysobj, centers, rc, alphas, energies)
        sysobj.proj = proj
        sysobj.pphase_of = lambda k: np.exp(-1j * (diffs @ np.asarray(k, np.float64)))
    return sysobj


def periodic_kb_projectors(s: PeriodicSystem, centers, rc: float, alphas, energies, lmax: int = 1):
    """KB projector matrices of atoms in a periodic cell (as projectors.kb_projector_set,
    with sphere points mapped to their images).  Returns the set and, per sphere point in
    matrix order, x_actual - x_image (the projector phases are exp(-i k . that))."""
    m = s.mesh
    h = m.spacing
    offs, mats, maps, scals, diffs = [], [], [], [], []
    moff = poff = soff = 0
    for a, c in enumerate(np.asarray(centers, np.float64)):
        lo = np.floor((c - rc) / h + m.origin_shift).astype(np.int64) - 1
        hi = np.ceil((c + rc) / h + m.origin_shift).astype(np.int64) + 1
        ax = [np.arange(lo[d], hi[d] + 1) for d in range(3)]
        g = np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1).reshape(-1, 3)
        xa = (g.astype(np.float64) - m.origin_shift) * h
        rel = xa - c
        keep = np.einsum("ij,ij->i", rel, rel) <= rc * rc
        g, xa, rel = g[keep], xa[keep], rel[keep]
        img = s.image(g)
        if len(np.unique(img)) != len(img):
            raise ValueError("a projector sphere overlaps its own periodic image")
        order = np.argsort(img, kind="stable")
        g, xa, rel, img = g[order], xa[order], rel[order], img[order]
        r = np.sqrt(np.einsum("ij,ij->i", rel, rel))
        cols, sc = [], []
        for l in range(lmax + 1):
            f = r ** l * np.exp(-(r * r) / alphas[a, l] ** 2) * (1.0 - (r / rc) ** 2) ** 2
            for mm in range(-l, l + 1):
                cols.append(f * ylm_real(l, mm, rel))
                sc.append(energies[a, l] * m.vol_pp)
        P = np.stack(cols, axis=1)
        npt, npj = P.shape
        offs.append([npt, npj, moff, poff, soff, -1])
        mats.append(P.T.ravel())
        maps.append((img + 1).astype(np.int32))
        scals.append(np.asarray(sc))
        diffs.append(xa - m.x[img])
        moff += npt * npj
        poff += npt
        soff += npj
    proj = ProjectorSet(nmat=len(offs), offsets=np.asarray(offs, dtype=np.int32).reshape(-1, 6),
                        matrices=np.concatenate(mats), maps=np.concatenate(maps),
                        scals=np.concatenate(scals), centers=np.asarray(centers, np.float64).reshape(-1, 3))
    return proj, np.concatenate(diffs)
