"""Kleinman-Bylander projector matrices in the device layout Octopus builds.

* sphere points: inner mesh points with |x - R_a|^2 <= rc^2, in increasing mesh index
  (src/grid/submesh.F90:222-265, sorted at src/grid/submesh.F90:369-370); `map` is
  1-based like pmat%map;
* columns: l = 0..lmax (skipping lloc), m = -l..l, component ic; entry
  p = f_l(r) * Y_lm(r_hat) (src/hamiltonian/hamiltonian_elec_base.F90:604-634,
  src/hamiltonian/kb_projector.F90:61-103);
* real spherical harmonics: ylmr_real (src/math/math.F90:244-310): l=0: sqrt(1/4pi);
  l=1: m=-1: -c*y/r, m=0: c*z/r, m=1: -c*x/r with c = sqrt(3/4pi); 0 at r < 1e-15 for l>0;
* scal(j) = e_(l,ic) * vol_pp (src/hamiltonian/hamiltonian_elec_base.F90:628);
* flattened layout of build_accel (src/hamiltonian/hamiltonian_elec_base.F90:755-909):
  offsets(6, nmat) = POINTS, PROJS, MATRIX, MAP, SCAL, MIX(-1 = none); matrices are the
  column-major (npoints x nprojs) blocks back to back; maps and scals concatenated.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

OFFSET_SIZE = 6
POINTS, PROJS, MATRIX, MAP, SCAL, MIX = range(6)


def ylm_real(l: int, m: int, rel: np.ndarray) -> np.ndarray:
    r = np.sqrt(np.einsum("ij,ij->i", rel, rel))
    if l == 0:
        return np.full(rel.shape[0], np.sqrt(1.0 / (4.0 * np.pi)))
    if l != 1:
        raise NotImplementedError("only l <= 1 projectors are generated")
    c = np.sqrt(3.0 / (4.0 * np.pi))
    out = np.zeros(rel.shape[0])
    ok = r >= 1e-15
    rr = r[ok]
    if m == -1:
        out[ok] = (-c) * (rel[ok, 1] / rr)
    elif m == 0:
        out[ok] = c * (rel[ok, 2] / rr)
    else:
        out[ok] = (-c) * (rel[ok, 0] / rr)
    return out


@dataclass
class ProjectorSet:
    nmat: int
    offsets: np.ndarray     # (nmat, 6) int32
    matrices: np.ndarray    # concatenated column-major blocks, float64
    maps: np.ndarray        # concatenated 1-based mesh indices, int32
    scals: np.ndarray       # concatenated, float64
    centers: np.ndarray     # (nmat, 3)

    @property
    def full_projection_size(self) -> int:
        return int(self.offsets[:, PROJS].sum()) if self.nmat else 0

    def matrix(self, a: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        o = self.offsets[a]
        npt, npj = int(o[POINTS]), int(o[PROJS])
        P = self.matrices[o[MATRIX]: o[MATRIX] + npt * npj].reshape(npj, npt).T
        return P, self.maps[o[MAP]: o[MAP] + npt], self.scals[o[SCAL]: o[SCAL] + npj]

    def self_overlap(self) -> bool:
        return False


def sphere_points(mesh, center: np.ndarray, rc: float) -> np.ndarray:
    """0-based mesh indices of inner points within rc of center, increasing."""
    h = mesh.spacing
    lo = np.floor((center - rc) / h + mesh.origin_shift).astype(np.int64) - 1
    hi = np.ceil((center + rc) / h + mesh.origin_shift).astype(np.int64) + 1
    ax = [np.arange(lo[d], hi[d] + 1) for d in range(3)]
    g = np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1).reshape(-1, 3)
    idx = mesh.index_of(g)
    idx = idx[(idx >= 0) & (idx < mesh.np)]
    x = (mesh.coords[idx].astype(np.float64) - mesh.origin_shift) * h - center
    keep = np.einsum("ij,ij->i", x, x) <= rc * rc
    return np.sort(idx[keep])


def kb_projector_set(mesh, centers: np.ndarray, rc: float, alphas: np.ndarray,
                     energies: np.ndarray, lmax: int = 1) -> ProjectorSet:
    """Synthetic KB projectors: f_l(r) = r^l exp(-r^2/alpha^2) (1 - (r/rc)^2)^2 times Y_lm.

    alphas, energies: (natoms, lmax+1) per (atom, l); one component (n_c = 1)."""
    offs, mats, maps, scals = [], [], [], []
    moff = 0
    poff = 0
    soff = 0
    for a, c in enumerate(centers):
        idx = sphere_points(mesh, c, rc)
        rel = (mesh.coords[idx].astype(np.float64) - mesh.origin_shift) * mesh.spacing - c
        r = np.sqrt(np.einsum("ij,ij->i", rel, rel))
        cols, sc = [], []
        for l in range(lmax + 1):
            f = r ** l * np.exp(-(r * r) / alphas[a, l] ** 2) * (1.0 - (r / rc) ** 2) ** 2
            for m in range(-l, l + 1):
                cols.append(f * ylm_real(l, m, rel))
                sc.append(energies[a, l] * mesh.vol_pp)
        P = np.stack(cols, axis=1)  # (npoints, nprojs)
        npt, npj = P.shape
        offs.append([npt, npj, moff, poff, soff, -1])
        mats.append(P.T.ravel())    # column-major
        maps.append((idx + 1).astype(np.int32))
        scals.append(np.asarray(sc))
        moff += npt * npj
        poff += npt
        soff += npj
    return ProjectorSet(nmat=len(centers), offsets=np.asarray(offs, dtype=np.int32).reshape(-1, 6),
                        matrices=np.concatenate(mats) if mats else np.zeros(0),
                        maps=np.concatenate(maps) if maps else np.zeros(0, np.int32),
                        scals=np.concatenate(scals) if scals else np.zeros(0),
                        centers=np.asarray(centers, dtype=np.float64).reshape(-1, 3))
