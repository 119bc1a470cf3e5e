"""Small seeded systems and comparison helpers shared by the tests."""
from __future__ import annotations

import numpy as np

from octhip import workloads
from octhip.mesh import sphere_mesh, cube_mesh
from octhip.stencil import nl_operator_build, star_points, neighbours


def small_sphere(radius=3.0, spacing=0.3, natoms=3, nwells=4, seed=12345, sbs=32, rc=1.2):
    return workloads.hamiltonian_system(radius=radius, spacing=spacing, seed=seed, nwells=nwells,
                                        natoms=natoms, states_block_size=sbs, rc=rc, name="small")


def max_rel_err(x: np.ndarray, y: np.ndarray) -> float:
    """max |x - y| / max |y| per state column (columns = lanes or states), worst state."""
    x = np.asarray(x)
    y = np.asarray(y)
    num = np.max(np.abs(x - y), axis=0)
    den = np.max(np.abs(y), axis=0)
    den = np.where(den == 0, 1.0, den)
    return float(np.max(num / den))


def per_state_rel_err(x: np.ndarray, y: np.ndarray) -> float:
    """max over states of max_points |x - y| / max_points |y| for (nst, np) arrays."""
    num = np.max(np.abs(x - y), axis=1)
    den = np.max(np.abs(y), axis=1)
    den = np.where(den == 0, 1.0, den)
    return float(np.max(num / den))


def brute_laplacian(mesh, w, f_np_part: np.ndarray, factor=1.0) -> np.ndarray:
    """Direct 3-D star stencil through the lattice lookup, independent of ri."""
    nb = neighbours(mesh, star_points(4))
    return factor * np.einsum("pj,pjk->pk", np.broadcast_to(w, nb.shape), f_np_part[nb])


def random_packed(np_part: int, lanes: int, seed: int, np_inner: int | None = None) -> np.ndarray:
    r = np.random.default_rng(seed)
    a = r.standard_normal((np_part, lanes))
    if np_inner is not None:
        a[np_inner:] = 0.0
    return a


def synthetic_projectors(mesh, centers, rc: float, nprojs, seed: int):
    """Flat build_accel arrays (ProjectorSet) with arbitrary projector counts per matrix:
    matrix a covers the inner points within rc of centers[a] and holds nprojs[a] seeded
    random columns; scal ~ U[-1, 1] * vol_pp.  Spheres may overlap."""
    from octhip.projectors import ProjectorSet, sphere_points
    r = np.random.default_rng(seed)
    offs, mats, maps, scals = [], [], [], []
    moff = poff = soff = 0
    for c, npj in zip(centers, nprojs):
        idx = sphere_points(mesh, np.asarray(c, np.float64), rc)
        npt = len(idx)
        P = r.standard_normal((npt, npj))
        offs.append([npt, npj, moff, poff, soff, -1])
        mats.append(P.T.ravel())
        maps.append((idx + 1).astype(np.int32))
        scals.append(r.uniform(-1.0, 1.0, npj) * mesh.vol_pp)
        moff += npt * npj
        poff += npt
        soff += npj
    return ProjectorSet(nmat=len(nprojs), offsets=np.asarray(offs, dtype=np.int32).reshape(-1, 6),
                        matrices=np.concatenate(mats), maps=np.concatenate(maps),
                        scals=np.concatenate(scals), centers=np.asarray(centers, np.float64).reshape(-1, 3))
