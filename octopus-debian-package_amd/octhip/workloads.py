"""Seeded synthetic workloads for the H-apply path.

Workloads (all FP64, order-4 star Laplacian, zero Dirichlet boundary):

* ``laplacian_cube``  -- 64^3 cube, h=0.3, 8 real functions (pack 8), each a sum of 4
  Gaussians a*exp(-|r-c|^2/w^2), c ~ U(box), w ~ U[0.5,2], a ~ N(0,1).
* ``hamiltonian_sphere`` -- sphere R=12, h=0.25, complex states Re,Im ~ N(0,1) *
  exp(-r^2/4^2); v = sum of 16 wells -2 exp(-|r-R_i|^2/0.8^2), R_i ~ U(ball 9);
  16 atoms x 4 KB projectors (l=0,1) on rc=3 spheres, alpha ~ U[0.6,1.2],
  e ~ U[-1,1], scal = e h^3.
* ``taylor_sphere`` -- the same H with orthonormalised states (Gram-Schmidt, FP64).
* ``uneven_systems`` -- 8 spheres R = 6,8,...,14 with their own wells (4-24) and
  projector sets and 8..256 states.
* ``adjoint_sphere`` -- sphere R=10, h=0.3, 32 states, dipole coupling x * eps(t).

Streams of the counter-based generator are fixed per quantity (see `_S`).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import rng
from .mesh import cube_mesh, sphere_mesh
from .projectors import ProjectorSet, kb_projector_set
from .stencil import NLOperator, nl_operator_build, star_points

_S = dict(wells=1, atoms=2, alpha=3, energy=4, state_re=5, state_im=6,
          gauss_c=7, gauss_w=8, gauss_a=9, nwells=10, field=11, targets=12)


@dataclass
class System:
    mesh: object
    op: NLOperator
    v: np.ndarray | None            # (np,) local potential
    proj: ProjectorSet | None
    seed: int
    name: str

    @property
    def np(self) -> int:
        return self.mesh.np

    @property
    def np_part(self) -> int:
        return self.mesh.np_part


def gaussian_wells(mesh, centers: np.ndarray, depth: float = 2.0, width: float = 0.8) -> np.ndarray:
    x = mesh.x[: mesh.np]
    v = np.zeros(mesh.np)
    for c in centers:
        d = x - c
        v += -depth * np.exp(-np.einsum("ij,ij->i", d, d) / width ** 2)
    return v


def hamiltonian_system(radius=12.0, spacing=0.25, seed=None, nwells=16, natoms=16,
                       states_block_size=32, rc=3.0, name="hamiltonian_sphere", block=None) -> System:
    seed = rng.SEED_BASE + 1 if seed is None else seed
    pts = star_points(4)
    mesh = sphere_mesh(radius, spacing, pts, states_block_size=states_block_size, block=block)
    op = nl_operator_build(mesh, 4)
    inner_r = max(radius - rc, 0.0)
    wells = rng.in_ball(seed, _S["wells"], nwells, inner_r)
    v = gaussian_wells(mesh, wells)
    if natoms > 0:
        atoms = rng.in_ball(seed, _S["atoms"], natoms, inner_r)
        alphas = rng.uniform(seed, _S["alpha"], natoms * 2, 0.6, 1.2).reshape(natoms, 2)
        energies = rng.uniform(seed, _S["energy"], natoms * 2, -1.0, 1.0).reshape(natoms, 2)
        proj = kb_projector_set(mesh, atoms, rc, alphas, energies, lmax=1)
    else:
        proj = None
    return System(mesh=mesh, op=op, v=v, proj=proj, seed=seed, name=name)


def laplacian_system(npts=64, spacing=0.3, seed=None, states_block_size=8) -> System:
    seed = rng.SEED_BASE + 0 if seed is None else seed
    mesh = cube_mesh(npts, spacing, star_points(4), states_block_size=states_block_size)
    op = nl_operator_build(mesh, 4)
    return System(mesh=mesh, op=op, v=None, proj=None, seed=seed, name="laplacian_cube")


def gaussian_functions(system: System, nfun: int = 8, ngauss: int = 4) -> np.ndarray:
    """(nfun, np) real functions on inner points (boundary values are zero)."""
    m = system.mesh
    half = (np.max(m.coords[: m.np], axis=0) - np.min(m.coords[: m.np], axis=0)) * m.spacing / 2.0
    c = rng.uniform(system.seed, _S["gauss_c"], nfun * ngauss * 3, -1.0, 1.0).reshape(nfun, ngauss, 3) * half
    w = rng.uniform(system.seed, _S["gauss_w"], nfun * ngauss, 0.5, 2.0).reshape(nfun, ngauss)
    a = rng.normal(system.seed, _S["gauss_a"], nfun * ngauss).reshape(nfun, ngauss)
    x = m.x[: m.np]
    out = np.zeros((nfun, m.np))
    for i in range(nfun):
        for g in range(ngauss):
            d = x - c[i, g]
            out[i] += a[i, g] * np.exp(-np.einsum("ij,ij->i", d, d) / w[i, g] ** 2)
    return out


def gaussian_envelope_states(system: System, nst: int, width: float = 4.0, first: int = 0) -> np.ndarray:
    """(nst, np) complex states: (N(0,1) + i N(0,1)) * exp(-r^2/width^2)."""
    m = system.mesh
    x = m.x[: m.np]
    env = np.exp(-np.einsum("ij,ij->i", x, x) / width ** 2)
    out = np.empty((nst, m.np), dtype=np.complex128)
    for i in range(nst):
        re = rng.normal(system.seed, _S["state_re"] * 100000 + first + i, m.np)
        im = rng.normal(system.seed, _S["state_im"] * 100000 + first + i, m.np)
        out[i] = (re + 1j * im) * env
    return out


def orthonormalise(states: np.ndarray, vol_pp: float) -> np.ndarray:
    """Gram-Schmidt in the mesh inner product <a|b> = h^3 sum conj(a) b (via QR,
    R's diagonal made positive so the result equals classical Gram-Schmidt)."""
    A = states.T * np.sqrt(vol_pp)
    q, r = np.linalg.qr(A)
    d = np.sign(np.real(np.diag(r)))
    d[d == 0] = 1
    q = q * d[None, :]
    return (q / np.sqrt(vol_pp)).T.copy()


def pack(states: np.ndarray, np_part: int, pack_size: int | None = None) -> np.ndarray:
    """batch_t packed layout (src/grid/batch.F90:1017-1045): rows = points (np_part),
    columns = pad_pow2(nst) lanes, complex as interleaved (re, im) doubles."""
    nst, npt = states.shape
    ps = pack_size or (1 << max(0, int(np.ceil(np.log2(max(nst, 1))))))
    if np.iscomplexobj(states):
        out = np.zeros((np_part, ps), dtype=np.complex128)
        out[:npt, :nst] = states.T
        return out.view(np.float64)           # (np_part, 2*ps)
    out = np.zeros((np_part, ps), dtype=np.float64)
    out[:npt, :nst] = states.T
    return out


def unpack(packed: np.ndarray, nst: int, npt: int, is_complex: bool) -> np.ndarray:
    if is_complex:
        z = np.ascontiguousarray(packed).view(np.complex128)
        return z[:npt, :nst].T.copy()
    return packed[:npt, :nst].T.copy()


def uneven_systems(seed=None, spacing=0.25):
    """Eight independent spheres of increasing size and state count."""
    seed = rng.SEED_BASE + 3 if seed is None else seed
    radii = [6.0, 8.0, 9.0, 10.0, 11.0, 12.0, 13.0, 14.0]
    nsts = [8, 16, 24, 48, 64, 96, 128, 256]
    nwell = (4 + np.floor(rng.uniform(seed, _S["nwells"], 8) * 21)).astype(int)
    out = []
    for i, (R, n) in enumerate(zip(radii, nsts)):
        s = hamiltonian_system(radius=R, spacing=spacing, seed=seed * 16 + i, nwells=int(nwell[i]),
                               natoms=16, name=f"uneven[{i}]")
        out.append((s, n))
    return out


def dipole_field(seed: int, nsin: int = 5):
    """eps(t) = sum_i A_i sin(w_i t + phi_i), w ~ U[0.05,0.5], |A| <= 0.01."""
    u = rng.uniform(seed, _S["field"], 3 * nsin)
    w = 0.05 + 0.45 * u[:nsin]
    A = 0.01 * (2.0 * u[nsin:2 * nsin] - 1.0)
    phi = 2.0 * np.pi * u[2 * nsin:]
    return w, A, phi


def eps_of_t(t, w, A, phi):
    t = np.asarray(t, dtype=np.float64)
    return np.sum(A[None, :] * np.sin(w[None, :] * t[..., None] + phi[None, :]), axis=-1)


def adjoint_system(seed=None) -> System:
    """Sphere R=10, h=0.3 with the sphere-R12 Hamiltonian recipe (16 wells, 16 KB atoms)."""
    seed = rng.SEED_BASE + 4 if seed is None else seed
    return hamiltonian_system(radius=10.0, spacing=0.3, seed=seed, name="adjoint_sphere")


def adjoint_problem(system: System, nst: int = 32, nsteps: int = 100, dt: float = 0.02):
    """Initial states, targets (both orthonormalised, seeded) and eps at the step
    midpoints t_s + dt/2 (exponential-midpoint convention)."""
    vol = system.mesh.vol_pp
    psi0 = orthonormalise(gaussian_envelope_states(system, nst), vol)
    m = system.mesh
    x = m.x[: m.np]
    env = np.exp(-np.einsum("ij,ij->i", x, x) / 4.0 ** 2)
    tg = np.empty((nst, m.np), dtype=np.complex128)
    for i in range(nst):
        re = rng.normal(system.seed, _S["targets"] * 100000 + 2 * i, m.np)
        im = rng.normal(system.seed, _S["targets"] * 100000 + 2 * i + 1, m.np)
        tg[i] = (re + 1j * im) * env
    phi = orthonormalise(tg, vol)
    w, A, ph = dipole_field(system.seed)
    eps = eps_of_t((np.arange(nsteps) + 0.5) * dt, w, A, ph)
    return psi0, phi, eps
