"""Synthetic pressure / momentum systems of the benchmark configurations.

Meshes come from libcfdgpu's box generator (restating the reference's make_box_mesh and
Mesh::compute_geometry); coefficients follow the reference:
  * pressure Laplacian: off-diagonals -|S_f| / |C_N - C_P| (tests/test_amg.cpp:52-61), diagonal
    the negative sum of the off-diagonals, zero-gradient walls, the pressure pinned at cell 0
    by set_reference (src/fv/src/ldu_matrix.cpp:105-109);
  * variable coefficient: face value kappa_P kappa_N / (w kappa_N + (1 - w) kappa_P)
    (src/fv/src/energy.cpp:159-163) with w the face weight (src/fv/src/fv_geometry.cpp:52-58);
  * momentum: the same off-diagonals with diagonal 1.1 sum|off| + V / dt.
Random numbers are mt19937_64 draws (u >> 11) 2^-53, identical on every platform.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .krylov import CsrView, LduPattern

SEED_BASE = 20261202


@dataclass
class BoxMesh:
    nx: int
    ny: int
    nz: int
    lo: tuple
    hi: tuple
    jitter: float
    seed: int
    owner: np.ndarray
    neighbour: np.ndarray
    coeff: np.ndarray   # |S_f| / |C_N - C_P| per internal face
    weight: np.ndarray  # interpolation weight per internal face
    volume: np.ndarray
    centre: np.ndarray  # (n, 3)

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])


def _spec(nx, ny, nz, lo, hi, jitter, seed) -> L.BoxSpec:
    s = L.BoxSpec()
    s.nx, s.ny, s.nz = nx, ny, nz
    s.lo = (C.c_double * 3)(*lo)
    s.hi = (C.c_double * 3)(*hi)
    s.jitter = jitter
    s.seed = seed
    return s


def box_mesh(nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
             jitter: float = 0.0, seed: int = SEED_BASE) -> BoxMesh:
    spec = _spec(nx, ny, nz, lo, hi, jitter, seed)
    n, nf = C.c_int(), C.c_int()
    L.call("cfdg_box_sizes", C.byref(spec), C.byref(n), C.byref(nf))
    n, nf = n.value, nf.value
    own = np.zeros(nf, np.int32)
    nei = np.zeros(nf, np.int32)
    coeff = np.zeros(nf)
    w = np.zeros(nf)
    vol = np.zeros(n)
    cen = np.zeros(3 * n)
    L.call("cfdg_box_build", C.byref(spec), L.iptr(own), L.iptr(nei), L.dptr(coeff), L.dptr(w),
           L.dptr(vol), L.dptr(cen))
    return BoxMesh(nx, ny, nz, tuple(lo), tuple(hi), jitter, seed, own, nei, coeff, w, vol,
                   cen.reshape(n, 3))


def uniform(seed: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    out = np.zeros(n)
    L.call("cfdg_uniform", C.c_uint64(seed), C.c_int64(n), lo, hi, L.dptr(out))
    return out


def zero_mean_rhs(n: int, seed: int) -> np.ndarray:
    """b ~ U[-1, 1] with its mean removed."""
    b = uniform(seed, n)
    return b - b.sum() / n


@dataclass
class LduSystem:
    """An LDU matrix (upper[f] = a(owner, neighbour), lower[f] = a(neighbour, owner)) with its
    right-hand side (src/fv/include/cfd/fv/ldu_matrix.hpp:36-41)."""

    n: int
    owner: np.ndarray
    neighbour: np.ndarray
    diag: np.ndarray
    upper: np.ndarray
    lower: np.ndarray
    source: np.ndarray

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])


def accumulate_diag(n: int, owner, neighbour, c_owner, c_neigh) -> np.ndarray:
    return (np.bincount(owner, weights=c_owner, minlength=n) +
            np.bincount(neighbour, weights=c_neigh, minlength=n))


def pressure_system(mesh: BoxMesh, kappa: Optional[np.ndarray] = None, pin: bool = True,
                    rhs_seed: Optional[int] = None) -> LduSystem:
    """Pressure Poisson system: off-diagonals -kappa_f c_f, zero-gradient walls, pinned cell 0."""
    c = mesh.coeff
    if kappa is not None:
        kp = kappa[mesh.owner]
        kn = kappa[mesh.neighbour]
        w = mesh.weight
        kf = kp * kn / np.maximum(w * kn + (1.0 - w) * kp, 1e-300)
        c = kf * c
    n = mesh.n_cells
    diag = accumulate_diag(n, mesh.owner, mesh.neighbour, c, c)
    src = zero_mean_rhs(n, rhs_seed if rhs_seed is not None else mesh.seed + 1)
    sysm = LduSystem(n, mesh.owner, mesh.neighbour, diag, -c, -c.copy(), src)
    if pin:
        set_reference(sysm, 0, 0.0)
    return sysm


def momentum_system(mesh: BoxMesh, dt: float = 1.0, rhs_seed: Optional[int] = None) -> LduSystem:
    """Momentum-like system: off-diagonals -c_f, diagonal 1.1 sum|off| + V / dt."""
    c = mesh.coeff
    n = mesh.n_cells
    diag = 1.1 * accumulate_diag(n, mesh.owner, mesh.neighbour, c, c) + mesh.volume / dt
    src = zero_mean_rhs(n, rhs_seed if rhs_seed is not None else mesh.seed + 2)
    return LduSystem(n, mesh.owner, mesh.neighbour, diag, -c, -c.copy(), src)


def sphere_kappa(mesh: BoxMesh, seed: int, radius: float = 0.2, rho_in: float = 1000.0,
                 rho_out: float = 1.0) -> np.ndarray:
    """kappa = 1 / rho with rho_in inside a sphere whose centre is drawn in [0.2, 0.8]^3."""
    centre = 0.2 + 0.6 * uniform(seed, 3, 0.0, 1.0)
    d = np.linalg.norm(mesh.centre - centre[None, :], axis=1)
    rho = np.where(d < radius, rho_in, rho_out)
    return 1.0 / rho


def set_reference(s: LduSystem, cell: int, value: float) -> None:
    """LduMatrix::set_reference (ldu_matrix.cpp:105-109)."""
    s.source[cell] += s.diag[cell] * value
    s.diag[cell] += s.diag[cell]


def csr_pattern(n: int, owner: np.ndarray, neighbour: np.ndarray) -> LduPattern:
    """build_csr_pattern (ldu_matrix.cpp:572-612) computed by libcfdgpu."""
    nf = int(owner.shape[0])
    nnz = n + 2 * nf
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    dp = np.zeros(n, np.int32)
    up = np.zeros(nf, np.int32)
    lp = np.zeros(nf, np.int32)
    own = np.ascontiguousarray(owner, np.int32)
    nei = np.ascontiguousarray(neighbour, np.int32)
    L.call("cfdg_csr_pattern", n, nf, L.iptr(own), L.iptr(nei), L.iptr(rp), L.iptr(ci),
           L.iptr(dp), L.iptr(up), L.iptr(lp))
    return LduPattern(n, nf, nnz, rp, ci, dp, up, lp)


def ldu_to_csr(p: LduPattern, diag, upper, lower) -> CsrView:
    """Values of the assembled CSR matrix (device_solver.cpp:127-135)."""
    val = np.zeros(p.nnz)
    val[p.diag_pos] = diag
    val[p.upper_pos] = upper
    val[p.lower_pos] = lower
    return CsrView(p.n, p.nnz, p.row_ptr, p.col, val)


# ------------------------------------------------------------------------------------------
# Benchmark configurations
# ------------------------------------------------------------------------------------------

def config_pressure(n_side: int, jitter: float, seed: int) -> tuple:
    """Uniform (jitter 0) or jittered n^3 box on the unit cube, pinned pressure system."""
    mesh = box_mesh(n_side, n_side, n_side, jitter=jitter, seed=seed)
    return mesh, pressure_system(mesh, rhs_seed=seed + 1)


def config1():
    """32^3 uniform box, FP64 SA-AMG-PCG to 1e-8."""
    return config_pressure(32, 0.0, SEED_BASE + 0)


def config2(n_side: int = 128):
    """128^3 jittered box (+-20% of the spacing): pressure and momentum systems."""
    seed = SEED_BASE + 1
    mesh = box_mesh(n_side, n_side, n_side, jitter=0.2, seed=seed)
    return mesh, pressure_system(mesh, rhs_seed=seed + 1), momentum_system(mesh, 1.0, seed + 2)


def config3(n_side: int = 160):
    """160^3 jittered box, rho = 1000 inside a seeded sphere of radius 0.2, harmonic faces."""
    seed = SEED_BASE + 2
    mesh = box_mesh(n_side, n_side, n_side, jitter=0.2, seed=seed)
    kappa = sphere_kappa(mesh, seed + 3)
    return mesh, pressure_system(mesh, kappa=kappa, rhs_seed=seed + 1)


def batch_shapes(count: int = 256, n_min: int = 4096, n_max: int = 500_000,
                 seed: int = SEED_BASE + 3):
    """Cell counts log-uniform in [n_min, n_max]; even indices uniform boxes, odd indices
    jittered; cell aspect 1:1:r with r log-uniform in [1, 8], dims m x m x round(n / m^2) with
    m = round((n / r)^(1/3))."""
    u = uniform(seed, 2 * count, 0.0, 1.0)
    shapes = []
    for i in range(count):
        n = int(round(math.exp(math.log(n_min) + u[2 * i] * (math.log(n_max) - math.log(n_min)))))
        r = math.exp(u[2 * i + 1] * math.log(8.0))
        m = max(2, int(round((n / r) ** (1.0 / 3.0))))
        nz = max(2, int(round(n / (m * m))))
        jitter = 0.2 if i % 2 == 1 else 0.0
        # cell size h x h x r h: domain [0,1] x [0,1] x [0, nz r / m]
        shapes.append(dict(nx=m, ny=m, nz=nz, lz=nz * r / m, jitter=jitter, seed=seed + 17 * i + 1))
    return shapes


def serving_requests(count: int, n_min: float = 1e4, n_max: float = 4e6, max_steps: int = 20,
                     rate: float = 200.0, seed: int = 20240607):
    """Serving workload: `count` requests, each a jittered box of log-uniform n_min..n_max cells
    (m x m x nz cells on [0,1] x [0,1] x [0, nz / m]) with its pressure system and a run of
    1..max_steps time-step right-hand sides, each the previous one plus 5% (in norm) of zero-mean
    uniform noise; arrival times (seconds) of a Poisson process of the given rate. numpy's
    default generator with the given seed draws sizes, step counts, noise and arrivals.

    Returns (requests, arrivals); a request is a dict with the system's LduSystem `system`, its
    LduPattern `pattern`, the assembled CsrView `csr` and the list `rhs`."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        m = max(4, int(round(n ** (1.0 / 3.0))))
        nz = max(4, int(round(n / (m * m))))
        mesh = box_mesh(m, m, nz, hi=(1.0, 1.0, nz / m), jitter=0.2, seed=seed + 31 * i)
        s = pressure_system(mesh, rhs_seed=seed + 31 * i + 1)
        pat = csr_pattern(s.n, s.owner, s.neighbour)
        steps = int(rng.integers(1, max_steps + 1))
        rhs = [s.source]
        for _ in range(steps - 1):
            noise = rng.uniform(-1.0, 1.0, s.n)
            noise -= noise.mean()
            rhs.append(rhs[-1] + 0.05 * np.linalg.norm(rhs[-1]) / np.linalg.norm(noise) * noise)
        reqs.append(dict(system=s, pattern=pat, csr=ldu_to_csr(pat, s.diag, s.upper, s.lower),
                         rhs=rhs))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    return reqs, arrivals
