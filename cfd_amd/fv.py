"""Finite-volume cell kernels on the GPU (gpu::DeviceFv) and synthetic finite-volume meshes.

`DeviceFv` mirrors the reference's class (src/gpu/include/cfd/gpu/fv_device.hpp:34-71): the
mesh is copied to the device once; `gradient`, `vector_gradient`, `scalar_corrections` and
`momentum_corrections` move only the fields. Results equal the reference's serial CPU face loops.

`fv_box` builds the mesh arrays the class takes (FvMeshView, fv_device.hpp:16-32) for a hex box:
vertices optionally graded, sheared by an affine map and jittered; internal faces in the
reference's order (owner ascending, then neighbour), boundary faces after them, patch by patch;
face area vectors and centres from the four triangles around the face's vertex mean; cell volumes
and centres from the pyramids on the faces; interpolation weights, non-orthogonal correction
vectors and least-squares arrays by the formulas of fv::FvGeometry
(src/fv/src/fv_geometry.cpp:17-92, 210-257).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L

GAUSS, LEAST_SQUARES, GAUSS_CORRECTED = 0, 1, 2


@dataclass
class FvMesh:
    n_cells: int
    n_internal_faces: int
    n_faces: int
    owner: np.ndarray         # n_faces
    neighbour: np.ndarray     # n_internal_faces
    cell_offsets: np.ndarray  # n_cells + 1
    cell_faces: np.ndarray    # faces of each cell in ascending order
    weight: np.ndarray        # n_internal_faces
    sf: np.ndarray            # (n_faces, 3)
    face_centre: np.ndarray   # (n_faces, 3)
    cell_centre: np.ndarray   # (n_cells, 3)
    volume: np.ndarray        # n_cells
    correction: np.ndarray    # (n_faces, 3)
    lsq_d: np.ndarray         # (n_faces, 3)
    lsq_inverse: np.ndarray   # (n_cells, 9), column-major 3x3
    points: Optional[np.ndarray] = None
    delta_coeffs: Optional[np.ndarray] = None  # n_faces: |S|^2 / (S . d)

    @property
    def n_boundary(self) -> int:
        return self.n_faces - self.n_internal_faces

    def c(self) -> L.FvMesh:
        m = L.FvMesh()
        m.n_cells, m.n_internal_faces, m.n_faces = self.n_cells, self.n_internal_faces, self.n_faces
        m.owner, m.neighbour = L.iptr(self.owner), L.iptr(self.neighbour)
        m.cell_offsets, m.cell_faces = L.iptr(self.cell_offsets), L.iptr(self.cell_faces)
        m.weight, m.sf, m.face_centre = L.dptr(self.weight), L.dptr(self.sf), L.dptr(self.face_centre)
        m.cell_centre, m.volume = L.dptr(self.cell_centre), L.dptr(self.volume)
        m.correction, m.lsq_d = L.dptr(self.correction), L.dptr(self.lsq_d)
        m.lsq_inverse = L.dptr(self.lsq_inverse)
        return m


def _quad_geometry(p0, p1, p2, p3):
    """Area vector and centre of quads given by their corner arrays (m, 3)."""
    xm = 0.25 * (p0 + p1 + p2 + p3)
    area = np.zeros_like(xm)
    moment = np.zeros_like(xm)
    weight = np.zeros(xm.shape[0])
    for a, b in ((p0, p1), (p1, p2), (p2, p3), (p3, p0)):
        s = 0.5 * np.cross(a - xm, b - xm)
        mag = np.linalg.norm(s, axis=1)
        area += s
        moment += mag[:, None] * (xm + a + b) / 3.0
        weight += mag
    return area, moment / weight[:, None]


def fv_box(nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), jitter: float = 0.0,
           seed: int = 1, shear: Sequence[float] = (0.0, 0.0, 0.0), grading: Sequence[float] = (1.0, 1.0, 1.0),
           xy_skew: float = 0.0) -> FvMesh:
    """Hex box of nx * ny * nz cells. grading: ratio of the last to the first spacing per axis
    (geometric); xy_skew: vertices off the x and y walls move by xy_skew (hx sin s, hy cos 1.7 s),
    s = 17 x + 31 y, the same at every z (the skewed box of the reference's tests/test_gpu.cpp:337-357:
    planar faces, skewed and non-orthogonal cells); shear: x += shear[0] y, y += shear[1] z,
    x += shear[2] z (keeps faces planar); jitter moves interior vertices by U[-jitter, jitter]
    times the spacing (warped faces)."""
    def axis(n, a, b, ratio):
        if ratio == 1.0 or n == 1:
            t = np.linspace(0.0, 1.0, n + 1)
        else:
            q = ratio ** (1.0 / (n - 1))
            h = q ** np.arange(n)
            t = np.concatenate([[0.0], np.cumsum(h)]) / h.sum()
        return a + (b - a) * t

    xs, ys, zs = (axis(n, a, b, r) for n, a, b, r in zip((nx, ny, nz), lo, hi, grading))
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")          # (nx+1, ny+1, nz+1)
    pts = np.stack([X, Y, Z], axis=-1)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        sp = [np.diff(v).min() for v in (xs, ys, zs)]
        inner = (slice(1, nx), slice(1, ny), slice(1, nz))
        shape = pts[inner].shape[:-1]
        for d in range(3):
            pts[inner + (d,)] += rng.uniform(-jitter, jitter, shape) * sp[d]
    if xy_skew != 0.0:
        hx, hy = (hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny
        x, y = pts[1:nx, 1:ny, :, 0].copy(), pts[1:nx, 1:ny, :, 1].copy()
        sk = 17.0 * x + 31.0 * y
        pts[1:nx, 1:ny, :, 0] = x + xy_skew * hx * np.sin(sk)
        pts[1:nx, 1:ny, :, 1] = y + xy_skew * hy * np.cos(1.7 * sk)
    x, y, z = pts[..., 0].copy(), pts[..., 1].copy(), pts[..., 2].copy()
    pts[..., 0] = x + shear[0] * y + shear[2] * z
    pts[..., 1] = y + shear[1] * z
    P = pts.transpose(2, 1, 0, 3).reshape(-1, 3)               # vertex id = i + (nx+1)(j + (ny+1)k)

    def vid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    def cid(i, j, k):
        return i + nx * (j + ny * k)

    own, nei, quads = [], [], []
    cells_all = np.arange(nx * ny * nz)
    I, J, K = cells_all % nx, (cells_all // nx) % ny, cells_all // (nx * ny)
    # internal faces towards +x, +y, +z
    for axis_id in range(3):
        m = [I < nx - 1, J < ny - 1, K < nz - 1][axis_id]
        i, j, k = I[m], J[m], K[m]
        if axis_id == 0:
            q = (vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i + 1, j + 1, k + 1), vid(i + 1, j, k + 1))
            n = cid(i + 1, j, k)
        elif axis_id == 1:
            q = (vid(i, j + 1, k), vid(i, j + 1, k + 1), vid(i + 1, j + 1, k + 1), vid(i + 1, j + 1, k))
            n = cid(i, j + 1, k)
        else:
            q = (vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1), vid(i, j + 1, k + 1))
            n = cid(i, j, k + 1)
        own.append(cid(i, j, k))
        nei.append(n)
        quads.append(np.stack(q, axis=1))
    own_i, nei_i, q_i = np.concatenate(own), np.concatenate(nei), np.concatenate(quads)
    order = np.lexsort((nei_i, own_i))
    own_i, nei_i, q_i = own_i[order], nei_i[order], q_i[order]
    # boundary patches: x-, x+, y-, y+, z-, z+ (outward quads), cells in order within a patch
    b_own, b_quads = [], []
    for axis_id in range(3):
        for side in (0, 1):
            lim = (nx, ny, nz)[axis_id]
            m = [I, J, K][axis_id] == (0 if side == 0 else lim - 1)
            i, j, k = I[m], J[m], K[m]
            c = cid(i, j, k)
            o = np.argsort(c, kind="stable")
            i, j, k, c = i[o], j[o], k[o], c[o]
            if axis_id == 0:
                a = i + side
                q = (vid(a, j, k), vid(a, j + 1, k), vid(a, j + 1, k + 1), vid(a, j, k + 1))
            elif axis_id == 1:
                a = j + side
                q = (vid(i, a, k), vid(i, a, k + 1), vid(i + 1, a, k + 1), vid(i + 1, a, k))
            else:
                a = k + side
                q = (vid(i, j, a), vid(i + 1, j, a), vid(i + 1, j + 1, a), vid(i, j + 1, a))
            q = np.stack(q, axis=1)
            if side == 0:
                q = q[:, ::-1]
            b_own.append(c)
            b_quads.append(q)
    owner = np.concatenate([own_i] + b_own).astype(np.int32)
    neighbour = nei_i.astype(np.int32)
    quads = np.concatenate([q_i] + b_quads)
    nc, nif, nf = nx * ny * nz, neighbour.shape[0], owner.shape[0]

    sf, cf = _quad_geometry(P[quads[:, 0]], P[quads[:, 1]], P[quads[:, 2]], P[quads[:, 3]])
    # cell volumes and centres: pyramids from the cell's vertex mean to each face
    corners = [vid(I + a, J + b, K + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    apex = np.mean([P[v] for v in corners], axis=0)            # (nc, 3), cells in id order
    vol = np.zeros(nc)
    mom = np.zeros((nc, 3))
    for cells, sign, faces in ((owner, 1.0, slice(0, nf)), (neighbour, -1.0, slice(0, nif))):
        xf = cf[faces]
        h = xf - apex[cells]
        v = sign * np.einsum("ij,ij->i", sf[faces], h) / 3.0
        np.add.at(vol, cells, v)
        np.add.at(mom, cells, v[:, None] * (apex[cells] + 0.75 * h))
    cc = mom / vol[:, None]

    # fv::FvGeometry (fv_geometry.cpp:17-92): weights, non-orthogonal corrections, least squares
    d = np.empty((nf, 3))
    d[:nif] = cc[neighbour] - cc[owner[:nif]]
    d[nif:] = cf[nif:] - cc[owner[nif:]]
    area = np.linalg.norm(sf, axis=1)
    dn = np.linalg.norm(d, axis=1)
    s_dot_d = np.maximum(np.einsum("ij,ij->i", sf, d), 0.05 * area * dn)
    delta = area * area / s_dot_d
    kv = sf - d * delta[:, None]
    dp = np.einsum("ij,ij->i", sf[:nif], cf[:nif] - cc[owner[:nif]])
    dq = np.einsum("ij,ij->i", sf[:nif], cc[neighbour] - cf[:nif])
    tot = dp + dq
    w = np.where(tot > 0.0, np.clip(dq / np.where(tot > 0.0, tot, 1.0), 0.0, 1.0), 0.5)
    dd = d / (dn * dn)[:, None]
    mats = np.zeros((nc, 3, 3))
    outer = d[:, :, None] * d[:, None, :] / (dn * dn)[:, None, None]
    np.add.at(mats, owner, outer)
    np.add.at(mats, neighbour, outer[:nif])
    inv = np.linalg.inv(mats)
    lsq = np.ascontiguousarray(inv.transpose(0, 2, 1).reshape(nc, 9))   # column-major

    cells = np.concatenate([owner, neighbour])
    faces = np.concatenate([np.arange(nf), np.arange(nif)]).astype(np.int32)
    order = np.lexsort((faces, cells))
    cell_faces = faces[order].astype(np.int32)
    counts = np.bincount(cells, minlength=nc)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    c64 = lambda a: np.ascontiguousarray(a, np.float64)
    return FvMesh(nc, nif, nf, owner, neighbour, offsets, cell_faces, c64(w), c64(sf), c64(cf), c64(cc),
                  c64(vol), c64(kv), c64(dd), lsq, points=P, delta_coeffs=c64(delta))


# Boundary kinds of the pressure equation (FlowBoundary, flow_solver.hpp:16).
VELOCITY_INLET, WALL, PRESSURE_OUTLET, SYMMETRY, EMPTY = 0, 1, 2, 3, 4


@dataclass
class PressureInputs:
    """Inputs of FlowSolver::pressure_correction's assembly (flow_solver.cpp:688-796): the
    momentum system (shared off-diagonals, per-component diagonal and source), the velocity, the
    mesh's delta coefficients and the boundary conditions per boundary face."""
    upper: np.ndarray            # n_internal_faces
    lower: np.ndarray            # n_internal_faces
    diag: np.ndarray             # (n_cells, 3)
    source: np.ndarray           # (n_cells, 3)
    u: np.ndarray                # (n_cells, 3)
    delta_coeffs: np.ndarray     # n_faces
    boundary_kind: np.ndarray    # n_boundary, VELOCITY_INLET .. EMPTY
    fixed_velocity: np.ndarray   # (n_boundary, 3)
    outlet_pressure: np.ndarray  # n_boundary
    u_prev: Optional[np.ndarray] = None    # (n_cells, 3): with phi_prev and alpha_u < 1, the
    phi_prev: Optional[np.ndarray] = None  # n_faces:      relaxation-independent face flux
    alpha_u: float = 1.0
    pin_reference: bool = True   # set_reference(0, 0.0): no fixed-pressure boundary


@dataclass
class PressureEquation:
    diag: np.ndarray      # n_cells
    upper: np.ndarray     # n_internal_faces
    lower: np.ndarray     # n_internal_faces
    source: np.ndarray    # n_cells
    phi_hbya: np.ndarray  # n_faces
    rau_f: np.ndarray     # n_faces
    hbya: np.ndarray      # (n_cells, 3)
    rau: np.ndarray       # n_cells


def synthetic_pressure_inputs(mesh: FvMesh, seed: int, random_kinds: bool = False, relax: bool = False,
                              nu: float = 1e-3, dt: float = 0.1) -> PressureInputs:
    """A momentum system of the box's shape for the pressure assembly: a swirling velocity field
    with seeded noise, upwind convection plus diffusion off-diagonals (nu |S|^2 / (S.d) + the
    upwinded face flux), per-component diagonals of the row sums plus V / dt (component
    variations of a few percent), sources A u + noise. Boundaries: inlet on x-min, outlet on
    x-max, symmetry on y-min, walls elsewhere, or (random_kinds) seeded kinds of all five;
    relax adds u_prev / phi_prev with alpha_u = 0.7."""
    rng = np.random.default_rng(seed)
    nc, nif, nf, nb = mesh.n_cells, mesh.n_internal_faces, mesh.n_faces, mesh.n_boundary
    x = mesh.cell_centre
    u = np.stack([np.sin(np.pi * x[:, 1]) + 0.3, np.cos(np.pi * x[:, 0]) * 0.5, 0.2 * x[:, 2]], axis=1)
    u = u + 0.05 * rng.standard_normal((nc, 3))
    own, nei = mesh.owner, mesh.neighbour
    w = mesh.weight
    uf = w[:, None] * u[own[:nif]] + (1.0 - w[:, None]) * u[nei]
    flux = np.einsum("ij,ij->i", uf, mesh.sf[:nif])
    dcoef = mesh.delta_coeffs if mesh.delta_coeffs is not None else np.linalg.norm(mesh.sf, axis=1)
    diff = nu * dcoef[:nif]
    upper = -(diff + np.maximum(-flux, 0.0))
    lower = -(diff + np.maximum(flux, 0.0))
    rows = np.zeros(nc)
    np.add.at(rows, own[:nif], -upper)
    np.add.at(rows, nei, -lower)
    base = rows + mesh.volume / dt
    diag = base[:, None] * (1.0 + 0.03 * rng.random((nc, 3)))
    source = diag * u + 0.01 * rng.standard_normal((nc, 3)) * base[:, None]
    cf = mesh.face_centre[nif:]
    if random_kinds:
        kind = rng.integers(0, 5, nb).astype(np.int32)
    else:
        lo, hi = cf.min(axis=0), cf.max(axis=0)
        tol = 1e-9 * max(1.0, float(np.abs(hi - lo).max()))
        kind = np.full(nb, WALL, np.int32)
        kind[cf[:, 1] <= lo[1] + tol] = SYMMETRY
        kind[cf[:, 0] <= lo[0] + tol] = VELOCITY_INLET
        kind[cf[:, 0] >= hi[0] - tol] = PRESSURE_OUTLET
    fixed = np.zeros((nb, 3))
    fixed[kind == VELOCITY_INLET] = (1.0, 0.0, 0.0)
    fixed[kind == WALL] = 0.01 * rng.standard_normal((int((kind == WALL).sum()), 3))
    pout = rng.random(nb)
    u_prev = phi_prev = None
    alpha = 1.0
    if relax:
        u_prev = u + 0.02 * rng.standard_normal((nc, 3))
        phi_prev = np.concatenate([flux, np.einsum("ij,ij->i", u[own[nif:]], mesh.sf[nif:])])
        phi_prev = phi_prev + 0.01 * rng.standard_normal(nf)
        alpha = 0.7
    c64 = lambda a: np.ascontiguousarray(a, np.float64)
    return PressureInputs(c64(upper), c64(lower), c64(diag), c64(source), c64(u), c64(dcoef), kind, c64(fixed),
                          c64(pout), None if u_prev is None else c64(u_prev),
                          None if phi_prev is None else c64(phi_prev), alpha, True)


def _d(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64).reshape(-1)
    if n is not None and a.shape[0] != n:
        raise L.InvalidArgument(1, f"array of {a.shape[0]} values where {n} are expected")
    return a


class DeviceFv:
    """gpu::DeviceFv (fv_device.hpp:34-71) on the GPU; raises NoDevice without one."""

    def __init__(self, mesh: FvMesh):
        self._mesh = mesh
        self._view = mesh.c()
        h = C.c_void_p()
        L.call("cfdg_fv_create", C.byref(self._view), C.byref(h))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            L.call("cfdg_fv_destroy", h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def n_cells(self) -> int:
        return self._mesh.n_cells

    def gradient(self, phi, bv, scheme: int) -> np.ndarray:
        m = self._mesh
        phi, bv = _d(phi, m.n_cells), _d(bv, m.n_boundary)
        out = np.empty(3 * m.n_cells)
        L.call("cfdg_fv_gradient", self._h, L.dptr(phi), L.dptr(bv), int(scheme), L.dptr(out))
        return out.reshape(m.n_cells, 3)

    def vector_gradient(self, u, ub, scheme: int) -> np.ndarray:
        """Column-major 3x3 per cell: out[c, 3 i + j] = d u_i / d x_j."""
        m = self._mesh
        u, ub = _d(u, 3 * m.n_cells), _d(ub, 3 * m.n_boundary)
        out = np.empty(9 * m.n_cells)
        L.call("cfdg_fv_vector_gradient", self._h, L.dptr(u), L.dptr(ub), int(scheme), L.dptr(out))
        return out.reshape(m.n_cells, 9)

    def scalar_corrections(self, phi, bv, scheme: int, flux, gamma, linear_upwind: bool, source) -> np.ndarray:
        """Returns source + the explicit internal-face fluxes (flux, gamma per internal face)."""
        m = self._mesh
        phi, bv = _d(phi, m.n_cells), _d(bv, m.n_boundary)
        flux, gamma = _d(flux)[:m.n_internal_faces].copy(), _d(gamma)[:m.n_internal_faces].copy()
        out = _d(source, m.n_cells).copy()
        L.call("cfdg_fv_scalar_corrections", self._h, L.dptr(phi), L.dptr(bv), int(scheme), L.dptr(flux),
               L.dptr(gamma), int(bool(linear_upwind)), L.dptr(out))
        return out

    def momentum_corrections(self, flux, nu_face, linear_upwind: bool, limited: bool, venkat: bool, eps2,
                             central: float, transpose_term: bool, source) -> np.ndarray:
        """Returns source (3 per cell) + the explicit momentum face fluxes, with the field and
        gradient of the last vector_gradient call."""
        m = self._mesh
        flux, nu = _d(flux)[:m.n_internal_faces].copy(), _d(nu_face)[:m.n_internal_faces].copy()
        e = _d(eps2 if eps2 is not None else np.zeros(3), 3)
        out = _d(source, 3 * m.n_cells).copy()
        L.call("cfdg_fv_momentum_corrections", self._h, L.dptr(flux), L.dptr(nu), int(bool(linear_upwind)),
               int(bool(limited)), int(bool(venkat)), L.dptr(e), float(central), int(bool(transpose_term)),
               L.dptr(out))
        return out.reshape(m.n_cells, 3)

    def pressure_equation(self, inp: PressureInputs) -> PressureEquation:
        """Assembles the pressure equation on the device (kept there for
        DeviceSolver.assemble_pressure) and returns it."""
        m = self._mesh
        nc, nif, nf, nb = m.n_cells, m.n_internal_faces, m.n_faces, m.n_boundary
        keep = [_d(inp.upper, nif), _d(inp.lower, nif), _d(inp.diag, 3 * nc), _d(inp.source, 3 * nc),
                _d(inp.u, 3 * nc), _d(inp.delta_coeffs, nf)]
        kind = np.ascontiguousarray(inp.boundary_kind, np.int32).reshape(-1)
        if kind.shape[0] != nb:
            raise L.InvalidArgument(1, f"boundary_kind has {kind.shape[0]} values where {nb} are expected")
        fixed, pout = _d(inp.fixed_velocity, 3 * nb), _d(inp.outlet_pressure, nb)
        up = None if inp.u_prev is None else _d(inp.u_prev, 3 * nc)
        pp = None if inp.phi_prev is None else _d(inp.phi_prev, nf)
        c = L.PressureInputs(*[L.dptr(a) for a in keep], L.iptr(kind), L.dptr(fixed), L.dptr(pout),
                             L.dptr(up) if up is not None else None, L.dptr(pp) if pp is not None else None,
                             float(inp.alpha_u), int(bool(inp.pin_reference)))
        eq = PressureEquation(np.empty(nc), np.empty(nif), np.empty(nif), np.empty(nc), np.empty(nf),
                              np.empty(nf), np.empty((nc, 3)), np.empty(nc))
        o = L.PressureOutputs(*[L.dptr(a) for a in (eq.diag, eq.upper, eq.lower, eq.source, eq.phi_hbya,
                                                    eq.rau_f, eq.hbya, eq.rau)])
        L.call("cfdg_fv_pressure_equation", self._h, C.byref(c), C.byref(o))
        return eq

    def on_device(self) -> bool:
        v = C.c_int()
        L.call("cfdg_fv_on_device", self._h, C.byref(v))
        return bool(v.value)

    def timing(self, enable: bool) -> None:
        L.call("cfdg_fv_timing", self._h, int(bool(enable)))

    def last_kernel_ms(self) -> float:
        ms = C.c_double()
        L.call("cfdg_fv_last_kernel_ms", self._h, C.byref(ms))
        return ms.value
