"""ctypes binding of the oracle (liboracle.so) and of the reference's own host solvers
(_ref/libcfdref.so, present only where the reference checkout was available at build time).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libcfdref.so"

P_INT = C.POINTER(C.c_int)
P_DBL = C.POINTER(C.c_double)
PP_INT = C.POINTER(P_INT)
PP_DBL = C.POINTER(P_DBL)


class OrcStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("residual", C.c_double)]


class OrcFvMesh(C.Structure):
    _fields_ = [("nc", C.c_int), ("nif", C.c_int), ("nf", C.c_int),
                ("own", P_INT), ("nei", P_INT), ("start", P_INT), ("faces", P_INT),
                ("w", P_DBL), ("sf", P_DBL), ("cf", P_DBL), ("cc", P_DBL), ("vol", P_DBL),
                ("kv", P_DBL), ("lsq_d", P_DBL), ("lsq", P_DBL)]


_orc = None
_ref = None


def _i(a):
    return np.ascontiguousarray(a, np.int32)


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def ip(a):
    return a.ctypes.data_as(P_INT)


def dp(a):
    return a.ctypes.data_as(P_DBL)


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not ORACLE_PATH.exists():
            raise ImportError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        h = C.CDLL(str(ORACLE_PATH))
        sig = {
            "orc_set_threads": (None, [C.c_int]),
            "orc_ldu_multiply": (None, [C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL, P_DBL, P_DBL]),
            "orc_ldu_residual": (None, [C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL, P_DBL, P_DBL, P_DBL]),
            "orc_ldu_normalised_residual": (C.c_double, [C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL, P_DBL, P_DBL]),
            "orc_ldu_set_reference": (None, [P_DBL, P_DBL, C.c_int, C.c_double]),
            "orc_csr_pattern": (C.c_int, [C.c_int, C.c_int, P_INT, P_INT, P_INT, P_INT, P_INT, P_INT, P_INT]),
            "orc_ldu_to_csr": (None, [C.c_int, C.c_int, P_INT, P_INT, P_INT, P_DBL, P_DBL, P_DBL, P_DBL]),
            "orc_assemble": (None, [C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL, C.c_double, P_DBL, P_DBL, P_DBL, P_DBL]),
            "orc_spmv": (None, [C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL]),
            "orc_dot": (C.c_double, [C.c_int, P_DBL, P_DBL]),
            "orc_solve": (OrcStats, [C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL, C.c_int, C.c_double, C.c_int]),
            "orc_aggregate": (C.c_int, [C.c_int, P_INT, P_INT, P_DBL, P_DBL, C.c_double, P_INT]),
            "orc_amg_setup": (C.c_void_p, [C.c_int, P_INT, P_INT, P_DBL, C.c_double, C.c_int, C.c_int]),
            "orc_amg_setup_decay": (C.c_void_p, [C.c_int, P_INT, P_INT, P_DBL, C.c_double, C.c_double,
                                                 C.c_int, C.c_int]),
            "orc_amg_free": (None, [C.c_void_p]),
            "orc_amg_levels": (C.c_int, [C.c_void_p]),
            "orc_amg_size": (None, [C.c_void_p, C.c_int, C.c_int, P_INT, P_INT, P_INT]),
            "orc_amg_copy": (None, [C.c_void_p, C.c_int, C.c_int, P_INT, P_INT, P_DBL]),
            "orc_amg_aggregates": (None, [C.c_void_p, C.c_int, P_INT]),
            "orc_amg_coarse_inverse": (C.c_int, [C.c_void_p, P_DBL]),
            "orc_amg_cg": (OrcStats, [C.c_int, P_INT, PP_INT, PP_INT, PP_DBL, PP_INT, PP_INT, PP_DBL,
                                      PP_INT, PP_INT, PP_DBL, P_DBL, C.c_int, C.c_double, P_DBL, P_DBL,
                                      C.c_double, C.c_int]),
            "orc_amg_cg_h": (OrcStats, [C.c_void_p, P_INT, P_INT, P_DBL, C.c_int, C.c_double, P_DBL,
                                        P_DBL, C.c_double, C.c_int]),
            "orc_sgs_pcg": (OrcStats, [C.c_void_p, P_DBL, P_DBL, C.c_double, C.c_int]),
            "orc_sgs_apply": (None, [C.c_void_p, P_DBL, P_DBL]),
            "orc_spgemm": (C.c_int, [C.c_int, C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_INT, P_INT, P_DBL,
                                     C.c_int, P_INT, P_INT, P_DBL]),
            "orc_fv_gradient": (None, [C.POINTER(OrcFvMesh), P_DBL, P_DBL, C.c_int, P_DBL]),
            "orc_fv_vector_gradient": (None, [C.POINTER(OrcFvMesh), P_DBL, P_DBL, C.c_int, P_DBL]),
            "orc_fv_scalar_corrections": (None, [C.POINTER(OrcFvMesh), P_DBL, P_DBL, C.c_int, P_DBL, P_DBL,
                                                 C.c_int, P_DBL]),
            "orc_fv_limiter": (None, [C.POINTER(OrcFvMesh), P_DBL, P_DBL, P_DBL, C.c_int, P_DBL, P_DBL]),
            "orc_pressure_equation": (None, [C.POINTER(OrcFvMesh)] + [P_DBL] * 6 + [P_INT, P_DBL, P_DBL, P_DBL,
                                             P_DBL, C.c_double, C.c_int] + [P_DBL] * 8),
            "orc_fv_momentum_corrections": (None, [C.POINTER(OrcFvMesh), P_DBL, P_DBL, P_DBL, P_DBL, P_DBL,
                                                   C.c_int, C.c_int, C.c_int, P_DBL, C.c_double, C.c_int,
                                                   P_DBL]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _orc = h
    return _orc


def ref_available() -> bool:
    return REF_PATH.exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        if not REF_PATH.exists():
            raise ImportError(f"{REF_PATH} missing (built only where the reference checkout exists)")
        h = C.CDLL(str(REF_PATH))
        stats = [P_INT, P_INT, P_DBL]
        amg_args = [C.c_int, P_INT, P_INT, PP_INT, PP_INT, PP_DBL, P_INT, PP_INT, PP_INT, PP_DBL,
                    P_INT, PP_INT, PP_INT, PP_DBL, P_DBL, C.c_int, C.c_double]
        sig = {
            "ref_backend_name": (C.c_int, [C.c_char_p, C.c_int]),
            "ref_solve_host": (None, [C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL, C.c_int,
                                      C.c_double, C.c_int] + stats),
            "ref_solve_amg_cg_host": (None, amg_args + [P_DBL, P_DBL, C.c_double, C.c_int] + stats),
            "ref_assemble": (None, [C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_DBL, P_DBL, C.c_double,
                                    P_DBL, P_DBL, P_DBL, P_DBL]),
            "ref_device_solver_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, P_INT, P_INT, P_INT,
                                                      P_INT, P_INT]),
            "ref_device_solver_destroy": (None, [C.c_void_p]),
            "ref_device_solver_set_amg": (None, [C.c_void_p] + amg_args),
            "ref_device_solver_solve": (None, [C.c_void_p, P_DBL, P_DBL, P_DBL, P_DBL, P_DBL, C.c_int,
                                               C.c_int, C.c_double, C.c_int] + stats),
            "ref_fv_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, P_INT, P_INT, P_INT, P_INT] + [P_DBL] * 8),
            "ref_fv_destroy": (None, [C.c_void_p]),
            "ref_fv_gradient": (None, [C.c_void_p, P_DBL, P_DBL, C.c_int, P_DBL]),
            "ref_fv_vector_gradient": (None, [C.c_void_p, P_DBL, P_DBL, C.c_int, P_DBL]),
            "ref_fv_scalar_corrections": (None, [C.c_void_p, P_DBL, P_DBL, C.c_int, P_DBL, P_DBL, C.c_int,
                                                 P_DBL]),
            "ref_fv_momentum_corrections": (None, [C.c_void_p, P_DBL, P_DBL, C.c_int, C.c_int, C.c_int, P_DBL,
                                                   C.c_double, C.c_int, P_DBL]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _ref = h
    return _ref


def set_threads(n: int) -> None:
    orc().orc_set_threads(int(n))


@dataclass
class Csr:
    n_rows: int
    n_cols: int
    rp: np.ndarray
    ci: np.ndarray
    va: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.rp[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.va, self.ci, self.rp), shape=(self.n_rows, self.n_cols))


@dataclass
class Stats:
    iterations: int
    converged: bool
    residual: float


def _stats(s: OrcStats) -> Stats:
    return Stats(s.iterations, bool(s.converged), s.residual)


# ---------------------------------------------------------------------------------------------
def csr_pattern(n: int, owner, neighbour):
    owner, neighbour = _i(owner), _i(neighbour)
    nf = owner.shape[0]
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(n + 2 * nf, np.int32)
    dpos = np.zeros(n, np.int32)
    upos = np.zeros(nf, np.int32)
    lpos = np.zeros(nf, np.int32)
    nnz = orc().orc_csr_pattern(n, nf, ip(owner), ip(neighbour), ip(rp), ip(ci), ip(dpos), ip(upos),
                                ip(lpos))
    return nnz, rp, ci, dpos, upos, lpos


def ldu_multiply(owner, neighbour, diag, upper, lower, x):
    owner, neighbour = _i(owner), _i(neighbour)
    diag, upper, lower, x = _d(diag), _d(upper), _d(lower), _d(x)
    y = np.zeros_like(diag)
    orc().orc_ldu_multiply(diag.shape[0], owner.shape[0], ip(owner), ip(neighbour), dp(diag),
                           dp(upper), dp(lower), dp(x), dp(y))
    return y


def ldu_normalised_residual(owner, neighbour, diag, upper, lower, source, x) -> float:
    owner, neighbour = _i(owner), _i(neighbour)
    arrs = [_d(v) for v in (diag, upper, lower, source, x)]
    return orc().orc_ldu_normalised_residual(arrs[0].shape[0], owner.shape[0], ip(owner),
                                             ip(neighbour), *[dp(v) for v in arrs])


def assemble(owner, neighbour, flux, gamma, delta, convection, diag, divergence):
    """Serial reference assembly; returns (upper, lower) and accumulates diag / divergence."""
    owner, neighbour = _i(owner), _i(neighbour)
    flux, gamma, delta = _d(flux), _d(gamma), _d(delta)
    nf = owner.shape[0]
    up = np.zeros(nf)
    lo = np.zeros(nf)
    orc().orc_assemble(diag.shape[0], nf, ip(owner), ip(neighbour), dp(flux), dp(gamma), dp(delta),
                       float(convection), dp(up), dp(lo), dp(diag), dp(divergence))
    return up, lo


def spmv(a: Csr, x) -> np.ndarray:
    x = _d(x)
    y = np.zeros(a.n_rows)
    orc().orc_spmv(a.n_rows, ip(a.rp), ip(a.ci), dp(a.va), dp(x), dp(y))
    return y


def solve(a: Csr, b, x, method: int, tol: float, max_it: int) -> Stats:
    b = _d(b)
    return _stats(orc().orc_solve(a.n_rows, ip(a.rp), ip(a.ci), dp(a.va), dp(b), dp(x), method,
                                  tol, max_it))


def aggregate(a: Csr, diag, theta: float = 0.08):
    diag = _d(diag)
    agg = np.zeros(a.n_rows, np.int32)
    n_agg = orc().orc_aggregate(a.n_rows, ip(a.rp), ip(a.ci), dp(a.va), dp(diag), theta, ip(agg))
    return agg, n_agg


class Hierarchy:
    """AmgPreconditioner::setup restated (amg.cpp:122-190)."""

    def __init__(self, a: Csr, theta: float = 0.08, max_levels: int = 25, coarse_size: int = 400,
                 decay: float = 1.0):
        self._keep = a
        self.h = orc().orc_amg_setup_decay(a.n_rows, ip(a.rp), ip(a.ci), dp(a.va), theta, decay,
                                           max_levels, coarse_size)

    def __del__(self):
        try:
            if self.h:
                orc().orc_amg_free(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def n_levels(self) -> int:
        return orc().orc_amg_levels(self.h)

    def matrix(self, level: int, which: int = 0) -> Csr:
        r, c, z = C.c_int(), C.c_int(), C.c_int()
        orc().orc_amg_size(self.h, level, which, C.byref(r), C.byref(c), C.byref(z))
        rp = np.zeros(r.value + 1, np.int32)
        ci = np.zeros(max(z.value, 1), np.int32)
        va = np.zeros(max(z.value, 1))
        orc().orc_amg_copy(self.h, level, which, ip(rp), ip(ci), dp(va))
        return Csr(r.value, c.value, rp, ci[: z.value], va[: z.value])

    def aggregates(self, level: int) -> np.ndarray:
        n = self.matrix(level).n_rows
        agg = np.zeros(n, np.int32)
        orc().orc_amg_aggregates(self.h, level, ip(agg))
        return agg

    def coarse_inverse(self) -> np.ndarray:
        nc = self.matrix(self.n_levels - 1).n_rows
        out = np.zeros(nc * nc)
        ok = orc().orc_amg_coarse_inverse(self.h, dp(out))
        if not ok:
            raise RuntimeError("oracle coarse LDL^T failed")
        return out.reshape(nc, nc)

    def level_sizes(self) -> List[int]:
        return [self.matrix(lvl).n_rows for lvl in range(self.n_levels)]

    def operator_complexity(self) -> float:
        nnz = [self.matrix(lvl).nnz for lvl in range(self.n_levels)]
        return float(sum(nnz)) / float(nnz[0])

    def amg_cg(self, b, x, tol: float, max_it: int, sweeps: int = 2, weight: float = 0.7,
               level0: Optional[Csr] = None) -> Stats:
        """Jacobi-V-cycle AMG-PCG; level0 replaces the finest operator (stale hierarchy)."""
        b = _d(b)
        if level0 is None:
            return _stats(orc().orc_amg_cg_h(self.h, None, None, None, sweeps, weight, dp(b), dp(x),
                                             tol, max_it))
        return _stats(orc().orc_amg_cg_h(self.h, ip(level0.rp), ip(level0.ci), dp(level0.va),
                                         sweeps, weight, dp(b), dp(x), tol, max_it))

    def sgs_pcg(self, b, x, tol: float, max_it: int) -> Stats:
        b = _d(b)
        return _stats(orc().orc_sgs_pcg(self.h, dp(b), dp(x), tol, max_it))

    def sgs_apply(self, r) -> np.ndarray:
        """One SGS V-cycle z = M^-1 r from z = 0 (amg.cpp:218-225)."""
        r = _d(r)
        z = np.zeros_like(r)
        orc().orc_sgs_apply(self.h, dp(r), dp(z))
        return z


def spgemm(a: Csr, b: Csr) -> Csr:
    """The setup's conservative sparse product A B, explicit zeros kept."""
    cap = max(1, a.n_rows * b.n_cols)
    rp = np.zeros(a.n_rows + 1, np.int32)
    ci = np.zeros(cap, np.int32)
    va = np.zeros(cap)
    nnz = orc().orc_spgemm(a.n_rows, a.n_cols, b.n_cols, ip(_i(a.rp)), ip(_i(a.ci)), dp(_d(a.va)), ip(_i(b.rp)),
                           ip(_i(b.ci)), dp(_d(b.va)), cap, ip(rp), ip(ci), dp(va))
    if nnz < 0:
        raise RuntimeError("spgemm: output larger than rows x cols")
    return Csr(a.n_rows, b.n_cols, rp, ci[:nnz], va[:nnz])


class _LevelArrays:
    """Pointer arrays of a hierarchy given as per-level CSR lists (None where absent)."""

    def __init__(self, a_list, p_list, r_list):
        nl = len(a_list)
        self.keep = []
        self.n = np.array([m.n_rows for m in a_list], np.int32)

        def ptrs(lst, attr, kind):
            arr = (P_INT * nl)() if kind == "i" else (P_DBL * nl)()
            for i, m in enumerate(lst):
                if m is None:
                    continue
                v = getattr(m, attr)
                self.keep.append(v)
                arr[i] = ip(v) if kind == "i" else dp(v)
            return arr

        def nnz(lst):
            return np.array([0 if m is None else m.nnz for m in lst], np.int32)

        self.nnz_a, self.nnz_p, self.nnz_r = nnz(a_list), nnz(p_list), nnz(r_list)
        self.a = (ptrs(a_list, "rp", "i"), ptrs(a_list, "ci", "i"), ptrs(a_list, "va", "d"))
        self.p = (ptrs(p_list, "rp", "i"), ptrs(p_list, "ci", "i"), ptrs(p_list, "va", "d"))
        self.r = (ptrs(r_list, "rp", "i"), ptrs(r_list, "ci", "i"), ptrs(r_list, "va", "d"))

    def args(self):
        return (len(self.n), ip(self.n), ip(self.nnz_a), *self.a, ip(self.nnz_p), *self.p,
                ip(self.nnz_r), *self.r)


def amg_cg_lists(a_list, p_list, r_list, coarse_inverse, b, x, tol, max_it, sweeps=2,
                 weight=0.7) -> Stats:
    """The oracle's Jacobi-V-cycle AMG-PCG on a hierarchy given as per-level CSR lists and a
    dense row-major coarse inverse (krylov_kernels.inc:182-301)."""
    arrs = _LevelArrays(a_list, p_list, r_list)
    ci = _d(coarse_inverse).reshape(-1)
    b = _d(b)
    return _stats(orc().orc_amg_cg(len(a_list), ip(arrs.n), *arrs.a, *arrs.p, *arrs.r, dp(ci),
                                   int(sweeps), float(weight), dp(b), dp(x), float(tol), int(max_it)))


def ref_solve_amg_cg_host(a_list, p_list, r_list, coarse_inverse, b, x, tol, max_it,
                          sweeps=2, weight=0.7) -> Stats:
    """The reference's cfd::gpu::solve_amg_cg_host on the given hierarchy."""
    arrs = _LevelArrays(a_list, p_list, r_list)
    ci = _d(coarse_inverse).reshape(-1)
    b = _d(b)
    it, cv, res = C.c_int(), C.c_int(), C.c_double()
    ref().ref_solve_amg_cg_host(*arrs.args(), dp(ci), sweeps, weight, dp(b), dp(x), tol, max_it,
                                C.byref(it), C.byref(cv), C.byref(res))
    return Stats(it.value, bool(cv.value), res.value)


def ref_solve_host(a: Csr, b, x, method: int, tol: float, max_it: int) -> Stats:
    b = _d(b)
    it, cv, res = C.c_int(), C.c_int(), C.c_double()
    ref().ref_solve_host(a.n_rows, a.nnz, ip(a.rp), ip(a.ci), dp(a.va), dp(b), dp(x), method, tol,
                         max_it, C.byref(it), C.byref(cv), C.byref(res))
    return Stats(it.value, bool(cv.value), res.value)


def ref_assemble(owner, neighbour, flux, gamma, delta, convection, diag, divergence):
    owner, neighbour = _i(owner), _i(neighbour)
    flux, gamma, delta = _d(flux), _d(gamma), _d(delta)
    nf = owner.shape[0]
    up = np.zeros(nf)
    lo = np.zeros(nf)
    ref().ref_assemble(diag.shape[0], nf, ip(owner), ip(neighbour), dp(flux), dp(gamma), dp(delta),
                       float(convection), dp(up), dp(lo), dp(diag), dp(divergence))
    return up, lo


def hierarchy_lists(h: Hierarchy):
    nl = h.n_levels
    a = [h.matrix(lvl, 0) for lvl in range(nl)]
    p = [h.matrix(lvl, 1) if lvl + 1 < nl else None for lvl in range(nl)]
    r = [h.matrix(lvl, 2) if lvl + 1 < nl else None for lvl in range(nl)]
    return a, p, r


# ---- Finite-volume cell kernels (gpu::DeviceFv) ----------------------------------------------

class FvOracle:
    """The oracle's restatement of gpu::DeviceFv for one mesh (any object with fv_box's fields);
    momentum_corrections takes the field and its gradient explicitly."""

    def __init__(self, mesh):
        self.mesh = mesh
        self._keep = [_i(mesh.owner), _i(mesh.neighbour), _i(mesh.cell_offsets), _i(mesh.cell_faces)] + \
            [_d(a) for a in (mesh.weight, mesh.sf, mesh.face_centre, mesh.cell_centre, mesh.volume,
                             mesh.correction, mesh.lsq_d, mesh.lsq_inverse)]
        k = self._keep
        self.m = OrcFvMesh(mesh.n_cells, mesh.n_internal_faces, mesh.n_faces, ip(k[0]), ip(k[1]), ip(k[2]),
                           ip(k[3]), dp(k[4]), dp(k[5]), dp(k[6]), dp(k[7]), dp(k[8]), dp(k[9]), dp(k[10]),
                           dp(k[11]))

    @property
    def nc(self):
        return self.mesh.n_cells

    def gradient(self, phi, bv, scheme):
        out = np.zeros(3 * self.nc)
        phi, bv = _d(phi), _d(bv)
        orc().orc_fv_gradient(C.byref(self.m), dp(phi), dp(bv), int(scheme), dp(out))
        return out.reshape(self.nc, 3)

    def vector_gradient(self, u, ub, scheme):
        out = np.zeros(9 * self.nc)
        u, ub = _d(u).reshape(-1), _d(ub).reshape(-1)
        orc().orc_fv_vector_gradient(C.byref(self.m), dp(u), dp(ub), int(scheme), dp(out))
        return out.reshape(self.nc, 9)

    def scalar_corrections(self, phi, bv, scheme, flux, gamma, linear_upwind, source):
        out = _d(source).reshape(-1).copy()
        phi, bv, flux, gamma = _d(phi), _d(bv), _d(flux), _d(gamma)
        orc().orc_fv_scalar_corrections(C.byref(self.m), dp(phi), dp(bv), int(scheme), dp(flux), dp(gamma),
                                        int(bool(linear_upwind)), dp(out))
        return out

    def pressure_equation(self, inp):
        """orc_pressure_equation on inputs with PressureInputs' fields; returns a dict of the
        outputs (diag, upper, lower, source, phi_hbya, rau_f, hbya, rau)."""
        m = self.mesh
        nc, nif, nf = m.n_cells, m.n_internal_faces, m.n_faces
        a = [_d(v).reshape(-1) for v in (inp.upper, inp.lower, inp.diag, inp.source, inp.u, inp.delta_coeffs)]
        kind = _i(np.asarray(inp.boundary_kind).reshape(-1))
        fixed, pout = _d(inp.fixed_velocity).reshape(-1), _d(inp.outlet_pressure).reshape(-1)
        up = None if inp.u_prev is None else _d(inp.u_prev).reshape(-1)
        pp = None if inp.phi_prev is None else _d(inp.phi_prev).reshape(-1)
        out = {k: np.zeros(n) for k, n in (("diag", nc), ("upper", nif), ("lower", nif), ("source", nc),
                                           ("phi_hbya", nf), ("rau_f", nf), ("hbya", 3 * nc), ("rau", nc))}
        orc().orc_pressure_equation(C.byref(self.m), *[dp(v) for v in a], ip(kind), dp(fixed), dp(pout),
                                    dp(up) if up is not None else None, dp(pp) if pp is not None else None,
                                    C.c_double(inp.alpha_u), int(bool(inp.pin_reference)),
                                    *[dp(out[k]) for k in ("diag", "upper", "lower", "source", "phi_hbya",
                                                           "rau_f", "hbya", "rau")])
        out["hbya"] = out["hbya"].reshape(nc, 3)
        return out

    def limiter(self, u, ub, grad, venkat, eps2):
        out = np.zeros(3 * self.nc)
        u, ub, grad, e = _d(u).reshape(-1), _d(ub).reshape(-1), _d(grad).reshape(-1), _d(eps2)
        orc().orc_fv_limiter(C.byref(self.m), dp(u), dp(ub), dp(grad), int(bool(venkat)), dp(e), dp(out))
        return out.reshape(self.nc, 3)

    def momentum_corrections(self, u, ub, grad, flux, nu_face, linear_upwind, limited, venkat, eps2, central,
                             transpose_term, source):
        out = _d(source).reshape(-1).copy()
        u, ub, grad = _d(u).reshape(-1), _d(ub).reshape(-1), _d(grad).reshape(-1)
        flux, nu, e = _d(flux), _d(nu_face), _d(eps2)
        orc().orc_fv_momentum_corrections(C.byref(self.m), dp(u), dp(ub), dp(grad), dp(flux), dp(nu),
                                          int(bool(linear_upwind)), int(bool(limited)), int(bool(venkat)), dp(e),
                                          float(central), int(bool(transpose_term)), dp(out))
        return out.reshape(self.nc, 3)


class FvReference:
    """The reference's own gpu::DeviceFv (host threads), from _ref/libcfdref.so."""

    def __init__(self, mesh):
        self.mesh = mesh
        self._keep = [_i(mesh.owner), _i(mesh.neighbour), _i(mesh.cell_offsets), _i(mesh.cell_faces)] + \
            [_d(a) for a in (mesh.weight, mesh.sf, mesh.face_centre, mesh.cell_centre, mesh.volume,
                             mesh.correction, mesh.lsq_d, mesh.lsq_inverse)]
        k = self._keep
        self.h = ref().ref_fv_create(mesh.n_cells, mesh.n_internal_faces, mesh.n_faces, ip(k[0]), ip(k[1]),
                                     ip(k[2]), ip(k[3]), *[dp(a) for a in k[4:]])

    def close(self):
        if self.h:
            ref().ref_fv_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def nc(self):
        return self.mesh.n_cells

    def gradient(self, phi, bv, scheme):
        out = np.zeros(3 * self.nc)
        phi, bv = _d(phi), _d(bv)
        ref().ref_fv_gradient(self.h, dp(phi), dp(bv), int(scheme), dp(out))
        return out.reshape(self.nc, 3)

    def vector_gradient(self, u, ub, scheme):
        out = np.zeros(9 * self.nc)
        u, ub = _d(u).reshape(-1), _d(ub).reshape(-1)
        ref().ref_fv_vector_gradient(self.h, dp(u), dp(ub), int(scheme), dp(out))
        return out.reshape(self.nc, 9)

    def scalar_corrections(self, phi, bv, scheme, flux, gamma, linear_upwind, source):
        out = _d(source).reshape(-1).copy()
        phi, bv, flux, gamma = _d(phi), _d(bv), _d(flux), _d(gamma)
        ref().ref_fv_scalar_corrections(self.h, dp(phi), dp(bv), int(scheme), dp(flux), dp(gamma),
                                        int(bool(linear_upwind)), dp(out))
        return out

    def momentum_corrections(self, flux, nu_face, linear_upwind, limited, venkat, eps2, central, transpose_term,
                             source):
        out = _d(source).reshape(-1).copy()
        flux, nu, e = _d(flux), _d(nu_face), _d(eps2)
        ref().ref_fv_momentum_corrections(self.h, dp(flux), dp(nu), int(bool(linear_upwind)), int(bool(limited)),
                                          int(bool(venkat)), dp(e), float(central), int(bool(transpose_term)),
                                          dp(out))
        return out.reshape(self.nc, 3)


# ------------------------------------------------------------------------------------------
# The reference's caller policy for the device solver, restated over the oracle's primitives:
# LduMatrix::solve with a CSR pattern and an AmgCache (src/fv/src/ldu_matrix.cpp:111-131,
# solve_persistent 287-366; AmgCache src/fv/include/cfd/fv/ldu_matrix.hpp:61-75), and the
# host AmgCg path's stale-hierarchy rule (src/fv/src/ldu_matrix.cpp:225-258) when
# stale_rebuild is set.
# ------------------------------------------------------------------------------------------
class AmgCacheState:
    """AmgCache (ldu_matrix.hpp:61-75): the hierarchy, its age and the iterations of the first
    solve after the last setup."""

    def __init__(self, max_age: int = 20, stale_rebuild: bool = False, decay: float = 1.0):
        self.h: Optional[Hierarchy] = None
        self.age = 0
        self.setup_iterations = 0
        self.max_age = max_age
        self.stale_rebuild = stale_rebuild
        self.decay = decay
        self.setups = 0


def ldu_csr(pattern, diag, upper, lower) -> Csr:
    """CSR values placed through the pattern's positions (ldu_matrix.cpp:315-323)."""
    nnz, rp, ci, dpos, upos, lpos = pattern
    va = np.zeros(nnz)
    va[dpos] = diag
    va[upos] = upper
    va[lpos] = lower
    return Csr(len(diag), len(diag), rp, ci[:nnz], va)


def ldu_solve(cache: AmgCacheState, owner, neighbour, pattern, diag, upper, lower, source, x,
              method: int, tolerance: float, rel_tolerance: float, max_iterations: int) -> dict:
    """LduMatrix::solve (ldu_matrix.cpp:111-131) for GpuCg (method 0) / GpuBiCgStab (method 1)
    with a pattern and a cache: solve_persistent (:287-366), x in/out."""
    n = len(diag)
    initial = ldu_normalised_residual(owner, neighbour, diag, upper, lower, source, x)
    out = dict(initial=initial, final=initial, iterations=0, solved=False, used_amg=False, setups=0,
               target=None)
    if initial < tolerance or n == 0:  # :121-124
        return out
    r = _d(source) - ldu_multiply(owner, neighbour, diag, upper, lower, x)
    r0 = float(np.cumsum(r * r)[-1])  # the serial sum of :293-297
    target = max(rel_tolerance, tolerance / max(initial, 1e-20)) * float(np.sqrt(r0))
    a = ldu_csr(pattern, diag, upper, lower)
    cg = method == 0
    fresh = False
    if cg and (cache.h is None or cache.age >= cache.max_age):  # :314-326
        cache.h = Hierarchy(a, decay=cache.decay)
        cache.age = 0
        cache.setups += 1
        out["setups"] += 1
        fresh = True
    if cg:  # :355-357
        cache.age += 1
    use_amg = cg and cache.h is not None and cache.h.n_levels > 1  # :358
    tol = max(target, 1e-300)
    if use_amg:
        st = cache.h.amg_cg(source, x, tol, max_iterations, level0=a)
    else:
        st = solve(a, source, x, 0 if cg else 1, tol, max_iterations)
    its = st.iterations
    if cache.stale_rebuild and cg:  # :242-256
        if fresh:
            cache.setup_iterations = max(its, 1)
        elif its > 2 * cache.setup_iterations + 2:
            cache.h = Hierarchy(a, decay=cache.decay)
            cache.setups += 1
            out["setups"] += 1
            use_amg = cache.h.n_levels > 1
            more = (cache.h.amg_cg(source, x, tol, max_iterations, level0=a) if use_amg
                    else solve(a, source, x, 0, tol, max_iterations))
            its += more.iterations
            cache.age = 1
            cache.setup_iterations = max(more.iterations, 1)
    out.update(final=ldu_normalised_residual(owner, neighbour, diag, upper, lower, source, x),
               iterations=its, solved=True, used_amg=use_amg, target=tol)
    return out
