"""Python bindings of the C ABI in include/cstfem_hip.h (ctypes over libcstfem_hip.so).

The names follow the reference's C++ interface (src/assembly.cpp, src/bc.cpp,
src/solver.cpp): ``build_stiffness_pattern``, ``assemble_stiffness``, ``apply_dirichlet``,
``setup`` / ``solve`` of a Jacobi-CG solver.  Errors raise :class:`CstfemError` (the
reference's ``std::runtime_error``) or :class:`CstfemLogicError` (``std::logic_error``)
with the reference's message prefixes.  There is no CPU fallback: when the HIP library is
missing or no GPU is visible, every call fails loudly.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import meshes  # noqa: F401

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libcstfem_hip.so")

OK, NOT_CONVERGED = 0, 2
ERR_ARGUMENT, ERR_DEVICE, ERR_BAD_DIAGONAL, ERR_MISSING_ENTRY = -1, -2, -3, -4
ERR_OUT_OF_RANGE, ERR_ALLOC, ERR_TRACTION_EDGE, ERR_CONFLICT, ERR_ORDER = -5, -6, -7, -8, -9
PLANE_STRESS, PLANE_STRAIN = 0, 1
METHOD_CG_FP64, METHOD_IR_FP32, METHOD_FCG_LEARNED, METHOD_CG_FP64_REFINED = 0, 1, 2, 3
METHOD_FCG_LEARNED_FP32 = 4


class CstfemError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class CstfemLogicError(CstfemError):
    pass


class SolverOptions(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("record_history", ctypes.c_int), ("instrument", ctypes.c_int),
                ("method", ctypes.c_int), ("inner_rtol", ctypes.c_double),
                ("max_outer", ctypes.c_int), ("refresh", ctypes.c_int)]


class SolverStatsC(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("breakdown", ctypes.c_int), ("zero_rhs", ctypes.c_int),
                ("outer_iterations", ctypes.c_int), ("fallback", ctypes.c_int),
                ("warning", ctypes.c_int), ("pad_", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("true_rel_residual", ctypes.c_double),
                ("breakdown_pap", ctypes.c_double), ("setup", ctypes.c_double),
                ("spmv_local", ctypes.c_double), ("dot_local", ctypes.c_double),
                ("vector_ops", ctypes.c_double), ("total", ctypes.c_double),
                ("spmv_calls", ctypes.c_longlong), ("reductions", ctypes.c_longlong)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("waves_per_group", ctypes.c_int), ("groups_per_cu", ctypes.c_int),
                ("max_team", ctypes.c_int)]


class KirschMetricsC(ctypes.Structure):
    _fields_ = [("l2_rel_stress_error", ctypes.c_double),
                ("l2_rel_displacement_error", ctypes.c_double),
                ("energy_rel_error", ctypes.c_double), ("scf_element", ctypes.c_int),
                ("pad_", ctypes.c_int), ("scf_x", ctypes.c_double), ("scf_y", ctypes.c_double),
                ("scf_numerical", ctypes.c_double), ("scf_analytical_centroid", ctypes.c_double),
                ("scf_theoretical", ctypes.c_double)]


class BatchResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("refinements", ctypes.c_int), ("status", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("true_rel_residual", ctypes.c_double)]


@dataclass
class SolverStats:
    """Mirror of SolverStats (solver.hpp:99-107) with the reference's message strings."""
    iterations: int = 0
    rel_residual: float = 0.0
    true_rel_residual: float = 0.0
    converged: bool = False
    message: str = ""
    warning: str = ""
    residual_history: list = field(default_factory=list)
    outer_iterations: int = 0
    fallback: int = 0  # 1 fell back to Jacobi-CG, 2 to the Chebyshev-smoothed multilevel cycle
    breakdown: dict = field(default_factory=dict)


# cstfem_traction_fn: (user, count, pts[4*count] = (x, y, nx, ny), t[2*count]) -> None
TRACTION_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))

_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_I64 = ctypes.POINTER(ctypes.c_int64)
_F = ctypes.POINTER(ctypes.c_float)

SIGNATURES = {
    "cstfem_last_error": (ctypes.c_char_p, []),
    "cstfem_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "cstfem_ctx_destroy": (ctypes.c_int, [_P]),
    "cstfem_device_info": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, _I]),
    "cstfem_upload_mesh": (ctypes.c_int, [_P, ctypes.c_int, _D, ctypes.c_int64, _I]),
    "cstfem_set_material": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_double, _D]),
    "cstfem_set_constitutive": (ctypes.c_int, [_P, _D, ctypes.c_double]),
    "cstfem_build_pattern": (ctypes.c_int, [_P, _I64]),
    "cstfem_download_pattern": (ctypes.c_int, [_P, _I, _I]),
    "cstfem_assemble": (ctypes.c_int, [_P]),
    "cstfem_download_values": (ctypes.c_int, [_P, _D]),
    "cstfem_upload_csr": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, _D]),
    "cstfem_set_rhs": (ctypes.c_int, [_P, _D, ctypes.c_int]),
    "cstfem_download_rhs": (ctypes.c_int, [_P, _D, ctypes.c_int]),
    "cstfem_add_body_force": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double]),
    "cstfem_add_traction": (ctypes.c_int, [_P, ctypes.c_int, _I, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double]),
    "cstfem_add_traction_fn": (ctypes.c_int, [_P, ctypes.c_int, _I, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "cstfem_element_stiffness": (ctypes.c_int, [_P, ctypes.c_int64, _D, _D, ctypes.c_double,
                                                _D]),
    "cstfem_apply_dirichlet": (ctypes.c_int, [_P, ctypes.c_int, _I, _D]),
    "cstfem_apply_dirichlet_noalloc": (ctypes.c_int, [_P, ctypes.c_int, _I, _D]),
    "cstfem_setup": (ctypes.c_int, [_P]),
    "cstfem_solve": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions),
                                    ctypes.POINTER(SolverStatsC), _D, _D]),
    "cstfem_download_solution": (ctypes.c_int, [_P, _D, ctypes.c_int]),
    "cstfem_spmv": (ctypes.c_int, [_P, _D, _D, ctypes.c_int]),
    "cstfem_dot": (ctypes.c_int, [_P, _D, _D, ctypes.c_int, _D]),
    "cstfem_profile_kernel": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _D]),
    "cstfem_set_material_fields": (ctypes.c_int, [_P, _D, _D, ctypes.c_int, ctypes.c_double]),
    "cstfem_batch_define": (ctypes.c_int, [_P, ctypes.c_int, _I]),
    "cstfem_batch_solve": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions),
                                          ctypes.POINTER(BatchConfig),
                                          ctypes.POINTER(BatchResult), _D, _D]),
    "cstfem_batch_autotune": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), ctypes.c_char_p,
                                             ctypes.POINTER(BatchConfig), _D]),
    "cstfem_mpnn_num_params": (ctypes.c_int, []),
    "cstfem_mpnn_init": (ctypes.c_int, [_P, ctypes.c_int]),
    "cstfem_mpnn_set_params": (ctypes.c_int, [_P, _F, ctypes.c_int]),
    "cstfem_mpnn_get_params": (ctypes.c_int, [_P, _F, ctypes.c_int]),
    "cstfem_mpnn_build": (ctypes.c_int, [_P]),
    "cstfem_mpnn_harvest": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I]),
    "cstfem_mpnn_set_pairs": (ctypes.c_int, [_P, ctypes.c_int, _F, _F]),
    "cstfem_mpnn_project_pairs": (ctypes.c_int, [_P]),
    "cstfem_mpnn_add_white_pairs": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "cstfem_batch_harvest": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), ctypes.c_int,
                                            ctypes.c_int, _I]),
    "cstfem_mpnn_loss_grad": (ctypes.c_int, [_P, _D]),
    "cstfem_mpnn_train_step": (ctypes.c_int, [_P, ctypes.c_float, _D]),
    "cstfem_mpnn_download": (ctypes.c_int, [_P, ctypes.c_int, _F, ctypes.c_int64]),
    "cstfem_element_stresses": (ctypes.c_int, [_P, _D, _D, _D, _D, _D]),
    "cstfem_kirsch_metrics": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double,
                                             ctypes.POINTER(KirschMetricsC)]),
    "cstfem_mg_setup": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int]),
    "cstfem_mg_prepare": (ctypes.c_int, [_P]),
    "cstfem_mg_levels": (ctypes.c_int, [_P, _I]),
    "cstfem_mg_level_info": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, _I64, _D]),
    "cstfem_mg_tail": (ctypes.c_int, [_P, _I, _I]),
    "cstfem_mg_download_aggregates": (ctypes.c_int, [_P, ctypes.c_int, _I, ctypes.c_int]),
    "cstfem_mg_download_matrix": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, _D]),
    "cstfem_mg_apply": (ctypes.c_int, [_P, _D, _D, ctypes.c_int]),
}


def lib():
    """Loads libcstfem_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def _check(rc: int):
    if rc in (OK, NOT_CONVERGED):
        return rc
    msg = lib().cstfem_last_error().decode()
    if rc in (ERR_MISSING_ENTRY, ERR_ORDER):
        raise CstfemLogicError(rc, msg)
    raise CstfemError(rc, msg)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Device:
    """One context on one GPU: mesh, matrix, right-hand side and solver state in HBM."""

    def __init__(self, device: int = 0):
        self._h = _P()
        _check(lib().cstfem_ctx_create(device, ctypes.byref(self._h)))
        self.n = 0
        self.nnz = 0
        self.mesh = None

    def close(self):
        if self._h:
            lib().cstfem_ctx_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        buf = ctypes.create_string_buffer(256)
        cu = ctypes.c_int()
        _check(lib().cstfem_device_info(self._h, buf, 256, ctypes.byref(cu)))
        return buf.value.decode(), cu.value

    # ---- mesh and material
    def upload_mesh(self, mesh):
        self.mesh = mesh
        xy, tri = _f64(mesh.xy), _i32(mesh.tri)
        _check(lib().cstfem_upload_mesh(self._h, int(xy.shape[0]), _ptr(xy, ctypes.c_double),
                                        int(tri.shape[0]), _ptr(tri, ctypes.c_int)))
        self.n = 2 * int(xy.shape[0])

    def set_material(self, E=1.0, nu=0.3, plane=PLANE_STRESS, thickness=1.0, E_elem=None):
        Ee = None if E_elem is None else _f64(E_elem)
        _check(lib().cstfem_set_material(self._h, float(E), float(nu), int(plane),
                                         float(thickness), _ptr(Ee, ctypes.c_double)))

    def set_constitutive(self, D, thickness=1.0):
        """Uniform 3x3 constitutive matrix, as assemble_stiffness(mesh, D, t, K) takes it."""
        D = _f64(D).reshape(-1)
        _check(lib().cstfem_set_constitutive(self._h, _ptr(D, ctypes.c_double), float(thickness)))

    # ---- pattern and assembly (assembly.cpp:9-74)
    def build_stiffness_pattern(self):
        nnz = ctypes.c_int64()
        _check(lib().cstfem_build_pattern(self._h, ctypes.byref(nnz)))
        self.nnz = nnz.value
        return self.nnz

    def download_pattern(self):
        row_ptr = np.empty(self.n + 1, dtype=np.int32)
        cols = np.empty(self.nnz, dtype=np.int32)
        _check(lib().cstfem_download_pattern(self._h, _ptr(row_ptr, ctypes.c_int),
                                             _ptr(cols, ctypes.c_int)))
        return row_ptr, cols

    def assemble_stiffness(self):
        _check(lib().cstfem_assemble(self._h))

    def download_values(self):
        v = np.empty(self.nnz, dtype=np.float64)
        _check(lib().cstfem_download_values(self._h, _ptr(v, ctypes.c_double)))
        return v

    def upload_csr(self, row_ptr, cols, vals):
        rp, c, v = _i32(row_ptr), _i32(cols), _f64(vals)
        n = int(rp.shape[0]) - 1
        _check(lib().cstfem_upload_csr(self._h, n, _ptr(rp, ctypes.c_int), _ptr(c, ctypes.c_int),
                                       _ptr(v, ctypes.c_double)))
        self.n, self.nnz = n, int(rp[-1])

    # ---- loads and Dirichlet (bc.cpp:137-253)
    def set_rhs(self, b):
        """The load vector; None sets it to zero on the device."""
        if b is None:
            _check(lib().cstfem_set_rhs(self._h, None, int(self.n)))
            return
        b = _f64(b)
        _check(lib().cstfem_set_rhs(self._h, _ptr(b, ctypes.c_double), int(b.shape[0])))

    def download_rhs(self):
        b = np.empty(self.n, dtype=np.float64)
        _check(lib().cstfem_download_rhs(self._h, _ptr(b, ctypes.c_double), self.n))
        return b

    def add_traction(self, edges, tx=0.0, ty=0.0, kirsch=None):
        e = _i32(edges).reshape(-1)
        if kirsch is not None:
            mode, p0, p1 = 1, float(kirsch[0]), float(kirsch[1])
        else:
            mode, p0, p1 = 0, float(tx), float(ty)
        _check(lib().cstfem_add_traction(self._h, int(e.size // 2), _ptr(e, ctypes.c_int), mode,
                                         p0, p1))

    def add_traction_fn(self, edges, fn):
        """The reference's TractionFn form (bc.hpp:71-72): fn(x, y, nx, ny) -> (tx, ty) with
        numpy arrays of the Gauss points' coordinates and outward normals."""
        e = _i32(edges).reshape(-1)

        def cb(_user, count, pts, t):
            p = np.ctypeslib.as_array(pts, shape=(count, 4))
            tx, ty = fn(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), p[:, 3].copy())
            out = np.ctypeslib.as_array(t, shape=(count, 2))
            out[:, 0] = tx
            out[:, 1] = ty

        cfn = TRACTION_FN(cb)
        _check(lib().cstfem_add_traction_fn(self._h, int(e.size // 2), _ptr(e, ctypes.c_int),
                                            ctypes.cast(cfn, ctypes.c_void_p), None))

    def element_stiffness(self, xy, D, thickness=1.0):
        """cst_stiffness (element.cpp:37-55) for a batch: xy [count x 6] -> Ke [count x 6 x 6]."""
        xy = _f64(xy).reshape(-1, 6)
        D = _f64(D).reshape(-1)
        Ke = np.empty((xy.shape[0], 6, 6), dtype=np.float64)
        _check(lib().cstfem_element_stiffness(self._h, int(xy.shape[0]), _ptr(xy, ctypes.c_double),
                                              _ptr(D, ctypes.c_double), float(thickness),
                                              _ptr(Ke, ctypes.c_double)))
        return Ke

    def add_body_force(self, bx, by):
        _check(lib().cstfem_add_body_force(self._h, float(bx), float(by)))

    def apply_dirichlet(self, dofs, values):
        d, v = _i32(dofs), _f64(values)
        _check(lib().cstfem_apply_dirichlet(self._h, int(d.size), _ptr(d, ctypes.c_int),
                                            _ptr(v, ctypes.c_double)))

    def apply_dirichlet_noalloc(self, dofs, values):
        """apply_dirichlet without device allocation (fixed per-GPU workspace, limits in
        include/cstfem_hip.h)."""
        d, v = _i32(dofs), _f64(values)
        _check(lib().cstfem_apply_dirichlet_noalloc(self._h, int(d.size), _ptr(d, ctypes.c_int),
                                                    _ptr(v, ctypes.c_double)))

    # ---- solve (solver.cpp:109-157)
    def setup(self):
        _check(lib().cstfem_setup(self._h))

    def solve(self, rtol=1e-10, max_iter=20000, record_history=False, instrument=False,
              method=METHOD_CG_FP64, inner_rtol=0.0, max_outer=0, refresh=0, download=True):
        o = SolverOptions(rtol, max_iter, int(record_history), int(instrument), method,
                          inner_rtol, max_outer, refresh)
        s = SolverStatsC()
        x = np.empty(self.n, dtype=np.float64) if download else None
        hist = np.zeros(max_iter + 1, dtype=np.float64) if record_history else None
        rc = _check(lib().cstfem_solve(self._h, ctypes.byref(o), ctypes.byref(s),
                                       _ptr(x, ctypes.c_double), _ptr(hist, ctypes.c_double)))
        st = _stats_from(s, o, rc, hist)
        return x, st

    def download_solution(self):
        x = np.empty(self.n, dtype=np.float64)
        _check(lib().cstfem_download_solution(self._h, _ptr(x, ctypes.c_double), self.n))
        return x

    # ---- batches of independent plates
    def set_material_fields(self, E_elem, nu_elem, plane=PLANE_STRESS, thickness=1.0):
        E, nu = _f64(E_elem), _f64(nu_elem)
        _check(lib().cstfem_set_material_fields(self._h, _ptr(E, ctypes.c_double),
                                                _ptr(nu, ctypes.c_double), int(plane),
                                                float(thickness)))

    def batch_define(self, node_off):
        off = _i32(node_off)
        self.batch_count = int(off.size - 1)
        _check(lib().cstfem_batch_define(self._h, self.batch_count, _ptr(off, ctypes.c_int)))

    def batch_solve(self, rtol=1e-10, max_iter=200000, method=METHOD_CG_FP64_REFINED,
                    config=None, inner_rtol=0.0, download=True):
        """Returns (x, [BatchResult], device seconds)."""
        o = SolverOptions(rtol, max_iter, 0, 0, method, inner_rtol, 0, 0)
        cfg = None if config is None else BatchConfig(*config)
        res = (BatchResult * self.batch_count)()
        x = np.empty(self.n, dtype=np.float64) if download else None
        t = ctypes.c_double()
        _check(lib().cstfem_batch_solve(self._h, ctypes.byref(o),
                                        ctypes.byref(cfg) if cfg is not None else None, res,
                                        _ptr(x, ctypes.c_double), ctypes.byref(t)))
        return x, list(res), t.value

    def batch_autotune(self, rtol=1e-10, max_iter=200000, method=METHOD_CG_FP64_REFINED,
                       cache_path=None):
        o = SolverOptions(rtol, max_iter, 0, 0, method, 0.0, 0, 0)
        best = BatchConfig()
        t = ctypes.c_double()
        path = cache_path.encode() if cache_path else None
        _check(lib().cstfem_batch_autotune(self._h, ctypes.byref(o), path, ctypes.byref(best),
                                           ctypes.byref(t)))
        return (best.waves_per_group, best.groups_per_cu, best.max_team), t.value

    def profile_kernel(self, kind, reps=20, nbytes=0):
        """Average seconds per launch of a hot kernel (0 fused SpMV, 1 SpMV, 2 update,
        3 direction, 4 copy of nbytes, 5 FP32 SpMV, 6-9 flexible-CG kernels, 10 read of
        nbytes)."""
        t = ctypes.c_double()
        _check(lib().cstfem_profile_kernel(self._h, int(kind), int(reps), int(nbytes),
                                           ctypes.byref(t)))
        return t.value

    # ---- stresses and Kirsch metrics (postprocess.cpp)
    def element_stresses(self, ne):
        """(sig [ne x 3], von_mises, szz, centroids [ne x 2], area) of the last solution."""
        sig = np.empty(3 * ne)
        vm, szz, area = np.empty(ne), np.empty(ne), np.empty(ne)
        cxy = np.empty(2 * ne)
        _check(lib().cstfem_element_stresses(self._h, _ptr(sig, ctypes.c_double),
                                             _ptr(vm, ctypes.c_double), _ptr(szz, ctypes.c_double),
                                             _ptr(cxy, ctypes.c_double),
                                             _ptr(area, ctypes.c_double)))
        return sig.reshape(ne, 3), vm, szz, cxy.reshape(ne, 2), area

    def kirsch_metrics(self, sigma_inf, a):
        m = KirschMetricsC()
        _check(lib().cstfem_kirsch_metrics(self._h, float(sigma_inf), float(a), ctypes.byref(m)))
        return {k: getattr(m, k) for k, _ in KirschMetricsC._fields_ if k != "pad_"}

    # ---- learned preconditioner
    def mpnn_num_params(self):
        return int(lib().cstfem_mpnn_num_params())

    def mpnn_init(self, seed=0):
        _check(lib().cstfem_mpnn_init(self._h, int(seed)))

    def mpnn_set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        _check(lib().cstfem_mpnn_set_params(self._h, _ptr(t, ctypes.c_float), int(t.size)))

    def mpnn_get_params(self):
        t = np.empty(self.mpnn_num_params(), dtype=np.float32)
        _check(lib().cstfem_mpnn_get_params(self._h, _ptr(t, ctypes.c_float), int(t.size)))
        return t

    def mpnn_build(self):
        _check(lib().cstfem_mpnn_build(self._h))

    def mpnn_harvest(self, k_first=1, k_last=50, stride=1):
        got = ctypes.c_int()
        _check(lib().cstfem_mpnn_harvest(self._h, int(k_first), int(k_last), int(stride),
                                         ctypes.byref(got)))
        return got.value

    def batch_harvest(self, k_first=1, k_last=50, rtol=1e-10, max_iter=200000):
        """Pairs (r_k, e_k) of every plate of the batch after plain Jacobi-CG iterations
        k_first..k_last (x* by the batched FP64 solve to rtol); returns the pair count."""
        o = SolverOptions(rtol, max_iter, 0, 0, METHOD_CG_FP64_REFINED, 0.0, 0, 0)
        got = ctypes.c_int()
        _check(lib().cstfem_batch_harvest(self._h, ctypes.byref(o), int(k_first), int(k_last),
                                          ctypes.byref(got)))
        return got.value

    def mpnn_add_white_pairs(self, count, seed=0):
        """Appends `count` white-noise pairs (e ~ N(0, I) drawn on the device, r = K e)."""
        _check(lib().cstfem_mpnn_add_white_pairs(self._h, int(count), int(seed)))

    def mpnn_project_pairs(self):
        """Pairs -> (K e', e'), e' = the error the multilevel coarse correction leaves."""
        _check(lib().cstfem_mpnn_project_pairs(self._h))

    def mpnn_set_pairs(self, R, E):
        R = None if R is None else np.ascontiguousarray(R, dtype=np.float32).reshape(-1, self.n)
        E = None if E is None else np.ascontiguousarray(E, dtype=np.float32).reshape(-1, self.n)
        K = (R if R is not None else E).shape[0]
        _check(lib().cstfem_mpnn_set_pairs(self._h, int(K), _ptr(R, ctypes.c_float),
                                           _ptr(E, ctypes.c_float)))

    def mpnn_loss_grad(self):
        loss = ctypes.c_double()
        _check(lib().cstfem_mpnn_loss_grad(self._h, ctypes.byref(loss)))
        return loss.value

    def mpnn_train_step(self, lr=1e-3):
        loss = ctypes.c_double()
        _check(lib().cstfem_mpnn_train_step(self._h, float(lr), ctypes.byref(loss)))
        return loss.value

    def mpnn_download(self, what, count):
        out = np.empty(int(count), dtype=np.float32)
        _check(lib().cstfem_mpnn_download(self._h, int(what), _ptr(out, ctypes.c_float),
                                          int(count)))
        return out

    # ---- multilevel preconditioner
    def mg_setup(self, max_levels=10, coarse_max=500, fine_smoother=0, kcycle_levels=3):
        """fine_smoother: 0 the learned M (mpnn_build first), 1 Chebyshev on D^-1 K;
        coarse levels 1..kcycle_levels use the K-cycle, deeper ones the V-cycle."""
        _check(lib().cstfem_mg_setup(self._h, int(max_levels), int(coarse_max),
                                     int(fine_smoother), int(kcycle_levels)))

    def mg_prepare(self):
        """Starts the fine level's aggregation (pattern only) on a host thread; the next
        mg_setup uses it.  Call after build_stiffness_pattern."""
        _check(lib().cstfem_mg_prepare(self._h))

    def mg_levels(self):
        n = ctypes.c_int()
        _check(lib().cstfem_mg_levels(self._h, ctypes.byref(n)))
        return n.value

    def mg_level_info(self, level):
        """(DOFs, aggregates, stored entries, smoothing lambda_max) of one level."""
        n, na, nnz, lam = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_double()
        _check(lib().cstfem_mg_level_info(self._h, int(level), ctypes.byref(n), ctypes.byref(na),
                                          ctypes.byref(nnz), ctypes.byref(lam)))
        return n.value, na.value, nnz.value, lam.value

    def mg_tail(self):
        """(first level of the fused single-workgroup tail or -1, packed operands 0/1)."""
        first, packed = ctypes.c_int(), ctypes.c_int()
        _check(lib().cstfem_mg_tail(self._h, ctypes.byref(first), ctypes.byref(packed)))
        return first.value, packed.value

    def mg_download_aggregates(self, level):
        n = self.mg_level_info(level)[0]
        nodes = n // (2 if level == 0 else 3)
        agg = np.empty(nodes, dtype=np.int32)
        _check(lib().cstfem_mg_download_aggregates(self._h, int(level), _ptr(agg, ctypes.c_int),
                                                   nodes))
        return agg

    def mg_download_matrix(self, level):
        n, _, nnz, _ = self.mg_level_info(level)
        rp = np.empty(n + 1, dtype=np.int32)
        cols = np.empty(nnz, dtype=np.int32)
        vals = np.empty(nnz, dtype=np.float64)
        _check(lib().cstfem_mg_download_matrix(self._h, int(level), _ptr(rp, ctypes.c_int),
                                               _ptr(cols, ctypes.c_int),
                                               _ptr(vals, ctypes.c_double)))
        return rp, cols, vals

    def mg_apply(self, r):
        r = _f64(r)
        z = np.empty(self.n, dtype=np.float64)
        _check(lib().cstfem_mg_apply(self._h, _ptr(r, ctypes.c_double), _ptr(z, ctypes.c_double),
                                     self.n))
        return z

    def dot(self, a, b):
        a, b = _f64(a), _f64(b)
        out = ctypes.c_double()
        _check(lib().cstfem_dot(self._h, _ptr(a, ctypes.c_double), _ptr(b, ctypes.c_double),
                                int(a.size), ctypes.byref(out)))
        return out.value

    def spmv(self, x):
        x = _f64(x)
        y = np.empty(self.n, dtype=np.float64)
        _check(lib().cstfem_spmv(self._h, _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double),
                                 self.n))
        return y


def _stats_from(s: SolverStatsC, o: SolverOptions, rc: int, hist) -> SolverStats:
    """Builds the reference's SolverStats, messages included (solver.cpp:46-130)."""
    st = SolverStats(iterations=s.iterations, rel_residual=s.rel_residual,
                     true_rel_residual=s.true_rel_residual, converged=bool(s.converged),
                     outer_iterations=s.outer_iterations, fallback=int(s.fallback))
    if s.zero_rhs:
        st.message = "zero right-hand side, x = 0"
    elif s.breakdown:
        st.message = (f"CG breakdown at iteration {s.iterations + 1}: p'Ap = "
                      f"{s.breakdown_pap:g} (matrix not SPD?)")
    elif s.converged:
        st.message = (f"converged in {s.iterations} iterations, relative residual "
                      f"{s.rel_residual:g}")
    else:
        st.message = (f"NOT converged: reached max_iter = {o.max_iter}, relative residual "
                      f"{s.rel_residual:g} (rtol {o.rtol:g})")
    if s.warning:
        st.warning = (f"WARNING: CG recursive residual converged ({s.rel_residual:g}) but the "
                      f"true residual ||b - Ax|| / ||b|| = {s.true_rel_residual:g} exceeds "
                      f"10 * rtol = {10 * o.rtol:g}")
    if hist is not None:
        cnt = 1 if s.zero_rhs else s.iterations + 1
        st.residual_history = list(hist[:cnt])
    st.breakdown = dict(setup=s.setup, spmv_local=s.spmv_local, dot_local=s.dot_local,
                        vector_ops=s.vector_ops, total=s.total, spmv_calls=s.spmv_calls,
                        reductions=s.reductions)
    return st
