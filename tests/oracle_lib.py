"""ctypes bindings of the CPU checkers: oracle/libcst_oracle.so (the C restatement) and
oracle/_ref/libcstfem_ref.so (the reference's own core behind oracle/ref_harness.cpp).
Both export the same signatures with prefixes ``oc_`` and ``ref_``."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "libcst_oracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libcstfem_ref.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class Stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("breakdown", ctypes.c_int), ("zero_rhs", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("true_rel_residual", ctypes.c_double),
                ("breakdown_pap", ctypes.c_double)]


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


class Oracle:
    def __init__(self, path, prefix):
        self.L = ctypes.CDLL(path)
        self.prefix = prefix
        L, f = self.L, lambda n: getattr(self.L, f"{prefix}_{n}")
        self._constitutive = f("constitutive")
        self._constitutive.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, _D]
        self._stiffness = f("stiffness")
        self._stiffness.argtypes = [_D, _D, ctypes.c_double, _D]
        self._pattern = f("build_pattern")
        self._pattern.restype = ctypes.c_int64
        self._pattern.argtypes = [ctypes.c_int, ctypes.c_int64, _I, _I, _I]
        self._assemble = f("assemble")
        self._assemble.restype = ctypes.c_int
        self._assemble.argtypes = [ctypes.c_int, ctypes.c_int64, _I, _D, _D, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_double, _I, _I, _D,
                                   _I, _I]
        self._spmv = f("spmv")
        self._spmv.argtypes = [ctypes.c_int, _I, _I, _D, _D, _D]
        self._dirichlet = f("apply_dirichlet")
        self._dirichlet.restype = ctypes.c_int
        self._dirichlet.argtypes = [ctypes.c_int, _I, _I, _D, _D, ctypes.c_int, _I, _D, _I]
        self._traction = f("add_traction")
        self._traction.restype = ctypes.c_int
        self._traction.argtypes = [ctypes.c_int, ctypes.c_int64, _I, _D, ctypes.c_int, _I,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, _D, _I]
        self._pcg = f("pcg_jacobi")
        self._pcg.restype = ctypes.c_int
        self._pcg.argtypes = [ctypes.c_int, _I, _I, _D, _D, _D, ctypes.c_double, ctypes.c_int,
                              _D, ctypes.POINTER(Stats)]
        self._stresses = f("element_stresses")
        self._stresses.argtypes = [ctypes.c_int64, _I, _D, _D, _D, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, _D, _D, _D]
        if prefix == "oc":
            self._pcg_mt = L.oc_pcg_jacobi_mt
            self._pcg_mt.restype = ctypes.c_int
            self._pcg_mt.argtypes = [ctypes.c_int, _I, _I, _D, _D, _D, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(Stats)]
            self._body = L.oc_add_body_force
            self._body.argtypes = [ctypes.c_int64, _I, _D, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _D]
            self._kirsch = L.oc_kirsch_metrics
            self._kirsch.argtypes = [ctypes.c_int64, _I, _D, _D, _D, _D, _D, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, _D]
            self._kstress = L.oc_kirsch_stress
            self._kstress.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _D]
        else:
            self._kirsch_ref = L.ref_kirsch_metrics
            self._kirsch_ref.argtypes = [ctypes.c_int, ctypes.c_int64, _I, _D, _D,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _D]
            L.ref_last_error.restype = ctypes.c_char_p

    # element.cpp
    def constitutive(self, E, nu, strain=0):
        D = np.zeros(9)
        self._constitutive(E, nu, strain, _p(D))
        return D

    def stiffness(self, xy, D, t):
        xy, D = np.ascontiguousarray(xy, float), np.ascontiguousarray(D, float)
        K = np.zeros(36)
        self._stiffness(_p(xy), _p(D), t, _p(K))
        return K.reshape(6, 6)

    # assembly.cpp
    def build_pattern(self, nn, tri):
        tri = np.ascontiguousarray(tri, np.int32)
        ne = tri.shape[0]
        row_ptr = np.zeros(2 * nn + 1, np.int32)
        nnz = self._pattern(nn, ne, _p(tri, ctypes.c_int), _p(row_ptr, ctypes.c_int), None)
        cols = np.zeros(nnz, np.int32)
        self._pattern(nn, ne, _p(tri, ctypes.c_int), _p(row_ptr, ctypes.c_int),
                      _p(cols, ctypes.c_int))
        return row_ptr, cols

    def assemble(self, xy, tri, row_ptr, cols, E=1.0, nu=0.3, strain=0, t=1.0, E_elem=None):
        xy = np.ascontiguousarray(xy, float).reshape(-1)
        tri = np.ascontiguousarray(tri, np.int32)
        vals = np.zeros(cols.shape[0])
        Ee = None if E_elem is None else np.ascontiguousarray(E_elem, float)
        br, bc = ctypes.c_int(), ctypes.c_int()
        rc = self._assemble(xy.size // 2, tri.shape[0], _p(tri, ctypes.c_int), _p(xy), _p(Ee),
                            E, nu, strain, t, _p(row_ptr, ctypes.c_int), _p(cols, ctypes.c_int),
                            _p(vals), ctypes.byref(br), ctypes.byref(bc))
        if rc != 0:
            raise RuntimeError(f"assemble failed ({rc})")
        return vals

    def spmv(self, row_ptr, cols, vals, x):
        x = np.ascontiguousarray(x, float)
        y = np.zeros(row_ptr.shape[0] - 1)
        self._spmv(y.size, _p(row_ptr, ctypes.c_int), _p(cols, ctypes.c_int), _p(vals), _p(x), _p(y))
        return y

    # bc.cpp
    def apply_dirichlet(self, row_ptr, cols, vals, b, dofs, values):
        vals, b = vals.copy(), b.copy()
        dofs = np.ascontiguousarray(dofs, np.int32)
        values = np.ascontiguousarray(values, float)
        bad = ctypes.c_int()
        rc = self._dirichlet(b.size, _p(row_ptr, ctypes.c_int), _p(cols, ctypes.c_int), _p(vals),
                             _p(b), dofs.size, _p(dofs, ctypes.c_int), _p(values),
                             ctypes.byref(bad))
        if rc != 0:
            raise RuntimeError(f"apply_dirichlet failed ({rc}) at {bad.value}")
        return vals, b

    def add_traction(self, xy, tri, edges, f, t=1.0, tx=0.0, ty=0.0, kirsch=None):
        xy = np.ascontiguousarray(xy, float).reshape(-1)
        tri = np.ascontiguousarray(tri, np.int32)
        e = np.ascontiguousarray(edges, np.int32).reshape(-1)
        mode, p0, p1 = (1, kirsch[0], kirsch[1]) if kirsch is not None else (0, tx, ty)
        bad = ctypes.c_int()
        rc = self._traction(xy.size // 2, tri.shape[0], _p(tri, ctypes.c_int), _p(xy), e.size // 2,
                            _p(e, ctypes.c_int), t, mode, p0, p1, _p(f), ctypes.byref(bad))
        if rc != 0:
            raise RuntimeError(f"add_traction failed ({rc})")
        return f

    def add_body_force(self, xy, tri, f, t, bx, by):
        xy = np.ascontiguousarray(xy, float).reshape(-1)
        tri = np.ascontiguousarray(tri, np.int32)
        self._body(tri.shape[0], _p(tri, ctypes.c_int), _p(xy), t, bx, by, _p(f))
        return f

    # solver.cpp
    def pcg_jacobi(self, row_ptr, cols, vals, b, rtol=1e-10, max_iter=20000, history=False):
        x = np.zeros(b.size)
        h = np.zeros(max_iter + 1) if history else None
        st = Stats()
        rc = self._pcg(b.size, _p(row_ptr, ctypes.c_int), _p(cols, ctypes.c_int), _p(vals),
                       _p(np.ascontiguousarray(b, float)), _p(x), rtol, max_iter, _p(h),
                       ctypes.byref(st))
        if rc < 0:
            raise RuntimeError(f"pcg_jacobi failed ({rc})")
        out = dict(iterations=st.iterations, converged=bool(st.converged),
                   breakdown=bool(st.breakdown), zero_rhs=bool(st.zero_rhs),
                   rel_residual=st.rel_residual, true_rel_residual=st.true_rel_residual)
        if history:
            out["history"] = h[: (1 if st.zero_rhs else st.iterations + 1)].copy()
        return x, out

    def pcg_jacobi_mt(self, row_ptr, cols, vals, b, rtol=1e-10, max_iter=20000, threads=1):
        """The same recurrence on `threads` OpenMP threads (timed CPU baseline)."""
        x = np.zeros(b.size)
        st = Stats()
        rc = self._pcg_mt(b.size, _p(row_ptr, ctypes.c_int), _p(cols, ctypes.c_int), _p(vals),
                          _p(np.ascontiguousarray(b, float)), _p(x), rtol, max_iter,
                          int(threads), ctypes.byref(st))
        if rc < 0:
            raise RuntimeError(f"pcg_jacobi_mt failed ({rc})")
        return x, dict(iterations=st.iterations, converged=bool(st.converged),
                       breakdown=bool(st.breakdown), zero_rhs=bool(st.zero_rhs),
                       rel_residual=st.rel_residual, true_rel_residual=st.true_rel_residual)

    # postprocess.cpp
    def element_stresses(self, xy, tri, u, E, nu, strain=0, E_elem=None):
        xy = np.ascontiguousarray(xy, float).reshape(-1)
        tri = np.ascontiguousarray(tri, np.int32)
        ne = tri.shape[0]
        sig, cxy, area = np.zeros(3 * ne), np.zeros(2 * ne), np.zeros(ne)
        Ee = None if E_elem is None else np.ascontiguousarray(E_elem, float)
        self._stresses(ne, _p(tri, ctypes.c_int), _p(xy), _p(np.ascontiguousarray(u, float)),
                       _p(Ee), E, nu, strain, _p(sig), _p(cxy), _p(area))
        return sig.reshape(ne, 3), cxy.reshape(ne, 2), area

    def kirsch_metrics(self, xy, tri, u, sigma_inf, a, E, nu, strain=0):
        xy = np.ascontiguousarray(xy, float).reshape(-1)
        tri = np.ascontiguousarray(tri, np.int32)
        u = np.ascontiguousarray(u, float)
        out = np.zeros(8)
        if self.prefix == "oc":
            sig, cxy, area = self.element_stresses(xy, tri, u, E, nu, strain)
            self._kirsch(tri.shape[0], _p(tri, ctypes.c_int), _p(xy), _p(u),
                         _p(np.ascontiguousarray(sig.reshape(-1))),
                         _p(np.ascontiguousarray(cxy.reshape(-1))), _p(area), sigma_inf, a, E,
                         nu, strain, _p(out))
        else:
            self._kirsch_ref(xy.size // 2, tri.shape[0], _p(tri, ctypes.c_int), _p(xy), _p(u),
                             sigma_inf, a, E, nu, strain, _p(out))
        return dict(l2_rel_stress_error=out[0], l2_rel_displacement_error=out[1],
                    energy_rel_error=out[2], scf_element=int(out[3]), scf_x=out[4],
                    scf_y=out[5], scf_numerical=out[6], scf_analytical_centroid=out[7])


def dirichlet_dofs(mesh, spec):
    """spec: list of (group, component, value) -> (dofs, values) in the reference's visiting
    order (add_dirichlet, bc.cpp:118-135), duplicates kept (values agree)."""
    dofs, vals = [], []
    for g, comp, v in spec:
        nodes = mesh.groups[g].reshape(-1)
        dofs.append(2 * nodes + comp)
        vals.append(np.full(nodes.size, float(v)))
    return np.concatenate(dofs).astype(np.int32), np.concatenate(vals)
