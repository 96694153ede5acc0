"""ctypes binding of the oracle libraries.  TEST INFRASTRUCTURE.

All arrays are numpy float64 arrays over deal.II's full DoF set of the application's
mesh (periodic images included, x fastest), block vectors concatenated.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_longlong, c_uint, c_void_p

import numpy as np

from .build import library_path

_dp = POINTER(c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class OracleError(RuntimeError):
    pass


class Oracle:
    _libs: dict[str, ctypes.CDLL] = {}

    def __init__(self, app: str, prm_text: str, threads: int | None = None):
        if app not in Oracle._libs:
            path = library_path(app)
            if not path.exists():
                raise FileNotFoundError(f"{path} missing: run `python -m oracle.build {app}`")
            lib = ctypes.CDLL(str(path))
            lib.oracle_create.argtypes = [c_char_p, POINTER(c_void_p)]
            lib.oracle_destroy.argtypes = [c_void_p]
            lib.oracle_destroy.restype = None
            lib.oracle_last_error.argtypes = [c_void_p]
            lib.oracle_last_error.restype = c_char_p
            lib.oracle_set_threads.argtypes = [c_int]
            lib.oracle_info.argtypes = [c_void_p, POINTER(c_longlong), POINTER(c_int), POINTER(c_int)]
            lib.oracle_set_option.argtypes = [c_void_p, c_char_p, c_char_p]
            lib.oracle_set_field.argtypes = [c_void_p, c_int, c_int, _dp, c_longlong]
            lib.oracle_get_field.argtypes = [c_void_p, c_int, c_int, _dp, c_longlong]
            lib.oracle_set_time.argtypes = [c_void_p, c_uint, c_double]
            lib.oracle_apply_operator.argtypes = [c_void_p, c_int, c_int, c_int, _dp, _dp]
            lib.oracle_inverse_diagonal.argtypes = [c_void_p, c_int, c_int, _dp]
            lib.oracle_solve_block.argtypes = [c_void_p, c_int]
            lib.oracle_rotate_old.argtypes = [c_void_p, c_int]
            lib.oracle_solve_increment.argtypes = [c_void_p, POINTER(c_int)]
            lib.oracle_run.argtypes = [c_void_p, c_uint, POINTER(c_int)]
            lib.oracle_linear_solve.argtypes = [
                c_void_p, c_int, POINTER(c_int), _dp, _dp, _dp, POINTER(c_int),
            ]
            lib.oracle_last_newton_stats.argtypes = [
                c_void_p, c_int, POINTER(c_int), POINTER(c_int), _dp, _dp, POINTER(c_int),
            ]
            lib.oracle_field_norms.argtypes = [c_void_p, c_int, c_int, _dp, _dp]
            lib.oracle_output_log.argtypes = [c_void_p]
            lib.oracle_output_log.restype = c_char_p
            lib.oracle_lumped_mass.argtypes = [c_void_p, _dp]
            lib.oracle_adjoint_forward.argtypes = [c_void_p, _dp, c_double, c_double, c_uint, _dp]
            lib.oracle_adjoint_vjp.argtypes = [
                c_void_p, _dp, _dp, c_double, c_double, c_uint, _dp, _dp, _dp,
            ]
            lib.oracle_mg_level_ndofs.argtypes = [c_void_p, c_int, c_int, POINTER(c_longlong)]
            lib.oracle_mg_transfer.argtypes = [c_void_p, c_int, c_int, c_int, _dp, _dp]
            lib.oracle_mg_inv_diag.argtypes = [c_void_p, c_int, c_int, _dp]
            lib.oracle_mg_smooth.argtypes = [c_void_p, c_int, c_int, _dp, _dp, _dp, _dp]
            lib.oracle_mg_estimate.argtypes = [c_void_p, c_int, c_int, _dp, _dp, _dp]
            lib.oracle_mg_vcycle.argtypes = [c_void_p, c_int, _dp, _dp]
            lib.oracle_mg_level_vmult.argtypes = [c_void_p, c_int, c_int, _dp, _dp]
            lib.oracle_precondition.argtypes = [c_void_p, c_int, _dp, _dp]
            lib.oracle_krylov_steps.argtypes = [c_void_p, c_int, _dp, _dp, c_uint]
            Oracle._libs[app] = lib
        self.lib = Oracle._libs[app]
        if threads:
            self.lib.oracle_set_threads(threads)
        h = c_void_p()
        if self.lib.oracle_create(prm_text.encode(), ctypes.byref(h)) != 0:
            raise OracleError(self.lib.oracle_last_error(None).decode())
        self.h = h
        nd, nf, nb = c_longlong(), c_int(), c_int()
        self._check(self.lib.oracle_info(self.h, ctypes.byref(nd), ctypes.byref(nf), ctypes.byref(nb)))
        self.ndofs, self.nfields, self.nblocks = nd.value, nf.value, nb.value

    def _check(self, st):
        if st != 0:
            raise OracleError(self.lib.oracle_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_option(self, key, value):
        self._check(self.lib.oracle_set_option(self.h, key.encode(), str(value).encode()))

    def set_field(self, f, values, age=0):
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._check(self.lib.oracle_set_field(self.h, f, age, _ptr(a), a.size))

    def get_field(self, f, age=0):
        out = np.empty(self.ndofs)
        self._check(self.lib.oracle_get_field(self.h, f, age, _ptr(out), out.size))
        return out

    def set_time(self, increment, time):
        self._check(self.lib.oracle_set_time(self.h, increment, time))

    def apply_operator(self, block_id, which, scaling, src, nblock=1):
        dst = np.zeros(self.ndofs * nblock)
        s = np.ascontiguousarray(src if src is not None else np.zeros(self.ndofs * nblock), dtype=np.float64)
        self._check(self.lib.oracle_apply_operator(self.h, block_id, which, scaling, _ptr(s), _ptr(dst)))
        return dst

    def inverse_diagonal(self, block_id, scaling, nblock=1):
        dst = np.zeros(self.ndofs * nblock)
        self._check(self.lib.oracle_inverse_diagonal(self.h, block_id, scaling, _ptr(dst)))
        return dst

    def solve_block(self, block_id):
        self._check(self.lib.oracle_solve_block(self.h, block_id))

    def rotate_old(self, block_id):
        self._check(self.lib.oracle_rotate_old(self.h, block_id))

    def solve_increment(self):
        st = c_int()
        self._check(self.lib.oracle_solve_increment(self.h, ctypes.byref(st)))
        return st.value

    def run(self, n):
        st = c_int()
        self._check(self.lib.oracle_run(self.h, n, ctypes.byref(st)))
        return st.value

    def linear_solve(self, block_id):
        its, conv = c_int(), c_int()
        r0, r, tol = c_double(), c_double(), c_double()
        self._check(self.lib.oracle_linear_solve(
            self.h, block_id, ctypes.byref(its), ctypes.byref(r0), ctypes.byref(r),
            ctypes.byref(tol), ctypes.byref(conv)))
        return {"iterations": its.value, "initial_residual": r0.value,
                "final_residual": r.value, "tolerance": tol.value, "converged": bool(conv.value)}

    def last_newton_stats(self, block_id):
        its, lin, conv = c_int(), c_int(), c_int()
        r, tol = c_double(), c_double()
        self._check(self.lib.oracle_last_newton_stats(
            self.h, block_id, ctypes.byref(its), ctypes.byref(lin), ctypes.byref(r),
            ctypes.byref(tol), ctypes.byref(conv)))
        return {"iterations": its.value, "linear_iterations": lin.value, "residual": r.value,
                "tolerance": tol.value, "converged": bool(conv.value)}

    def field_norms(self, f, age=0):
        l2, integral = c_double(), c_double()
        self._check(self.lib.oracle_field_norms(self.h, f, age, ctypes.byref(l2), ctypes.byref(integral)))
        return l2.value, integral.value

    def output_log(self):
        return self.lib.oracle_output_log(self.h).decode()

    def lumped_mass(self):
        out = np.empty(self.ndofs)
        self._check(self.lib.oracle_lumped_mass(self.h, _ptr(out)))
        return out

    def adjoint_forward(self, phi0, L, kappa, n):
        p = np.ascontiguousarray(phi0, dtype=np.float64)
        out = np.empty(self.ndofs)
        self._check(self.lib.oracle_adjoint_forward(self.h, _ptr(p), L, kappa, n, _ptr(out)))
        return out

    def adjoint_vjp(self, phi0, lam, L, kappa, n):
        p = np.ascontiguousarray(phi0, dtype=np.float64)
        lm = np.ascontiguousarray(lam, dtype=np.float64)
        dl, dk = c_double(), c_double()
        out = np.empty(self.ndofs)
        self._check(self.lib.oracle_adjoint_vjp(self.h, _ptr(p), _ptr(lm), L, kappa, n,
                                                ctypes.byref(dl), ctypes.byref(dk), _ptr(out)))
        return dl.value, dk.value, out

    def mg_level_ndofs(self, block_id, level):
        n = c_longlong()
        self._check(self.lib.oracle_mg_level_ndofs(self.h, block_id, level, ctypes.byref(n)))
        return n.value

    def mg_prolongate(self, block_id, level, coarse):
        """fine = P coarse between V-cycle levels level-1 -> level (coarsest first)."""
        c = np.ascontiguousarray(coarse, dtype=np.float64)
        out = np.empty(self.mg_level_ndofs(block_id, level))
        self._check(self.lib.oracle_mg_transfer(self.h, block_id, level, 0, _ptr(c), _ptr(out)))
        return out

    def mg_restrict(self, block_id, level, fine):
        """coarse = P^T fine between V-cycle levels level -> level-1."""
        f = np.ascontiguousarray(fine, dtype=np.float64)
        out = np.empty(self.mg_level_ndofs(block_id, level - 1))
        self._check(self.lib.oracle_mg_transfer(self.h, block_id, level, 1, _ptr(f), _ptr(out)))
        return out

    def mg_inv_diag(self, block_id, level):
        """Inverse diagonal of V-cycle level `level` (coarsest first), as its smoother uses it."""
        out = np.empty(self.mg_level_ndofs(block_id, level))
        self._check(self.lib.oracle_mg_inv_diag(self.h, block_id, level, _ptr(out)))
        return out

    def mg_smooth(self, block_id, level, rhs):
        """x = S rhs: one application of the level's Chebyshev smoother from zero;
        returns (x, lmax, lmin)."""
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.empty_like(r)
        lmax, lmin = c_double(), c_double()
        self._check(self.lib.oracle_mg_smooth(self.h, block_id, level, _ptr(r), _ptr(x),
                                              ctypes.byref(lmax), ctypes.byref(lmin)))
        return x, lmax.value, lmin.value

    def mg_estimate(self, block_id, level, x0):
        """The smoother's eigenvalue estimate from start vector x0: (lmax, lmin)."""
        v = np.ascontiguousarray(x0, dtype=np.float64)
        lmax, lmin = c_double(), c_double()
        self._check(self.lib.oracle_mg_estimate(self.h, block_id, level, _ptr(v),
                                                ctypes.byref(lmax), ctypes.byref(lmin)))
        return lmax.value, lmin.value

    def mg_vcycle(self, block_id, defect):
        """u = V(defect): one V-cycle on the finest level."""
        d = np.ascontiguousarray(defect, dtype=np.float64)
        u = np.empty_like(d)
        self._check(self.lib.oracle_mg_vcycle(self.h, block_id, _ptr(d), _ptr(u)))
        return u

    def mg_level_vmult(self, block_id, level, src):
        x = np.ascontiguousarray(src, dtype=np.float64)
        y = np.empty_like(x)
        self._check(self.lib.oracle_mg_level_vmult(self.h, block_id, level, _ptr(x), _ptr(y)))
        return y

    def precondition(self, block_id, g):
        """h = P^{-1} g with the block's preconditioner, set up as for a solve."""
        v = np.ascontiguousarray(g, dtype=np.float64)
        out = np.empty_like(v)
        self._check(self.lib.oracle_precondition(self.h, block_id, _ptr(v), _ptr(out)))
        return out

    def krylov_steps(self, block_id, x0, rhs, steps):
        """Exactly `steps` steps of the block's GMRES / FGMRES from x0."""
        x = np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        self._check(self.lib.oracle_krylov_steps(self.h, block_id, _ptr(x), _ptr(b), steps))
        return x
