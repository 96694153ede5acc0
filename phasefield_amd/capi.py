"""ctypes binding of the C-ABI in ``include/ppf.h``.

`Library` loads one application library (``phasefield_amd/lib/libppf_<app>.so``) and
declares the signatures; `Context` wraps a ``ppf_ctx`` with Python methods that raise
`PPFError` on a non-zero status.  Device arrays are passed as raw pointers (for example
``tensor.data_ptr()`` of a torch tensor on the GPU).  A missing library is an error:
there is no fallback implementation.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_longlong, c_uint, c_void_p
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"

OK = 0
OP_RHS, OP_LHS, OP_LHS_PLAIN = 0, 1, 2
SCALE_NONE, SCALE_INVM, SCALE_INVM_SQRT = 0, 1, 2
FP64, FP32 = 0, 1
AGE_LATEST = -1

# Every symbol include/ppf.h declares.
EXPORTED_SYMBOLS = (
    "ppf_app_name",
    "ppf_create",
    "ppf_destroy",
    "ppf_last_error",
    "ppf_get_info",
    "ppf_field_name",
    "ppf_set_option",
    "ppf_set_field",
    "ppf_get_field",
    "ppf_field_device_ptr",
    "ppf_run",
    "ppf_solve_increment",
    "ppf_solve_block",
    "ppf_rotate_old",
    "ppf_set_time",
    "ppf_apply_operator",
    "ppf_apply_operator_noalloc",
    "ppf_inverse_diagonal",
    "ppf_linear_solve",
    "ppf_last_solve_stats",
    "ppf_last_newton_stats",
    "ppf_field_norms",
    "ppf_output_log",
    "ppf_stream",
    "ppf_synchronize",
    "ppf_adjoint_forward",
    "ppf_adjoint_vjp",
    "ppf_lumped_mass",
)


class PPFError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class Info(ctypes.Structure):
    _fields_ = [
        ("dim", c_int),
        ("degree", c_int),
        ("cells", c_int * 3),
        ("nodes", c_int * 3),
        ("periodic", c_int * 3),
        ("lower", c_double * 3),
        ("h", c_double * 3),
        ("n_nodes", c_longlong),
        ("n_dofs", c_longlong),
        ("n_fields", c_int),
        ("n_blocks", c_int),
        ("dt", c_double),
        ("time", c_double),
        ("increment", c_uint),
        ("n_increments", c_uint),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int),
        ("initial_residual", c_double),
        ("final_residual", c_double),
        ("tolerance", c_double),
        ("converged", c_int),
    ]


class NewtonStats(ctypes.Structure):
    """ppf_newton_stats: Newton updates, CG steps over all of them, last residual l2 norm,
    tolerance, converged flag (reference newton_solver.h:140-147 logs the same)."""

    _fields_ = [
        ("iterations", c_int),
        ("linear_iterations", c_int),
        ("residual", c_double),
        ("tolerance", c_double),
        ("converged", c_int),
    ]


def library_path(app: str) -> Path:
    return LIB_DIR / f"libppf_{app}.so"


class Library:
    """One loaded application library."""

    _cache: dict[str, "Library"] = {}

    def __init__(self, app: str):
        path = library_path(app)
        if not path.exists():
            raise FileNotFoundError(
                f"{path} is missing: build it with `python -m phasefield_amd.build {app}`"
            )
        self.app = app
        self.path = path
        self.lib = ctypes.CDLL(str(path))
        L = self.lib
        vp = c_void_p
        dp = POINTER(c_double)
        L.ppf_app_name.restype = c_char_p
        L.ppf_create.argtypes = [c_char_p, POINTER(vp)]
        L.ppf_destroy.argtypes = [vp]
        L.ppf_destroy.restype = None
        L.ppf_last_error.argtypes = [vp]
        L.ppf_last_error.restype = c_char_p
        L.ppf_get_info.argtypes = [vp, POINTER(Info)]
        L.ppf_field_name.argtypes = [vp, c_int, c_char_p, c_int]
        L.ppf_set_option.argtypes = [vp, c_char_p, c_char_p]
        L.ppf_set_field.argtypes = [vp, c_int, c_int, dp, c_longlong]
        L.ppf_get_field.argtypes = [vp, c_int, c_int, dp, c_longlong]
        L.ppf_field_device_ptr.argtypes = [vp, c_int, c_int, POINTER(vp)]
        L.ppf_run.argtypes = [vp, c_uint, POINTER(c_int)]
        L.ppf_solve_increment.argtypes = [vp, POINTER(c_int)]
        L.ppf_solve_block.argtypes = [vp, c_int]
        L.ppf_rotate_old.argtypes = [vp, c_int]
        L.ppf_set_time.argtypes = [vp, c_uint, c_double]
        L.ppf_apply_operator.argtypes = [vp, c_int, c_int, c_int, c_int, vp, vp]
        L.ppf_apply_operator_noalloc.argtypes = [vp, c_int, c_int, c_int, c_int, vp, vp]
        L.ppf_inverse_diagonal.argtypes = [vp, c_int, c_int, c_int, vp]
        L.ppf_linear_solve.argtypes = [vp, c_int, POINTER(SolveStats)]
        L.ppf_last_solve_stats.argtypes = [vp, c_int, POINTER(SolveStats)]
        L.ppf_last_newton_stats.argtypes = [vp, c_int, POINTER(NewtonStats)]
        L.ppf_field_norms.argtypes = [vp, c_int, c_int, dp, dp]
        L.ppf_output_log.argtypes = [vp]
        L.ppf_output_log.restype = c_char_p
        L.ppf_stream.argtypes = [vp]
        L.ppf_stream.restype = vp
        L.ppf_synchronize.argtypes = [vp]
        L.ppf_adjoint_forward.argtypes = [vp, vp, c_double, c_double, c_uint, c_int, vp]
        L.ppf_adjoint_vjp.argtypes = [
            vp, vp, vp, c_double, c_double, c_uint, c_uint, c_int, dp, dp, vp,
        ]
        L.ppf_lumped_mass.argtypes = [vp, vp]

    @classmethod
    def load(cls, app: str) -> "Library":
        if app not in cls._cache:
            cls._cache[app] = Library(app)
        return cls._cache[app]

    def name(self) -> str:
        return self.lib.ppf_app_name().decode()

    def missing_symbols(self) -> list[str]:
        return [s for s in EXPORTED_SYMBOLS if not hasattr(self.lib, s)]


class Context:
    """A ``ppf_ctx``: one simulation of the application on the current GPU."""

    def __init__(self, app: str, prm_text: str):
        self.library = Library.load(app)
        self.L = self.library.lib
        handle = c_void_p()
        st = self.L.ppf_create(prm_text.encode(), ctypes.byref(handle))
        if st != OK:
            raise PPFError(st, self.L.ppf_last_error(None).decode())
        self.handle = handle
        self.info = self.get_info()

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.L.ppf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, status: int) -> None:
        if status != OK:
            raise PPFError(status, self.L.ppf_last_error(self.handle).decode())

    # ------------------------------------------------------------------ queries
    def get_info(self) -> Info:
        info = Info()
        self._check(self.L.ppf_get_info(self.handle, ctypes.byref(info)))
        return info

    def field_names(self) -> list[str]:
        names = []
        for f in range(self.info.n_fields):
            buf = ctypes.create_string_buffer(128)
            self._check(self.L.ppf_field_name(self.handle, f, buf, 128))
            names.append(buf.value.decode())
        return names

    def full_shape(self) -> tuple[int, ...]:
        """deal.II node grid (periodic images included), slowest axis first."""
        i = self.info
        n = [i.degree * i.cells[d] + 1 for d in range(i.dim)]
        return tuple(reversed(n))

    def stored_shape(self) -> tuple[int, ...]:
        i = self.info
        return tuple(reversed([i.nodes[d] for d in range(i.dim)]))

    # ------------------------------------------------------------------ options
    def set_option(self, key: str, value) -> None:
        self._check(self.L.ppf_set_option(self.handle, key.encode(), str(value).encode()))

    # ------------------------------------------------------------------- fields
    def set_field(self, field: int, values: np.ndarray, age: int = 0) -> None:
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._check(
            self.L.ppf_set_field(
                self.handle, field, age, a.ctypes.data_as(POINTER(c_double)), a.size
            )
        )

    def get_field(self, field: int, age: int = 0) -> np.ndarray:
        out = np.empty(self.info.n_dofs, dtype=np.float64)
        self._check(
            self.L.ppf_get_field(
                self.handle, field, age, out.ctypes.data_as(POINTER(c_double)), out.size
            )
        )
        return out.reshape(self.full_shape())

    def field_device_ptr(self, field: int, age: int = 0) -> int:
        p = c_void_p()
        self._check(self.L.ppf_field_device_ptr(self.handle, field, age, ctypes.byref(p)))
        return p.value

    # -------------------------------------------------------------------- solve
    def run(self, n_increments: int) -> int:
        status = c_int()
        self._check(self.L.ppf_run(self.handle, n_increments, ctypes.byref(status)))
        return status.value

    def solve_increment(self) -> int:
        status = c_int()
        self._check(self.L.ppf_solve_increment(self.handle, ctypes.byref(status)))
        return status.value

    def solve_block(self, block_id: int) -> None:
        self._check(self.L.ppf_solve_block(self.handle, block_id))

    def rotate_old(self, block_id: int) -> None:
        self._check(self.L.ppf_rotate_old(self.handle, block_id))

    def set_time(self, increment: int, time: float) -> None:
        self._check(self.L.ppf_set_time(self.handle, increment, time))

    def apply_operator(self, block_id, which, scaling, precision, d_src, d_dst) -> None:
        self._check(
            self.L.ppf_apply_operator(
                self.handle, block_id, which, scaling, precision,
                c_void_p(d_src or 0), c_void_p(d_dst),
            )
        )

    def apply_operator_noalloc(self, block_id, which, scaling, precision, d_src, d_dst) -> None:
        self._check(
            self.L.ppf_apply_operator_noalloc(
                self.handle, block_id, which, scaling, precision,
                c_void_p(d_src or 0), c_void_p(d_dst),
            )
        )

    def inverse_diagonal(self, block_id, scaling, precision, d_dst) -> None:
        self._check(
            self.L.ppf_inverse_diagonal(self.handle, block_id, scaling, precision, c_void_p(d_dst))
        )

    def linear_solve(self, block_id: int) -> SolveStats:
        st = SolveStats()
        self._check(self.L.ppf_linear_solve(self.handle, block_id, ctypes.byref(st)))
        return st

    def last_solve_stats(self, block_id: int) -> SolveStats:
        st = SolveStats()
        self._check(self.L.ppf_last_solve_stats(self.handle, block_id, ctypes.byref(st)))
        return st

    def last_newton_stats(self, block_id: int) -> NewtonStats:
        st = NewtonStats()
        self._check(self.L.ppf_last_newton_stats(self.handle, block_id, ctypes.byref(st)))
        return st

    def field_norms(self, field: int, age: int = 0) -> tuple[float, float]:
        l2, integral = c_double(), c_double()
        self._check(
            self.L.ppf_field_norms(self.handle, field, age, ctypes.byref(l2), ctypes.byref(integral))
        )
        return l2.value, integral.value

    def output_log(self) -> str:
        return self.L.ppf_output_log(self.handle).decode()

    def stream(self) -> int:
        return self.L.ppf_stream(self.handle)

    def synchronize(self) -> None:
        self._check(self.L.ppf_synchronize(self.handle))

    # ------------------------------------------------------------------ adjoint
    def adjoint_forward(self, d_phi0, L, kappa, n_steps, precision, d_phi_final) -> None:
        self._check(
            self.L.ppf_adjoint_forward(
                self.handle, c_void_p(d_phi0), L, kappa, n_steps, precision, c_void_p(d_phi_final)
            )
        )

    def adjoint_vjp(self, d_phi0, d_lam, L, kappa, n_steps, every, precision, d_dphi0):
        dl, dk = c_double(), c_double()
        self._check(
            self.L.ppf_adjoint_vjp(
                self.handle, c_void_p(d_phi0), c_void_p(d_lam), L, kappa, n_steps, every,
                precision, ctypes.byref(dl), ctypes.byref(dk), c_void_p(d_dphi0),
            )
        )
        return dl.value, dk.value

    def lumped_mass(self, d_mass) -> None:
        self._check(self.L.ppf_lumped_mass(self.handle, c_void_p(d_mass)))
