"""ctypes access to the thin HIP layer (include/ps26119_hip.h, libps26119_hip.so).

Used for kernel-level checks: a ``Context`` uploads a scaled problem and exposes the reference
Backend operations on device vector slots (upload/download/fill/copy/axpby/spmv/spmv_t/dot/
norms/steps/r2h steps/kkt) plus the fused chunk/check entry points.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.normpath(os.path.join(_HERE, "..", "..", "csrc", "hip", "libps26119_hip.so"))

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)

X, X0, XH, XB, DX, ATY, Y, Y0, YH, YB, DY, AX = range(12)
RED_TREE, RED_CPU, RED_DECISION = 0, 1, 2


class Problem(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("nnz", ctypes.c_longlong), ("a_ptr", I), ("a_col", I),
                ("a_val", D), ("a_val_orig", D), ("at_ptr", I), ("at_col", I), ("ato_col", I), ("at_val", D),
                ("ato_val", D), ("c", D), ("l", D), ("u", D), ("rl", D), ("ru", D), ("row_scale", D),
                ("col_scale", D), ("bound_scale", ctypes.c_double), ("obj_scale", ctypes.c_double), ("cl_o", D),
                ("cu_o", D), ("rl_o", D), ("ru_o", D), ("c_min", D)]


class Kkt(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("primal_residual", "dual_residual", "primal_obj", "dual_obj", "b_norm",
                                               "c_norm", "primal_max_rel", "dual_max_rel")]

    def as_list(self):
        return [getattr(self, k) for k, _ in self._fields_]


class Ray(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("rmax", "dual_obj", "dual_viol", "dual_mag", "dmax", "primal_cd",
                                               "primal_viol", "primal_cmag")]


class ChunkArgs(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("inner0", ctypes.c_longlong), ("tau", ctypes.c_double),
                ("sigma", ctypes.c_double), ("rho", ctypes.c_double), ("eta", ctypes.c_double),
                ("omega", ctypes.c_double), ("want_hat_last", ctypes.c_int), ("mode", ctypes.c_int)]


class CheckArgs(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_double), ("eta", ctypes.c_double), ("want_ray", ctypes.c_int),
                ("want_drift", ctypes.c_int), ("want_pid", ctypes.c_int), ("mode", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("px", ctypes.c_double), ("py", ctypes.c_double), ("cross", ctypes.c_double),
                ("r0_px", ctypes.c_double), ("r0_py", ctypes.c_double), ("r0_cross", ctypes.c_double),
                ("have_r0", ctypes.c_int), ("kkt", Kkt), ("ray", Ray), ("drift", Ray), ("dx_norm", ctypes.c_double),
                ("dy_norm", ctypes.c_double), ("finite", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB):
            raise OSError(f"{HIP_LIB} not built")
        L = ctypes.CDLL(HIP_LIB)
        P = ctypes.c_void_p
        sig = {
            "psh_device_count": ([], ctypes.c_int),
            "psh_create": ([ctypes.c_int, ctypes.POINTER(P)], ctypes.c_int),
            "psh_destroy": ([P], None),
            "psh_error": ([P], ctypes.c_char_p),
            "psh_sync": ([P], ctypes.c_int),
            "psh_load": ([P, ctypes.POINTER(Problem)], ctypes.c_int),
            "psh_set_fp32": ([P, ctypes.c_int], ctypes.c_int),
            "psh_upload": ([P, ctypes.c_int, D], ctypes.c_int),
            "psh_download": ([P, ctypes.c_int, D], ctypes.c_int),
            "psh_fill": ([P, ctypes.c_int, ctypes.c_double], ctypes.c_int),
            "psh_copy": ([P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "psh_axpby": ([P, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int], ctypes.c_int),
            "psh_spmv": ([P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "psh_spmv_t": ([P, ctypes.c_int, ctypes.c_int], ctypes.c_int),
            "psh_dot": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, D], ctypes.c_int),
            "psh_diff_norm2": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, D], ctypes.c_int),
            "psh_primal_step": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int], ctypes.c_int),
            "psh_dual_step": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int], ctypes.c_int),
            "psh_r2h_primal": ([P] + [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int] * 2, ctypes.c_int),
            "psh_r2h_dual": ([P] + [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int] * 2, ctypes.c_int),
            "psh_kkt_of": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Kkt)], ctypes.c_int),
            "psh_ray_of": ([P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Ray)], ctypes.c_int),
            "psh_ray_vectors": ([P, ctypes.c_int, D, D], ctypes.c_int),
            "psh_chunk": ([P, ctypes.POINTER(ChunkArgs)], ctypes.c_int),
            "psh_check": ([P, ctypes.POINTER(CheckArgs), ctypes.POINTER(Stats)], ctypes.c_int),
            "psh_restart": ([P], ctypes.c_int),
            "psh_lanczos_begin": ([P, D], ctypes.c_int),
            "psh_lanczos_step": ([P, ctypes.c_double, ctypes.c_int, D, D], ctypes.c_int),
            "psh_lanczos_advance": ([P, ctypes.c_double], ctypes.c_int),
            "psh_lanczos_product": ([P, D, D], ctypes.c_int),
            "psh_dot_noalloc": ([P, P, ctypes.c_longlong, ctypes.c_int, P, P], ctypes.c_int),
            "psh_noalloc_error": ([], ctypes.c_char_p),
            "psh_graphs": ([P, ctypes.c_int], ctypes.c_int),
            "psh_sliced": ([P, ctypes.c_int], ctypes.c_int),
            "psh_delta": ([P, ctypes.c_int], ctypes.c_int),
            "psh_scale": ([P, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D], ctypes.c_int),
            "psh_set_vectors": ([P, D, D, D, D, D, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Context:
    """One device context with a scaled problem loaded.

    ``scaled`` is a dict with the scaled CSR data (as produced by ``scaled_problem``)."""

    def __init__(self, scaled: dict, device: int = 0, sliced: int = 1):
        """sliced: phased products (0 row tiles only, 1 where the gathered vector exceeds 4 MiB,
        2 at any size), set before the problem is loaded."""
        L = lib()
        self._h = ctypes.c_void_p()
        rc = L.psh_create(device, ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError(f"psh_create failed ({rc})")
        self._ok(L.psh_sliced(self._h, sliced))
        self.m, self.n = scaled["m"], scaled["n"]
        self._keep = dict(scaled)
        s = self._keep
        p = Problem(self.m, self.n, int(s["a_ptr"][-1]), _p(s["a_ptr"], ctypes.c_int), _p(s["a_col"], ctypes.c_int),
                    _p(s["a_val"]), _p(s["a_val_orig"]), _p(s["at_ptr"], ctypes.c_int), _p(s["at_col"], ctypes.c_int),
                    _p(s["ato_col"], ctypes.c_int), _p(s["at_val"]), _p(s["ato_val"]), _p(s["c"]), _p(s["l"]),
                    _p(s["u"]), _p(s["rl"]), _p(s["ru"]), _p(s["row_scale"]), _p(s["col_scale"]),
                    float(s["bound_scale"]), float(s["obj_scale"]), _p(s["cl_o"]), _p(s["cu_o"]), _p(s["rl_o"]),
                    _p(s["ru_o"]), _p(s["c_min"]))
        self._ok(L.psh_load(self._h, ctypes.byref(p)))

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(f"psh error {rc}: {lib().psh_error(self._h).decode()}")

    def close(self):
        if self._h:
            lib().psh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def length(self, slot):
        return self.n if slot <= ATY else self.m

    def upload(self, slot, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._ok(lib().psh_upload(self._h, slot, _p(v)))

    def download(self, slot):
        out = np.empty(self.length(slot))
        self._ok(lib().psh_download(self._h, slot, _p(out)))
        return out

    def set_fp32(self, on):
        self._ok(lib().psh_set_fp32(self._h, int(on)))

    def fill(self, slot, a):
        self._ok(lib().psh_fill(self._h, slot, a))

    def copy(self, dst, src):
        self._ok(lib().psh_copy(self._h, dst, src))

    def axpby(self, a, x, b, y):
        self._ok(lib().psh_axpby(self._h, a, x, b, y))

    def spmv(self, x, out):
        self._ok(lib().psh_spmv(self._h, x, out))

    def spmv_t(self, y, out):
        self._ok(lib().psh_spmv_t(self._h, y, out))

    def dot(self, a, b, mode=RED_TREE):
        r = ctypes.c_double()
        self._ok(lib().psh_dot(self._h, a, b, mode, ctypes.byref(r)))
        return r.value

    def diff_norm2(self, a, b, mode=RED_TREE):
        r = ctypes.c_double()
        self._ok(lib().psh_diff_norm2(self._h, a, b, mode, ctypes.byref(r)))
        return r.value

    def primal_step(self, x, aty, tau, xh):
        self._ok(lib().psh_primal_step(self._h, x, aty, tau, xh))

    def dual_step(self, y, ax, sigma, yh):
        self._ok(lib().psh_dual_step(self._h, y, ax, sigma, yh))

    def r2h_primal(self, x, x0, aty, tau, w, rho, xh, xb):
        self._ok(lib().psh_r2h_primal(self._h, x, x0, aty, tau, w, rho, xh, xb))

    def r2h_dual(self, y, y0, ax, sigma, w, rho, yh, yb):
        self._ok(lib().psh_r2h_dual(self._h, y, y0, ax, sigma, w, rho, yh, yb))

    def kkt(self, xs, ys, mode=RED_TREE):
        k = Kkt()
        self._ok(lib().psh_kkt_of(self._h, xs, ys, mode, ctypes.byref(k)))
        return k.as_list()

    def chunk(self, iterations, inner0, tau, sigma, rho=1.0, want_hat_last=1, mode=RED_TREE):
        a = ChunkArgs(iterations, inner0, tau, sigma, rho, 0.0, 0.0, want_hat_last, mode)
        self._ok(lib().psh_chunk(self._h, ctypes.byref(a)))

    def check(self, omega, eta, want_ray=1, want_drift=1, want_pid=1, mode=RED_TREE):
        a = CheckArgs(omega, eta, want_ray, want_drift, want_pid, mode)
        s = Stats()
        self._ok(lib().psh_check(self._h, ctypes.byref(a), ctypes.byref(s)))
        return s

    def restart(self):
        self._ok(lib().psh_restart(self._h))

    def graphs(self, on):
        self._ok(lib().psh_graphs(self._h, int(on)))

    def sliced(self, on):
        self._ok(lib().psh_sliced(self._h, int(on)))

    def scale(self, geo_iterations=12, ruiz_iterations=10, pock_chambolle=1):
        """Scaling sweeps on the device (psh_scale) of a problem loaded with unscaled values
        (``unscaled_problem``); returns the row and column factors."""
        rs, cs = np.zeros(self.m), np.zeros(self.n)
        self._ok(lib().psh_scale(self._h, geo_iterations, ruiz_iterations, pock_chambolle, _p(rs), _p(cs)))
        return rs, cs

    def delta(self, on):
        self._ok(lib().psh_delta(self._h, int(on)))


def unscaled_problem(lp) -> dict:
    """The psh problem dict of `lp` with unit factors: the original values in Ã's and Ãᵀ's order
    (what psh_scale starts from) and unscaled vectors."""
    row_ptr, col, val0 = lp.csr()
    m, n = lp.m, lp.n
    order = np.lexsort((np.repeat(np.arange(m), np.diff(row_ptr)), col))
    val0 = np.asarray(val0, float)
    return scaled_problem({"a_vals": val0, "at_vals": val0[order], "c": lp.sense * np.asarray(lp.c, float),
                           "col_lower": lp.col_lower, "col_upper": lp.col_upper, "row_lower": lp.row_lower,
                           "row_upper": lp.row_upper, "row_scale": np.ones(m), "col_scale": np.ones(n),
                           "bound_scale": 1.0, "obj_scale": 1.0}, lp)


def scaled_problem(sc: dict, lp) -> dict:
    """The psh problem dict from an oracle/reference scaling result `sc` and the model `lp`."""
    row_ptr, col, val0 = lp.csr()
    m, n = lp.m, lp.n
    # Ãᵀ structure: transpose of the sorted CSR (rows of Ãᵀ sorted by row index)
    order = np.lexsort((np.repeat(np.arange(m), np.diff(row_ptr)), col))
    at_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(col, minlength=n), out=at_ptr[1:])
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(row_ptr))
    return {
        "m": m, "n": n,
        "a_ptr": row_ptr.astype(np.int32), "a_col": col.astype(np.int32), "a_val": np.asarray(sc["a_vals"], float),
        "a_val_orig": val0.astype(float),
        "at_ptr": at_ptr.astype(np.int32), "at_col": rows[order].astype(np.int32), "at_val": np.asarray(sc["at_vals"], float),
        "ato_col": np.asarray(lp.row_index, np.int32), "ato_val": np.asarray(lp.value, float),
        "c": np.asarray(sc["c"], float), "l": np.asarray(sc["col_lower"], float), "u": np.asarray(sc["col_upper"], float),
        "rl": np.asarray(sc["row_lower"], float), "ru": np.asarray(sc["row_upper"], float),
        "row_scale": np.asarray(sc["row_scale"], float), "col_scale": np.asarray(sc["col_scale"], float),
        "bound_scale": sc["bound_scale"], "obj_scale": sc["obj_scale"],
        "cl_o": np.asarray(lp.col_lower, float), "cu_o": np.asarray(lp.col_upper, float),
        "rl_o": np.asarray(lp.row_lower, float), "ru_o": np.asarray(lp.row_upper, float),
        "c_min": lp.sense * np.asarray(lp.c, float),
    }
