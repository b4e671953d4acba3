"""ctypes binding of libtdglhip.so (declared in include/tdgl_hip.h).

The library is the only compute path: importing this module does not need a
GPU, but creating a :class:`Context` does, and every failure (missing
library, no GPU, HIP error) raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libtdglhip.so"

TDGL_OK = 0
TDGL_E_INVALID = -1
TDGL_E_HIP = -2
TDGL_E_NOMEM = -3
TDGL_E_COMPLEX_PHI = -4
TDGL_E_NONFINITE = -5

PREC_F64 = 0
PREC_MIXED = 1

EXPORTS = [
    "tdgl_ctx_create", "tdgl_ctx_destroy", "tdgl_set_stream", "tdgl_last_error",
    "tdgl_set_stencil_kz", "tdgl_get_stencil_kz", "tdgl_set_sweep_tile", "tdgl_get_sweep_tile",
    "tdgl_state_len", "tdgl_n_interior", "tdgl_set_material", "tdgl_set_field",
    "tdgl_set_field_vectors", "tdgl_permute_in", "tdgl_permute_out", "tdgl_rhs",
    "tdgl_jvp_trap_fd", "tdgl_jvp_trap32", "tdgl_precond_cheb32", "tdgl_vjp_f", "tdgl_vjp_bz", "tdgl_adjoint_solve", "tdgl_dot", "tdgl_norm_inf", "tdgl_trap_step", "tdgl_newton_trap", "tdgl_tgcr_trap", "tdgl_euler_step",
    "tdgl_euler_step_noalloc", "tdgl_noalloc_max_len",
    "tdgl_observables", "tdgl_profile_enable", "tdgl_profile_read",
    "tdgl_batch_create", "tdgl_batch_destroy", "tdgl_batch_set_stream", "tdgl_batch_work_stream", "tdgl_batch_own_queue",
    "tdgl_batch_last_error",
    "tdgl_batch_total_len", "tdgl_batch_offset", "tdgl_batch_set_kz", "tdgl_batch_get_kz",
    "tdgl_batch_set_sweep_tile", "tdgl_batch_get_sweep_tile", "tdgl_batch_set_sweep_split",
    "tdgl_batch_get_sweep_split",
    "tdgl_batch_time_stencil", "tdgl_batch_set_graphs",
    "tdgl_batch_trap_step", "tdgl_batch_profile_enable", "tdgl_batch_profile_read",
]

KERNEL_CLASSES = ["pred", "newton", "jvp", "mgs", "mgs_last", "final_p", "update_r", "accum",
                  "cgs2", "cheb"]


class Grid(C.Structure):
    _fields_ = [("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
                ("periodic_x", C.c_int32), ("periodic_y", C.c_int32), ("periodic_z", C.c_int32),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("kappa", C.c_double)]


class SolverOpts(C.Structure):
    _fields_ = [("newton_tol_f", C.c_double), ("newton_tol_dx", C.c_double),
                ("newton_max_iter", C.c_int32), ("precision", C.c_int32),
                ("tol_gcr", C.c_double), ("eps_mf", C.c_double),
                ("tgcr_max_iter", C.c_int64), ("mixed_tol_floor", C.c_double),
                ("mixed_precond", C.c_int32), ("f64_precond", C.c_int32),
                ("x_prev", C.c_void_p), ("dt_prev", C.c_void_p)]


def _is_vec(t, dtype, numel: int) -> bool:
    """A contiguous device tensor of that dtype and length (what a kernel
    argument must be: a host tensor's pointer is no device address)."""
    return (getattr(t, "dtype", None) == dtype and t.numel() == numel and t.is_cuda
            and t.is_contiguous())


def _with_history(opts: "SolverOpts", x_prev, dt_prev, nfilm: int, total_len: int, torch):
    """A copy of opts whose predictor extrapolates from the last accepted step
    (x_prev: the state before it, dt_prev: its size per film); returns the copy
    and the host array that must outlive the call."""
    import numpy as np
    o = SolverOpts.from_buffer_copy(opts)
    if x_prev is None:
        return o, None
    if not _is_vec(x_prev, torch.float64, total_len):
        raise ValueError("x_prev must be a contiguous float64 device tensor shaped like the state")
    d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt_prev, dtype=np.float64), (nfilm,)))
    o.x_prev = x_prev.data_ptr()
    o.dt_prev = d.ctypes.data
    return o, d


class StepStats(C.Structure):
    _fields_ = [("converged", C.c_int32), ("newton_iters", C.c_int32),
                ("tgcr_failed", C.c_int32), ("tgcr_iters_last", C.c_int32),
                ("tgcr_iters_total", C.c_int32), ("rhs_evals", C.c_int32),
                ("jvp_evals", C.c_int32), ("reserved", C.c_int32),
                ("g_inf", C.c_double), ("dx_inf", C.c_double), ("bytes_moved", C.c_double),
                ("tgcr_r0", C.c_double), ("tgcr_rk", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


_lib = None


def load() -> C.CDLL:
    """Load libtdglhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m nanowire_tdgl_amd.build` "
            "(the HIP library is the only compute path)")
    lib = C.CDLL(str(LIB_PATH))
    P, D, I32, I64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64
    sig = {
        "tdgl_ctx_create": ([C.POINTER(Grid), C.POINTER(P)], C.c_int),
        "tdgl_ctx_destroy": ([P], C.c_int),
        "tdgl_set_stream": ([P, P], C.c_int),
        "tdgl_last_error": ([P], C.c_char_p),
        "tdgl_state_len": ([P], I64),
        "tdgl_n_interior": ([P], I64),
        "tdgl_set_material": ([P, P, P, P], C.c_int),
        "tdgl_set_field": ([P, D, D, D], C.c_int),
        "tdgl_set_field_vectors": ([P, P, P, P], C.c_int),
        "tdgl_permute_in": ([P, P, P], C.c_int),
        "tdgl_permute_out": ([P, P, P], C.c_int),
        "tdgl_rhs": ([P, P, P], C.c_int),
        "tdgl_jvp_trap_fd": ([P, P, P, P, D, D, P], C.c_int),
        "tdgl_dot": ([P, P, P, C.POINTER(D)], C.c_int),
        "tdgl_norm_inf": ([P, P, C.POINTER(D)], C.c_int),
        "tdgl_trap_step": ([P, P, D, C.POINTER(SolverOpts), C.POINTER(StepStats)], C.c_int),
        "tdgl_newton_trap": ([P, P, P, D, C.POINTER(SolverOpts), C.POINTER(StepStats)], C.c_int),
        "tdgl_tgcr_trap": ([P, P, P, D, C.POINTER(SolverOpts), P, C.POINTER(C.c_int),
                            C.POINTER(StepStats)], C.c_int),
        "tdgl_euler_step": ([P, P, D], C.c_int),
        "tdgl_euler_step_noalloc": ([P, P, D], C.c_int),
        "tdgl_noalloc_max_len": ([], I64),
        "tdgl_observables": ([P, P, P, P], C.c_int),
        "tdgl_profile_enable": ([P, C.c_int], C.c_int),
        "tdgl_profile_read": ([P, C.c_int, C.POINTER(D), C.POINTER(I64), C.POINTER(D)], C.c_int),
        "tdgl_jvp_trap32": ([P, P, P, D, P], C.c_int),
        "tdgl_precond_cheb32": ([P, P, P, D, C.c_int, P, P, P], C.c_int),
        "tdgl_set_stencil_kz": ([P, C.c_int], C.c_int),
        "tdgl_set_sweep_tile": ([P, C.c_int], C.c_int),
        "tdgl_get_sweep_tile": ([P], C.c_int),
        "tdgl_batch_set_sweep_tile": ([P, C.c_int], C.c_int),
        "tdgl_batch_get_sweep_tile": ([P], C.c_int),
        "tdgl_batch_set_sweep_split": ([P, C.c_int, C.c_int], C.c_int),
        "tdgl_batch_get_sweep_split": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "tdgl_vjp_f": ([P, P, P, P, P, P, P], C.c_int),
        "tdgl_vjp_bz": ([P, P, P, P], C.c_int),
        "tdgl_adjoint_solve": ([P, P, P, D, D, I64, C.c_int, P, C.POINTER(C.c_int)], C.c_int),
        "tdgl_get_stencil_kz": ([P], C.c_int),
        "tdgl_batch_create": ([P, C.c_int, C.POINTER(P)], C.c_int),
        "tdgl_batch_destroy": ([P], C.c_int),
        "tdgl_batch_set_stream": ([P, P], C.c_int),
        "tdgl_batch_work_stream": ([P, P], C.c_int),
        "tdgl_batch_own_queue": ([P], C.c_int),
        "tdgl_batch_last_error": ([P], C.c_char_p),
        "tdgl_batch_total_len": ([P], I64),
        "tdgl_batch_offset": ([P, C.c_int], I64),
        "tdgl_batch_set_kz": ([P, C.c_int, C.c_int], C.c_int),
        "tdgl_batch_get_kz": ([P, C.c_int], C.c_int),
        "tdgl_batch_time_stencil": ([P, C.c_int, C.POINTER(D)], C.c_int),
        "tdgl_batch_set_graphs": ([P, C.c_int], C.c_int),
        "tdgl_batch_trap_step": ([P, P, P, P, C.POINTER(SolverOpts), P], C.c_int),
        "tdgl_batch_profile_enable": ([P, C.c_int], C.c_int),
        "tdgl_batch_profile_read": ([P, C.c_int, C.POINTER(D), C.POINTER(I64), C.POINTER(D)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


_digest = None


def library_digest() -> str:
    """Digest of the built library (tuning results are only valid for one build)."""
    global _digest
    if _digest is None:
        import hashlib
        _digest = hashlib.sha256(LIB_PATH.read_bytes()).hexdigest()[:16]
    return _digest


def exported_symbols() -> list[str]:
    """Names from include/tdgl_hip.h that the built library exports."""
    lib = C.CDLL(str(LIB_PATH))
    return [s for s in EXPORTS if hasattr(lib, s)]


class TDGLError(RuntimeError):
    pass


def _raise(lib, ctx, rc: int, what: str) -> None:
    msg = lib.tdgl_last_error(ctx).decode() if ctx else ""
    text = f"{what} failed (code {rc}){': ' + msg if msg else ''}"
    if rc in (TDGL_E_INVALID, TDGL_E_COMPLEX_PHI):
        raise ValueError(text)
    raise TDGLError(text)


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


class Context:
    """One tdgl_ctx on the current device and torch's current stream."""

    def __init__(self, Nx, Ny, Nz, hx, hy, hz, kappa, periodic=(False, False, False)):
        import torch
        if not torch.cuda.is_available():
            raise TDGLError("no GPU visible: nanowire_tdgl_amd runs on MI355X (HIP) only")
        self.lib = load()
        self.torch = torch
        g = Grid(int(Nx), int(Ny), int(Nz), int(periodic[0]), int(periodic[1]), int(periodic[2]),
                 float(hx), float(hy), float(hz), float(kappa))
        h = C.c_void_p()
        rc = self.lib.tdgl_ctx_create(C.byref(g), C.byref(h))
        if rc != TDGL_OK:
            _raise(self.lib, None, rc, "tdgl_ctx_create")
        self.h = h
        self.grid = g
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = torch.cuda.current_stream(self.device)
        self._call("tdgl_set_stream", C.c_void_p(self.stream.cuda_stream))
        self.len = int(self.lib.tdgl_state_len(h))
        self.n = int(self.lib.tdgl_n_interior(h))
        self.nphi = 3 if Nz > 1 else 2

    def _call(self, name: str, *args) -> None:
        rc = getattr(self.lib, name)(self.h, *args)
        if rc != TDGL_OK:
            _raise(self.lib, self.h, rc, name)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.tdgl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stencil_kz(self, kz: int) -> None:
        self._call("tdgl_set_stencil_kz", C.c_int(int(kz)))

    def get_stencil_kz(self) -> int:
        return int(self.lib.tdgl_get_stencil_kz(self.h))

    def set_sweep_tile(self, shape: int) -> None:
        """Tile shape of the mixed-precision sweeps: 0 = 64x4, 1 = 32x8, 2 = 16x16,
        3 = the film's own best (fewest idle lanes and staged cells)."""
        self._call("tdgl_set_sweep_tile", C.c_int(int(shape)))

    def get_sweep_tile(self) -> int:
        return int(self.lib.tdgl_get_sweep_tile(self.h))

    # -- buffers ---------------------------------------------------------
    def empty_state(self):
        return self.torch.empty(self.len, dtype=self.torch.float64, device=self.device)

    def to_internal(self, ref_state):
        """complex128 reference-layout state (numpy or torch) -> internal FP64 buffer."""
        t = self.torch
        x = t.as_tensor(ref_state)
        if x.dtype != t.complex128:
            x = x.to(t.complex128)
        x = x.to(self.device).contiguous()
        if x.numel() != (1 + self.nphi) * self.n:
            raise ValueError(f"state length {x.numel()} != n_state {(1 + self.nphi) * self.n}")
        out = self.empty_state()
        self._call("tdgl_permute_in", C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()))
        return out

    def to_reference(self, st):
        """internal FP64 buffer -> complex128 reference-layout torch tensor (device)."""
        t = self.torch
        self._vecs64("to_reference", st)
        out = t.empty((1 + self.nphi) * self.n, dtype=t.complex128, device=self.device)
        self._call("tdgl_permute_out", C.c_void_p(st.data_ptr()), C.c_void_p(out.data_ptr()))
        return out

    # -- configuration -----------------------------------------------------
    def set_material(self, link_bits=None, kappa_int=None, sc_int=None) -> None:
        import numpy as np
        keep = []

        def arr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return C.c_void_p(a.ctypes.data)
        self._call("tdgl_set_material", arr(link_bits, np.uint8), arr(kappa_int, np.float64),
                   arr(sc_int, np.float64))

    def set_field(self, bx: float, by: float, bz: float) -> None:
        self._call("tdgl_set_field", C.c_double(bx), C.c_double(by), C.c_double(bz))

    def set_field_vectors(self, Bx, By, Bz) -> None:
        import numpy as np
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (Bx, By, Bz)]
        self._call("tdgl_set_field_vectors", *[C.c_void_p(v.ctypes.data) for v in a])

    # -- compute -------------------------------------------------------------
    def _vecs64(self, what: str, *vs) -> None:
        for v in vs:
            if not _is_vec(v, self.torch.float64, self.len):
                raise ValueError(f"{what} takes float64 device vectors of the internal state length")

    def rhs(self, x, f) -> None:
        self._vecs64("rhs", x, f)
        self._call("tdgl_rhs", C.c_void_p(x.data_ptr()), C.c_void_p(f.data_ptr()))

    def jvp_trap_fd(self, x, f0, p, eps: float, dt: float, ap) -> None:
        self._vecs64("jvp_trap_fd", x, f0, p, ap)
        self._call("tdgl_jvp_trap_fd", C.c_void_p(x.data_ptr()), C.c_void_p(f0.data_ptr()),
                   C.c_void_p(p.data_ptr()), C.c_double(eps), C.c_double(dt),
                   C.c_void_p(ap.data_ptr()))

    def jvp_trap32(self, x32, p32, dt: float, ap32) -> None:
        for v in (x32, p32, ap32):
            if not _is_vec(v, self.torch.float32, self.len):
                raise ValueError("jvp_trap32 takes float32 device vectors of the internal state length")
        self._call("tdgl_jvp_trap32", C.c_void_p(x32.data_ptr()), C.c_void_p(p32.data_ptr()),
                   C.c_double(dt), C.c_void_p(ap32.data_ptr()))

    def precond_cheb32(self, x32, v32, dt: float, degree: int, w32, z32) -> dict:
        """One application of the mixed TGCR's Chebyshev preconditioner:
        w32 = A M v, z32 = M v (A = I - dt/2 J(x32)).  Returns the interval
        (a, b), the degree d, and whether z32 was left unwritten (d = 1: z = v/theta)."""
        for v in (x32, v32, w32, z32):
            if not _is_vec(v, self.torch.float32, self.len):
                raise ValueError("precond_cheb32 takes float32 device vectors of the internal state length")
        info = (C.c_double * 4)()
        self._call("tdgl_precond_cheb32", C.c_void_p(x32.data_ptr()), C.c_void_p(v32.data_ptr()),
                   C.c_double(dt), C.c_int(int(degree)), C.c_void_p(w32.data_ptr()),
                   C.c_void_p(z32.data_ptr()), C.cast(info, C.c_void_p))
        return {"a": info[0], "b": info[1], "degree": int(info[2]), "z_is_v_over_theta": info[3] == 1.0}

    def vjp_f(self, x, w, gx, field=False, kappa=False, dsc=None):
        """gx = J(x)^T w; returns (d/d(Bx,By,Bz) or None, d/dkappa or None);
        dsc (optional) receives d<w, F>/d sc per interior node (n values)."""
        self._vecs64("vjp_f", x, w, gx)
        if dsc is not None and not _is_vec(dsc, self.torch.float64, self.n):
            raise ValueError("vjp_f: dsc must be a contiguous float64 device tensor of n_interior values")
        df = (C.c_double * 3)() if field else None
        dk = C.c_double() if kappa else None
        self._call("tdgl_vjp_f", C.c_void_p(x.data_ptr()), C.c_void_p(w.data_ptr()),
                   C.c_void_p(gx.data_ptr()), C.cast(df, C.c_void_p) if field else None,
                   C.byref(dk) if kappa else None, C.c_void_p(_ptr(dsc)) if dsc is not None else None)
        return (tuple(df) if field else None), (dk.value if kappa else None)

    def vjp_bz(self, gbz, gx, field=False):
        """gx = (dB_z/dx)^T gbz (gbz in reference order); returns d/d(Bx,By,Bz) or None."""
        if not _is_vec(gbz, self.torch.float64, self.n):
            raise ValueError("vjp_bz: gbz must be a contiguous float64 device tensor of n_interior values")
        self._vecs64("vjp_bz", gx)
        df = (C.c_double * 3)() if field else None
        self._call("tdgl_vjp_bz", C.c_void_p(gbz.data_ptr()), C.c_void_p(gx.data_ptr()),
                   C.cast(df, C.c_void_p) if field else None)
        return tuple(df) if field else None

    def adjoint_solve(self, x1, rhs, hdt: float, tol: float, max_iter: int = 0, out=None,
                      precond="chebyshev"):
        """mu with (I - hdt J(x1)^T) mu = rhs (TGCR, ||residual|| <= tol ||rhs||),
        right-preconditioned as `precond` says (as for the mixed method:
        "chebyshev" automatic degree, "none", or a degree); returns (mu, iterations)."""
        for v in (x1, rhs):
            if not _is_vec(v, self.torch.float64, self.len):
                raise ValueError("adjoint_solve takes float64 device vectors of the internal state length")
        out = self.torch.empty_like(rhs) if out is None else out
        self._vecs64("adjoint_solve", out)
        it = C.c_int(0)
        self._call("tdgl_adjoint_solve", C.c_void_p(x1.data_ptr()), C.c_void_p(rhs.data_ptr()),
                   C.c_double(hdt), C.c_double(tol), C.c_int64(int(max_iter)),
                   C.c_int(precond_code(precond)), C.c_void_p(out.data_ptr()), C.byref(it))
        return out, int(it.value)

    def dot(self, a, b) -> float:
        self._vecs64("dot", a, b)
        out = C.c_double()
        self._call("tdgl_dot", C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), C.byref(out))
        return out.value

    def norm_inf(self, a) -> float:
        self._vecs64("norm_inf", a)
        out = C.c_double()
        self._call("tdgl_norm_inf", C.c_void_p(a.data_ptr()), C.byref(out))
        return out.value

    def trap_step(self, X, dt: float, opts: SolverOpts, x_prev=None, dt_prev: float = 0.0) -> StepStats:
        """One step attempt; with x_prev (the state before the last accepted
        step, of size dt_prev > 0) Newton starts from the extrapolation
        X + dt/dt_prev (X - x_prev) instead of the reference's X + dt f(X)."""
        if not _is_vec(X, self.torch.float64, self.len):
            raise ValueError("X must be a contiguous float64 device tensor of the internal state length")
        st = StepStats()
        o, keep = _with_history(opts, x_prev, dt_prev, 1, self.len, self.torch)
        self._call("tdgl_trap_step", C.c_void_p(X.data_ptr()), C.c_double(dt), C.byref(o),
                   C.byref(st))
        del keep
        return st

    def newton_trap(self, x, gamma, dt: float, opts: SolverOpts) -> StepStats:
        """newton_gcr_trap on the device: x (in: x0, out: the last iterate)."""
        self._vecs64("newton_trap", x, gamma)
        st = StepStats()
        self._call("tdgl_newton_trap", C.c_void_p(x.data_ptr()), C.c_void_p(gamma.data_ptr()),
                   C.c_double(dt), C.byref(opts), C.byref(st))
        return st

    def tgcr_trap(self, x_lin, b, dt: float, opts: SolverOpts, dx) -> tuple:
        """tgcr_matrix_free_trap on the device: dx (written); returns
        (converged, stats)."""
        self._vecs64("tgcr_trap", x_lin, b, dx)
        st = StepStats()
        ok = C.c_int(0)
        self._call("tdgl_tgcr_trap", C.c_void_p(x_lin.data_ptr()), C.c_void_p(b.data_ptr()),
                   C.c_double(dt), C.byref(opts), C.c_void_p(dx.data_ptr()), C.byref(ok),
                   C.byref(st))
        return bool(ok.value), st

    def euler_step(self, X, dt: float) -> None:
        self._vecs64("euler_step", X)
        self._call("tdgl_euler_step", C.c_void_p(X.data_ptr()), C.c_double(dt))

    def euler_step_noalloc(self, X, dt: float) -> None:
        """euler_step staged in the library's per-GPU scratch array instead of
        a work buffer: allocates nothing (for graph capture and tight loops).
        States up to noalloc_max_len() doubles; overlapping calls on one GPU race."""
        self._vecs64("euler_step_noalloc", X)
        self._call("tdgl_euler_step_noalloc", C.c_void_p(X.data_ptr()), C.c_double(dt))

    def noalloc_max_len(self) -> int:
        return int(self.lib.tdgl_noalloc_max_len())

    def profile_enable(self, on: bool = True) -> None:
        self._call("tdgl_profile_enable", C.c_int(1 if on else 0))

    def profile_read(self) -> dict:
        """Per kernel class: summed event-timed milliseconds, launches, algorithmic bytes."""
        out = {}
        for k, name in enumerate(KERNEL_CLASSES):
            ms, n, by = C.c_double(), C.c_int64(), C.c_double()
            self._call("tdgl_profile_read", C.c_int(k), C.byref(ms), C.byref(n), C.byref(by))
            out[name] = {"ms": ms.value, "launches": n.value, "bytes": by.value}
        return out

    def observables(self, x, want_b=True, want_j=True):
        self._vecs64("observables", x)
        t = self.torch
        b = t.empty(3 * self.n, dtype=t.float64, device=self.device) if want_b else None
        j = t.empty(3 * self.n, dtype=t.float64, device=self.device) if want_j else None
        self._call("tdgl_observables", C.c_void_p(x.data_ptr()), C.c_void_p(_ptr(b)),
                   C.c_void_p(_ptr(j)))
        return b, j


class Batch:
    """A tdgl_batch over several Contexts (mixed-precision lockstep step).

    The contexts must stay alive as long as the batch; their fields and
    materials are read at every step.  A batched state is one float64 device
    tensor of ``total_len`` elements; film f is ``X[offset(f) : offset(f) +
    ctx.len]`` in the internal layout.
    """

    def __init__(self, contexts, stream=None):
        """``stream``: the torch stream the batch orders its work after and
        before (default: the first context's)."""
        import torch
        self.lib = load()
        self.torch = torch
        self.contexts = list(contexts)
        if not self.contexts:
            raise ValueError("a batch needs at least one film")
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        rc = self.lib.tdgl_batch_create(arr, len(self.contexts), C.byref(h))
        if rc != TDGL_OK:
            raise ValueError(f"tdgl_batch_create failed (code {rc}): films must all be 3-D or all 2-D")
        self.h = h
        self.device = self.contexts[0].device
        self.stream = self.contexts[0].stream if stream is None else stream
        self._call("tdgl_batch_set_stream", C.c_void_p(self.stream.cuda_stream))

    def _call(self, name: str, *args) -> None:
        rc = getattr(self.lib, name)(self.h, *args)
        if rc != TDGL_OK:
            msg = self.lib.tdgl_batch_last_error(self.h).decode() if self.h else ""
            text = f"{name} failed (code {rc}){': ' + msg if msg else ''}"
            if rc == TDGL_E_INVALID:
                raise ValueError(text)
            raise TDGLError(text)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.tdgl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nfilm(self) -> int:
        return len(self.contexts)

    @property
    def total_len(self) -> int:
        return int(self.lib.tdgl_batch_total_len(self.h))

    def offset(self, f: int) -> int:
        return int(self.lib.tdgl_batch_offset(self.h, int(f)))

    def set_kz(self, f: int, kz: int) -> None:
        self._call("tdgl_batch_set_kz", C.c_int(int(f)), C.c_int(int(kz)))

    def get_kz(self, f: int) -> int:
        return int(self.lib.tdgl_batch_get_kz(self.h, int(f)))

    def set_tile(self, shape: int) -> None:
        self._call("tdgl_batch_set_sweep_tile", C.c_int(int(shape)))

    def get_tile(self) -> int:
        return int(self.lib.tdgl_batch_get_sweep_tile(self.h))

    def set_split(self, max_planes: int, min_planes: int) -> None:
        """Depth range of the FP32 sweeps' pieces (same bits for every split)."""
        self._call("tdgl_batch_set_sweep_split", C.c_int(int(max_planes)), C.c_int(int(min_planes)))

    def get_split(self) -> tuple:
        a, b = C.c_int(), C.c_int()
        self._call("tdgl_batch_get_sweep_split", C.byref(a), C.byref(b))
        return a.value, b.value

    def set_stream(self, stream) -> None:
        """Order the batch's steps after and before ``stream`` (a torch stream)."""
        self.stream = stream
        self._call("tdgl_batch_set_stream", C.c_void_p(stream.cuda_stream))

    def own_queue(self) -> None:
        """Move the batch's work to a stream with a hardware queue of its own
        (tdgl_batch_own_queue); take :meth:`work_stream` again afterwards."""
        self._call("tdgl_batch_own_queue")

    def work_stream(self):
        """The batch's own stream (its kernels run there), as a torch stream;
        valid while the batch lives and until :meth:`own_queue` replaces it."""
        p = C.c_void_p()
        self._call("tdgl_batch_work_stream", C.byref(p))
        return self.torch.cuda.ExternalStream(p.value, device=self.device)

    def set_graphs(self, on: bool) -> None:
        self._call("tdgl_batch_set_graphs", C.c_int(1 if on else 0))

    def time_stencil(self, reps: int = 5) -> float:
        ms = C.c_double()
        self._call("tdgl_batch_time_stencil", C.c_int(int(reps)), C.byref(ms))
        return ms.value

    def empty_state(self):
        return self.torch.zeros(self.total_len, dtype=self.torch.float64, device=self.device)

    def pack(self, states):
        """Per-film internal-layout FP64 tensors -> one batched state."""
        X = self.empty_state()
        for f, (c, s) in enumerate(zip(self.contexts, states)):
            o = self.offset(f)
            X[o:o + c.len].copy_(s)
        return X

    def view(self, X, f: int):
        o = self.offset(f)
        return X[o:o + self.contexts[f].len]

    def trap_step(self, X, dt, opts: SolverOpts, active=None, x_prev=None, dt_prev=None) -> list:
        """One step attempt per active film; x_prev / dt_prev (per film, <= 0:
        none) as Context.trap_step's, packed like X."""
        import numpy as np
        F = self.nfilm
        if not _is_vec(X, self.torch.float64, self.total_len):
            raise ValueError("X must be a contiguous float64 device tensor of total_len elements")
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (F,)))
        a = None if active is None else np.ascontiguousarray(np.asarray(active, dtype=np.int32))
        o, keep = _with_history(opts, x_prev, 0.0 if dt_prev is None else dt_prev, F, self.total_len,
                                self.torch)
        stats = (StepStats * F)()
        self._call("tdgl_batch_trap_step", C.c_void_p(X.data_ptr()), C.c_void_p(d.ctypes.data),
                   C.c_void_p(a.ctypes.data) if a is not None else None, C.byref(o),
                   C.cast(stats, C.c_void_p))
        del keep
        return list(stats)

    def profile_enable(self, on: bool = True) -> None:
        self._call("tdgl_batch_profile_enable", C.c_int(1 if on else 0))

    def profile_read(self) -> dict:
        out = {}
        for k, name in enumerate(KERNEL_CLASSES):
            ms, n, by = C.c_double(), C.c_int64(), C.c_double()
            self._call("tdgl_batch_profile_read", C.c_int(k), C.byref(ms), C.byref(n), C.byref(by))
            out[name] = {"ms": ms.value, "launches": n.value, "bytes": by.value}
        return out


def precond_code(precond) -> int:
    """Map the mixed path's `precond` option to tdgl_solver_opts.mixed_precond:
    "chebyshev" (automatic degree) -> -1, "none" / None -> 0, an int d >= 1 ->
    a Chebyshev preconditioner with d stencil sweeps per TGCR iteration."""
    if precond is None or precond == "none":
        return 0
    if precond == "chebyshev":
        return -1
    if isinstance(precond, (int, np.integer)) and not isinstance(precond, bool) and precond >= 1:
        return int(precond)
    raise ValueError(f"precond must be 'chebyshev', 'none' or a degree >= 1, not {precond!r}")


def solver_opts(newton_tol_f=1e-3, newton_tol_dx=1e-3, newton_max_iter=20, tol_gcr=1e-4,
                eps_mf=1e-4, tgcr_max_iter=None, precision=PREC_F64,
                mixed_tol_floor=1e-6, precond="chebyshev", f64_precond="none") -> SolverOpts:
    """``precond``: the mixed method's TGCR preconditioner; ``f64_precond``:
    the FP64 method's ("none", the default, is the reference's TGCR;
    "chebyshev" or a degree: tdgl_solver_opts.f64_precond)."""
    if isinstance(precision, str):
        precision = {"f64": PREC_F64, "fp64": PREC_F64, "mixed": PREC_MIXED}[precision]
    return SolverOpts(float(newton_tol_f), float(newton_tol_dx), int(newton_max_iter),
                      int(precision), float(tol_gcr), float(eps_mf),
                      int(tgcr_max_iter) if tgcr_max_iter is not None else 0,
                      float(mixed_tol_floor), precond_code(precond), precond_code(f64_precond))

