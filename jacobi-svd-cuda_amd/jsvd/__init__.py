"""Python host binding of libjsvd.so (ctypes over the C ABI declared in include/jsvd.h).

This mirrors the reference's C++ entry points for scripting, tests and benchmarks:

* ``one_sided_jacobi_svd(a, config)`` -- jacobi::svd::one_sided_jacobi_svd
  (include/kernels.hpp:295-298): host row-major float64 in, ``JacobiSvdResult`` out.
* ``Plan`` -- a reusable device workspace for one shape (device-resident solves on torch tensors).
* kernel-level wrappers (``transpose``, ``pair_gram``, ``round_``, ``rotation_params``, ``offnorm``,
  ``u_sigma``) operating on torch device tensors, enqueued on torch's current stream.

The library is required: importing works without it, but every call raises ``RuntimeError`` when
lib/libjsvd.so has not been built.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import re
from typing import Optional

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libjsvd.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG), "include", "jsvd.h")

JSVD_OK = 0
JSVD_ERR_INVALID_ARGUMENT = 1
JSVD_ERR_DEVICE = 2
JSVD_ERR_INTERNAL = 3

LAYOUT_AUTO, LAYOUT_FORCE_ENABLE, LAYOUT_FORCE_DISABLE = 0, 1, 2


class CudaError(RuntimeError):
    """Device failure (the reference's jacobi::svd::CudaError, include/kernels.hpp:65-73)."""


class JsvdConfig(ctypes.Structure):
    """jsvd_config, field-for-field JacobiSvdConfig (include/kernels.hpp:78-124)."""
    _fields_ = [
        ("epsilon", ctypes.c_double),
        ("max_sweeps", ctypes.c_int32),
        ("threads_per_block", ctypes.c_int32),
        ("layout_transpose_mode", ctypes.c_int32),
        ("layout_transpose_min_columns", ctypes.c_int32),
        ("layout_transpose_min_elements", ctypes.c_uint64),
        ("layout_transpose_auto_tune", ctypes.c_int32),
        ("layout_transpose_benchmark_repetitions", ctypes.c_int32),
        ("layout_transpose_benchmark_sweeps", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


PATH_PLAIN_F64, PATH_BLOCK_F64, PATH_MIXED = 0, 1, 2
PATH_NAMES = {PATH_PLAIN_F64: "plain_f64", PATH_BLOCK_F64: "block_f64", PATH_MIXED: "mixed_f32_f64"}


class MixedConfig(ctypes.Structure):
    """jsvd_mixed_config (include/jsvd.h): solver path and FP32-phase knobs."""
    _fields_ = [
        ("path", ctypes.c_int32),
        ("fp32_max_sweeps", ctypes.c_int32),
        ("fp32_tolerance", ctypes.c_double),
        ("orthogonalize_steps", ctypes.c_int32),
        ("block_width", ctypes.c_int32),
        ("qr_precondition", ctypes.c_int32),
        ("inner_order", ctypes.c_int32),
        ("fp32_stop_fraction", ctypes.c_double),
    ]


INNER_CROSS, INNER_FULL = 0, 1


class SolveInfo(ctypes.Structure):
    """jsvd_solve_info (include/jsvd.h): per-solve report."""
    _fields_ = [
        ("path", ctypes.c_int32),
        ("sweeps_fp32", ctypes.c_int32),
        ("sweeps_fp64", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("offmax_fp32", ctypes.c_double),
        ("offmax_fp64", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def mixed_config(path: int = PATH_MIXED, fp32_max_sweeps: Optional[int] = None,
                 fp32_tolerance: Optional[float] = None, orthogonalize_steps: Optional[int] = None,
                 block_width: Optional[int] = None, qr_precondition: Optional[int] = None,
                 inner_order: Optional[int] = None, fp32_stop_fraction: Optional[float] = None) -> MixedConfig:
    mx = MixedConfig()
    lib().jsvd_mixed_config_default(ctypes.byref(mx))
    mx.path = path
    if inner_order is not None:
        mx.inner_order = inner_order
    if qr_precondition is not None:
        # 0 = sweep A, 1 = one QR (sweep R^T), 2 = two QRs (sweep R2^T), 3 = two QRs when n >= 512 and R's
        # diagonal spread is >= 100, else one; a bool selects 0 or 1
        mx.qr_precondition = int(qr_precondition)
    if block_width is not None:
        mx.block_width = block_width
    if fp32_max_sweeps is not None:
        mx.fp32_max_sweeps = fp32_max_sweeps
    if fp32_tolerance is not None:
        mx.fp32_tolerance = fp32_tolerance
    if orthogonalize_steps is not None:
        mx.orthogonalize_steps = orthogonalize_steps
    if fp32_stop_fraction is not None:
        mx.fp32_stop_fraction = fp32_stop_fraction
    return mx


@dataclasses.dataclass
class JacobiSvdConfig:
    """Python mirror of jacobi::svd::JacobiSvdConfig with the reference defaults (kernels.hpp:83-123)."""
    epsilon: float = 1.0e-9
    max_sweeps: int = 128
    threads_per_block: int = 256
    layout_transpose_mode: int = LAYOUT_AUTO
    layout_transpose_min_columns: int = 16
    layout_transpose_min_elements: int = 4096
    layout_transpose_auto_tune: bool = False
    layout_transpose_benchmark_repetitions: int = 2
    layout_transpose_benchmark_sweeps: int = 8

    def to_c(self) -> JsvdConfig:
        c = JsvdConfig()
        c.epsilon = self.epsilon
        c.max_sweeps = self.max_sweeps
        c.threads_per_block = self.threads_per_block
        c.layout_transpose_mode = self.layout_transpose_mode
        c.layout_transpose_min_columns = self.layout_transpose_min_columns
        c.layout_transpose_min_elements = self.layout_transpose_min_elements
        c.layout_transpose_auto_tune = int(bool(self.layout_transpose_auto_tune))
        c.layout_transpose_benchmark_repetitions = self.layout_transpose_benchmark_repetitions
        c.layout_transpose_benchmark_sweeps = self.layout_transpose_benchmark_sweeps
        return c


@dataclasses.dataclass
class JacobiSvdResult:
    """jacobi::svd::JacobiSvdResult (include/kernels.hpp:249-280): row-major U (m x n), sigma (n), V (n x n)."""
    rows: int
    columns: int
    u: np.ndarray
    sigma: np.ndarray
    v: np.ndarray
    sweeps: int


_lib = None
_dptr = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64


def library_path() -> str:
    return LIB_PATH


def declared_symbols() -> list[str]:
    """Every function name declared in include/jsvd.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(jsvd_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libjsvd.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.jsvd_config_default.argtypes = [ctypes.POINTER(JsvdConfig)]
        L.jsvd_config_default.restype = None
        L.jsvd_config_validate.argtypes = [ctypes.POINTER(JsvdConfig)]
        L.jsvd_last_error.restype = ctypes.c_char_p
        L.jsvd_version.restype = ctypes.c_char_p
        L.jsvd_plain_f64.argtypes = [_vp, _i64, _i64, ctypes.POINTER(JsvdConfig), _vp, _vp, _vp, _i32p]
        L.jsvd_plan_create.argtypes = [_i64, _i64, ctypes.POINTER(JsvdConfig), _vp, ctypes.POINTER(_vp)]
        L.jsvd_plan_destroy.argtypes = [_vp]
        L.jsvd_plan_solve_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _i32p]
        L.jsvd_plan_solve_host.argtypes = [_vp, _vp, _vp, _vp, _vp, _i32p]
        L.jsvd_plan_sweep_stats.argtypes = [_vp, _i32p, _dptr, ctypes.c_int32]
        L.jsvd_transpose_f64.argtypes = [_vp, _i64, _vp, _i64, _i64, _i64, _vp]
        L.jsvd_pair_gram_f64.argtypes = [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _vp, _vp]
        L.jsvd_round_f64.argtypes = [_vp, _i64, _i64, _vp, _i64, _i64, ctypes.c_int32, ctypes.c_double, _vp, _vp,
                                     _vp, _vp]
        L.jsvd_rotation_params_f64.argtypes = [_vp, _vp, _vp, _i64, ctypes.c_double, _vp, _vp, _vp, _vp]
        L.jsvd_round_robin_pairs.argtypes = [_i64, ctypes.c_int32, _i32p]
        L.jsvd_mixed_config_default.argtypes = [ctypes.POINTER(MixedConfig)]
        L.jsvd_mixed_config_default.restype = None
        L.jsvd_plan_create_ex.argtypes = [_i64, _i64, ctypes.POINTER(JsvdConfig), ctypes.POINTER(MixedConfig), _vp,
                                          ctypes.POINTER(_vp)]
        L.jsvd_plan_last_info.argtypes = [_vp, ctypes.POINTER(SolveInfo)]
        L.jsvd_solve_f64.argtypes = [_vp, _i64, _i64, ctypes.POINTER(JsvdConfig), ctypes.POINTER(MixedConfig), _vp,
                                     _vp, _vp, ctypes.POINTER(SolveInfo)]
        L.jsvd_offnorm_f64.argtypes = [_vp, _i64, _i64, _i64, _vp, _vp]
        L.jsvd_u_sigma_f64.argtypes = [_vp, _i64, _i64, _i64, ctypes.c_double, _vp, _vp, _vp]
        L.jsvd_hbm_copy.argtypes = [_vp, _vp, _i64, ctypes.c_int32, _vp]
        L.jsvd_block_eigen.argtypes = [ctypes.c_int32, ctypes.c_int32, _vp, _i64, ctypes.c_double, _vp, _vp, _vp]
        L.jsvd_block_eigen_ordered.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _i64,
                                               ctypes.c_double, _vp, _vp, ctypes.c_int32, _vp]
        L.jsvd_cholesky_f64.argtypes = [_vp, _i64, _i64, _vp, _vp]
        L.jsvd_gemm_f64.argtypes = [ctypes.c_int32, ctypes.c_int32, _i64, _i64, _i64, ctypes.c_double, _vp, _i64,
                                    _vp, _i64, ctypes.c_double, _vp, _i64, _vp, _i64, ctypes.c_int32, _vp]
        L.jsvd_trsm_right_upper_f64.argtypes = [_vp, _i64, _i64, _i64, _vp, _i64, _vp]
        L.jsvd_rounds_f64.argtypes = [_vp, _i64, _i64, _vp, _i64, _i64, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_double, _vp, _vp, ctypes.c_int32, _vp, _vp]
        for name in declared_symbols():
            if name not in ("jsvd_config_default", "jsvd_last_error", "jsvd_version", "jsvd_mixed_config_default"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def last_error() -> str:
    return (lib().jsvd_last_error() or b"").decode()


def check(status: int) -> None:
    if status == JSVD_OK:
        return
    msg = last_error()
    if status == JSVD_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == JSVD_ERR_DEVICE:
        raise CudaError(msg)
    raise RuntimeError(msg)


def default_config() -> JsvdConfig:
    c = JsvdConfig()
    lib().jsvd_config_default(ctypes.byref(c))
    return c


def _as_c_config(config) -> JsvdConfig:
    if config is None:
        return default_config()
    if isinstance(config, JsvdConfig):
        return config
    return config.to_c()


def validate_config(config) -> None:
    """jsvd_config_validate: raises ValueError with the reference's message for an invalid configuration."""
    check(lib().jsvd_config_validate(ctypes.byref(_as_c_config(config))))


def one_sided_jacobi_svd(a, config: Optional[JacobiSvdConfig] = None) -> JacobiSvdResult:
    """Host-array solve (jacobi::svd::one_sided_jacobi_svd).  a: row-major float64 m x n with m >= n."""
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if arr.ndim != 2:
        raise ValueError("Input buffer size mismatch.")
    m, n = arr.shape
    u = np.zeros((m, n))
    s = np.zeros(n)
    v = np.zeros((n, n))
    sweeps = ctypes.c_int32()
    cfg = _as_c_config(config)
    check(lib().jsvd_plain_f64(arr.ctypes.data, m, n, ctypes.byref(cfg), u.ctypes.data, s.ctypes.data,
                               v.ctypes.data, ctypes.byref(sweeps)))
    return JacobiSvdResult(m, n, u, s, v, sweeps.value)


def round_robin_pairs(n: int, rnd: int) -> list[tuple[int, int]]:
    """The closed-form schedule the kernels evaluate, computed on the host (jsvd_round_robin_pairs)."""
    buf = np.zeros(2 * ((n + 1) // 2 + 1), dtype=np.int32)
    k = lib().jsvd_round_robin_pairs(n, rnd, buf.ctypes.data_as(_i32p))
    if k < 0:
        check(-k)
    return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(k)]


def solve(a, config: Optional[JacobiSvdConfig] = None, path: int = PATH_MIXED,
          mixed: Optional[MixedConfig] = None):
    """One-shot host-array solve through the selected path (jsvd_solve_f64).  Returns (result, SolveInfo)."""
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    m, n = arr.shape
    u = np.zeros((m, n))
    s = np.zeros(n)
    v = np.zeros((n, n))
    info = SolveInfo()
    cfg = _as_c_config(config)
    mx = mixed if mixed is not None else mixed_config(path)
    check(lib().jsvd_solve_f64(arr.ctypes.data, m, n, ctypes.byref(cfg), ctypes.byref(mx), u.ctypes.data,
                               s.ctypes.data, v.ctypes.data, ctypes.byref(info)))
    return JacobiSvdResult(m, n, u, s, v, info.sweeps_fp64), info


def _stream_handle(stream=None) -> int:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _ptr(t) -> int:
    return int(t.data_ptr())


class Plan:
    """Reusable solver workspace for one shape (jsvd_plan_*)."""

    def __init__(self, m: int, n: int, config=None, stream=None, path: int = PATH_PLAIN_F64,
                 mixed: Optional[MixedConfig] = None):
        self.m, self.n = int(m), int(n)
        self._cfg = _as_c_config(config)
        self._mixed = mixed if mixed is not None else mixed_config(path)
        handle = _vp()
        s = _stream_handle(stream) if stream is not None else None
        check(lib().jsvd_plan_create_ex(self.m, self.n, ctypes.byref(self._cfg), ctypes.byref(self._mixed), s,
                                        ctypes.byref(handle)))
        self._h = handle

    def last_info(self) -> SolveInfo:
        info = SolveInfo()
        check(lib().jsvd_plan_last_info(self._h, ctypes.byref(info)))
        return info

    def launch_params(self) -> dict:
        """The plan's current launch parameters (jsvd_plan_launch_params)."""
        L = lib()
        L.jsvd_plan_launch_params.argtypes = [_vp] + [ctypes.POINTER(ctypes.c_int32)] * 5
        vals = [ctypes.c_int32() for _ in range(5)]
        check(L.jsvd_plan_launch_params(self._h, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("round_threads", "round_ea", "round_ev", "apply_workgroups", "block_width"),
                        (v.value for v in vals)))

    def set_round_span(self, waves: int, max_rounds: int = 8) -> None:
        """Round spans of a plain plan: `waves` = 8 or 16 wavefronts per workgroup and up to max_rounds rounds per
        launch, or waves = 0 for one launch per round (jsvd_plan_set_round_span)."""
        L = lib()
        L.jsvd_plan_set_round_span.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32]
        check(L.jsvd_plan_set_round_span(self._h, int(waves), int(max_rounds)))

    def round_span(self) -> dict:
        """The plan's round-span setting (jsvd_plan_round_span)."""
        L = lib()
        L.jsvd_plan_round_span.argtypes = [_vp] + [ctypes.POINTER(ctypes.c_int32)] * 3
        vals = [ctypes.c_int32() for _ in range(3)]
        check(L.jsvd_plan_round_span(self._h, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("waves", "max_rounds", "launches_per_sweep"), (v.value for v in vals)))

    def reshape(self, m: int, n: int, epsilon: float) -> None:
        """Re-targets the plan at m x n and epsilon inside its allocated capacity (jsvd_plan_reshape)."""
        L = lib()
        L.jsvd_plan_reshape.argtypes = [_vp, _i64, _i64, ctypes.c_double]
        check(L.jsvd_plan_reshape(self._h, int(m), int(n), float(epsilon)))
        self.m, self.n = int(m), int(n)

    def close(self) -> None:
        if self._h:
            check(lib().jsvd_plan_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_device(self, d_a, d_u, d_sigma, d_v) -> int:
        """Device-resident solve on torch float64 tensors (row-major, contiguous).

        The plan runs on its own stream, so pending torch work on the current stream (the producers of
        d_a, allocations of the outputs) is waited for first; the call returns after the solve completed.
        """
        import torch
        for t, shape in ((d_a, (self.m, self.n)), (d_u, (self.m, self.n)), (d_sigma, (self.n,)),
                         (d_v, (self.n, self.n))):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
                raise ValueError(f"expected a contiguous cuda float64 tensor of shape {shape}")
        torch.cuda.current_stream().synchronize()
        sweeps = ctypes.c_int32()
        check(lib().jsvd_plan_solve_device(self._h, _ptr(d_a), _ptr(d_u), _ptr(d_sigma), _ptr(d_v),
                                           ctypes.byref(sweeps)))
        return sweeps.value

    def solve_host(self, a: np.ndarray) -> JacobiSvdResult:
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        u = np.zeros((self.m, self.n))
        s = np.zeros(self.n)
        v = np.zeros((self.n, self.n))
        sweeps = ctypes.c_int32()
        check(lib().jsvd_plan_solve_host(self._h, arr.ctypes.data, u.ctypes.data, s.ctypes.data, v.ctypes.data,
                                         ctypes.byref(sweeps)))
        return JacobiSvdResult(self.m, self.n, u, s, v, sweeps.value)

    def sweep_stats(self, capacity: int = 1024):
        rot = np.zeros(capacity, dtype=np.int32)
        off = np.zeros(capacity)
        k = lib().jsvd_plan_sweep_stats(self._h, rot.ctypes.data_as(_i32p), off.ctypes.data_as(_dptr), capacity)
        if k < 0:
            check(-k)
        return rot[:k].copy(), off[:k].copy()


# ------------------------------------------------------------------------------------------------
# kernel-level wrappers (torch device tensors; enqueued on the current torch stream)
# ------------------------------------------------------------------------------------------------

def transpose(src, dst, rows: int, cols: int, lds: int, ldd: int, stream=None) -> None:
    check(lib().jsvd_transpose_f64(_ptr(src), lds, _ptr(dst), ldd, rows, cols, _stream_handle(stream)))


def pair_gram(a_col, lda: int, m: int, pairs, alpha, beta, gamma, stream=None) -> None:
    check(lib().jsvd_pair_gram_f64(_ptr(a_col), lda, m, _ptr(pairs), pairs.numel() // 2, _ptr(alpha), _ptr(beta),
                                   _ptr(gamma), _stream_handle(stream)))


def round_(a_col, lda: int, m: int, v_col, ldv: int, n: int, rnd: int, eps: float, rot_count=None, offmax=None,
           skip_if_zero=None, stream=None) -> None:
    check(lib().jsvd_round_f64(_ptr(a_col), lda, m, _ptr(v_col), ldv, n, rnd, eps,
                               _ptr(rot_count) if rot_count is not None else None,
                               _ptr(offmax) if offmax is not None else None,
                               _ptr(skip_if_zero) if skip_if_zero is not None else None,
                               _stream_handle(stream)))


def rounds(a_col, lda: int, m: int, v_col, ldv: int, n: int, first_round: int, count: int, eps: float,
           rot_count=None, offmax=None, skip_if_zero=None, stream=None, count_slots: int = 1) -> None:
    """`count` consecutive fused rounds from `first_round`, launched back to back (jsvd_rounds_f64); rot_count
    and offmax hold count_slots entries each (summed / maximised by the caller)."""
    for t in (rot_count, offmax):
        if t is not None and t.numel() < count_slots:
            raise ValueError("counter tensors need count_slots entries")
    check(lib().jsvd_rounds_f64(_ptr(a_col), lda, m, _ptr(v_col), ldv, n, first_round, count, eps,
                                _ptr(rot_count) if rot_count is not None else None,
                                _ptr(offmax) if offmax is not None else None, int(count_slots),
                                _ptr(skip_if_zero) if skip_if_zero is not None else None,
                                _stream_handle(stream)))


def rotation_params(alpha, beta, gamma, eps: float, c, s, rotate, stream=None) -> None:
    check(lib().jsvd_rotation_params_f64(_ptr(alpha), _ptr(beta), _ptr(gamma), alpha.numel(), eps, _ptr(c), _ptr(s),
                                         _ptr(rotate), _stream_handle(stream)))


def offnorm(a_col, lda: int, m: int, n: int, out, stream=None) -> None:
    check(lib().jsvd_offnorm_f64(_ptr(a_col), lda, m, n, _ptr(out), _stream_handle(stream)))


def u_sigma(a_col, lda: int, m: int, n: int, eps: float, u, sigma, stream=None) -> None:
    check(lib().jsvd_u_sigma_f64(_ptr(a_col), lda, m, n, eps, _ptr(u), _ptr(sigma), _stream_handle(stream)))


def hbm_copy(dst, src, variant: int = 0, stream=None) -> None:
    """jsvd_hbm_copy: the library's streaming device-to-device copy of src into dst (equal byte sizes, a multiple of
    16 bytes); variant 1 = non-temporal stores."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise ValueError("dst and src must have the same byte size")
    check(lib().jsvd_hbm_copy(_ptr(dst), _ptr(src), nbytes, int(variant), _stream_handle(stream)))


GEMM_UPPER_OUT, GEMM_MIRROR, GEMM_TRI_A, GEMM_TRI_B, GEMM_ZERO_LOWER, GEMM_ROW_MAJOR_CD = 1, 2, 4, 8, 16, 32


def gemm(trans_a: bool, trans_b: bool, m: int, n: int, k: int, alpha: float, a, lda: int, b, ldb: int, beta: float,
         c, ldc: int, d, ldd: int, flags: int = 0, stream=None) -> None:
    """jsvd_gemm_f64 on float64 CUDA tensors (flat storage, column-major with the given leading dimensions; c may
    be None when beta == 0, or the same tensor as d)."""
    check(lib().jsvd_gemm_f64(int(bool(trans_a)), int(bool(trans_b)), m, n, k, float(alpha), _ptr(a), lda, _ptr(b),
                              ldb, float(beta), _ptr(c) if c is not None else None, ldc, _ptr(d), ldd, int(flags),
                              _stream_handle(stream)))


def trsm_right_upper(b, ldb: int, rows: int, n: int, r, ldr: int, stream=None) -> None:
    """jsvd_trsm_right_upper_f64: b <- b r^-1 on float64 CUDA tensors (column-major, flat storage)."""
    check(lib().jsvd_trsm_right_upper_f64(_ptr(b), ldb, rows, n, _ptr(r), ldr, _stream_handle(stream)))


def cholesky(a, n: int, info, stream=None) -> None:
    """A = R^T R in place (jsvd_cholesky_f64) on a float64 CUDA tensor viewed column-major: a is (cols, lda)
    row-major storage = column-major lda x cols with lda >= n rounded up to 64 and cols >= that too; info is an
    int32 CUDA tensor of one element."""
    import torch
    if a.dtype != torch.float64 or not a.is_cuda or not a.is_contiguous() or a.dim() != 2:
        raise ValueError("a must be a contiguous float64 CUDA matrix")
    lda = a.shape[1]
    n64 = (n + 63) // 64 * 64
    if a.shape[0] < n64:
        raise ValueError("a needs n rounded up to 64 columns")
    check(lib().jsvd_cholesky_f64(_ptr(a), lda, n, _ptr(info), _stream_handle(stream)))


def block_eigen(gram, eps: float, q, rotations, stream=None, inner_order: int = INNER_FULL, repeats: int = 1) -> None:
    """One Gram-space inner sweep on each kw x kw matrix of gram (pairs x kw x kw, kw = 32, 64 or 128, float32 or
    float64 CUDA tensor): q receives the accumulated rotations, rotations (int32, pairs) the rotated-pair counts.
    inner_order INNER_FULL: the full cyclic sweep; INNER_CROSS (kw = 64): the block rounds' cross ordering.
    repeats > 1 enqueues the same launch that many times back to back (for timing)."""
    import torch
    if gram.dim() != 3 or gram.shape[1] != gram.shape[2] or gram.shape[1] not in (32, 64, 128) \
            or not gram.is_contiguous():
        raise ValueError("gram must be a contiguous pairs x kw x kw tensor, kw in (32, 64, 128)")
    if q.shape != gram.shape or q.dtype != gram.dtype or not q.is_contiguous():
        raise ValueError("q must match gram")
    if rotations.dtype != torch.int32 or rotations.numel() != gram.shape[0]:
        raise ValueError("rotations must be int32 with one entry per pair")
    precision = PRECISION_F32 if gram.dtype == torch.float32 else PRECISION_F64
    if gram.dtype not in (torch.float32, torch.float64):
        raise ValueError("gram must be float32 or float64")
    check(lib().jsvd_block_eigen_ordered(precision, gram.shape[1] // 2, int(inner_order), _ptr(gram), gram.shape[0],
                                         eps, _ptr(q), _ptr(rotations), int(repeats), _stream_handle(stream)))


# ------------------------------------------------------------------------------------------------
# batched path (jsvd_batch_*)
# ------------------------------------------------------------------------------------------------
PRECISION_F64, PRECISION_F32 = 0, 1
BATCH_MIN_COLUMNS = 4  # JSVD_BATCH_MIN_COLUMNS (include/jsvd.h)

BATCH_ITEM_DTYPE = np.dtype([
    ("a_offset", "<i8"), ("u_offset", "<i8"), ("sigma_offset", "<i8"), ("v_offset", "<i8"),
    ("rows", "<i4"), ("columns", "<i4"), ("precision", "<i4"), ("max_sweeps", "<i4"), ("tolerance", "<f8"),
])
assert BATCH_ITEM_DTYPE.itemsize == 56


class BatchBuffers(ctypes.Structure):
    """jsvd_batch_buffers (include/jsvd.h)."""
    _fields_ = [
        ("a", ctypes.c_void_p), ("u", ctypes.c_void_p), ("sigma", ctypes.c_void_p), ("v", ctypes.c_void_p),
        ("sweeps", ctypes.c_void_p),
        ("a_len", ctypes.c_int64), ("u_len", ctypes.c_int64), ("sigma_len", ctypes.c_int64), ("v_len", ctypes.c_int64),
    ]


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return
    L.jsvd_batch_create.argtypes = [_i64, _i64, _vp, ctypes.POINTER(_vp)]
    L.jsvd_batch_destroy.argtypes = [_vp]
    L.jsvd_batch_solve_device.argtypes = [_vp, _vp, _i64, ctypes.POINTER(BatchBuffers)]
    L.jsvd_batch_solve_async.argtypes = [_vp, _vp, _i64, ctypes.POINTER(BatchBuffers), _vp]
    L.jsvd_batch_set_shape.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32]
    L.jsvd_batch_get_shape.argtypes = [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32)]
    L.jsvd_batch_autotune.argtypes = [_vp, _vp, _i64, ctypes.POINTER(BatchBuffers), ctypes.c_int32,
                                      ctypes.POINTER(BatchTuneReport)]
    L.jsvd_batch_solve_host.argtypes = [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]
    L.jsvd_batch_solve_noalloc.argtypes = [_vp, _i64, ctypes.POINTER(BatchBuffers), _i64, _i64, _vp]
    L._batch_bound = True


def batch_solve_noalloc(d_items, count, d_a, d_u, d_s, d_v, d_sweeps, max_rows: int, max_cols: int, stream=None):
    """jsvd_batch_solve_noalloc: the batched solve with the library's fixed device scratch (no allocation, no
    batch object).  Device tensors as for Batch.solve_device; enqueued on `stream` (a torch stream, default the
    current one) without waiting."""
    import torch
    L = lib()
    _bind_batch(L)
    buf = BatchBuffers(d_a.data_ptr(), d_u.data_ptr(), d_s.data_ptr(), d_v.data_ptr(), d_sweeps.data_ptr(),
                       d_a.numel(), d_u.numel(), d_s.numel(), d_v.numel())
    s = stream if stream is not None else torch.cuda.current_stream()
    check(L.jsvd_batch_solve_noalloc(d_items.data_ptr(), int(count), ctypes.byref(buf), int(max_rows), int(max_cols),
                                     _stream_handle(s)))


BATCH_MAX_SHAPES = 8

# (threads per workgroup, LDS bytes per workgroup) of the batched kernel's launch shapes; the first is the default
BATCH_SHAPES = ((256, 76 * 1024), (128, 38 * 1024), (256, 38 * 1024), (512, 76 * 1024), (512, 152 * 1024),
                (1024, 152 * 1024))


class BatchTuneReport(ctypes.Structure):
    """jsvd_batch_tune_report (include/jsvd.h)."""
    _fields_ = [
        ("candidates", ctypes.c_int32), ("chosen", ctypes.c_int32),
        ("threads", ctypes.c_int32 * BATCH_MAX_SHAPES), ("lds_bytes", ctypes.c_int32 * BATCH_MAX_SHAPES),
        ("workgroups", ctypes.c_int32 * BATCH_MAX_SHAPES), ("reserved", ctypes.c_int32),
        ("ms", ctypes.c_double * BATCH_MAX_SHAPES),
    ]

    def as_dict(self) -> dict:
        k = self.candidates
        return {"chosen": self.chosen,
                "shapes": [{"threads": self.threads[i], "lds_bytes": self.lds_bytes[i],
                            "workgroups": self.workgroups[i], "ms": self.ms[i]} for i in range(k)]}


class BatchPacked:
    """Host packing of a list of matrices into the batch layout (offsets in doubles)."""

    def __init__(self, mats, precisions, tolerances, max_sweeps):
        k = len(mats)
        self.items = np.zeros(k, dtype=BATCH_ITEM_DTYPE)
        a_off = u_off = s_off = v_off = 0
        for i, a in enumerate(mats):
            m, n = a.shape
            it = self.items[i]
            it["a_offset"], it["u_offset"], it["sigma_offset"], it["v_offset"] = a_off, u_off, s_off, v_off
            it["rows"], it["columns"] = m, n
            it["precision"] = precisions[i]
            it["max_sweeps"] = max_sweeps if np.isscalar(max_sweeps) else max_sweeps[i]
            it["tolerance"] = tolerances[i]
            a_off += m * n
            u_off += m * n
            s_off += n
            v_off += n * n
        self.a = np.zeros(a_off)
        for i, a in enumerate(mats):
            o = int(self.items[i]["a_offset"])
            self.a[o:o + a.size] = np.ascontiguousarray(a, dtype=np.float64).ravel()
        self.lens = (a_off, u_off, s_off, v_off)

    def check_limits(self, max_rows: int, max_cols: int):
        """The batch item contract (JSVD_BATCH_MIN_COLUMNS <= n <= max_cols, n <= m <= max_rows); ValueError
        naming the first item outside it."""
        for i, it in enumerate(self.items):
            m, n = int(it["rows"]), int(it["columns"])
            if not (BATCH_MIN_COLUMNS <= n <= max_cols and n <= m <= max_rows):
                raise ValueError(f"batch item {i}: needs {BATCH_MIN_COLUMNS} <= columns <= rows within the batch "
                                 f"limits ({max_rows} x {max_cols}), got {m} x {n}")

    def cost_order(self):
        """Queue order: most expensive first (m n^2), stable."""
        cost = self.items["rows"].astype(np.float64) * self.items["columns"].astype(np.float64) ** 2
        return np.argsort(-cost, kind="stable")


class Batch:
    """Persistent batched solver (jsvd_batch_*) for matrices up to max_rows x max_cols."""

    def __init__(self, max_rows: int, max_cols: int, stream=None):
        L = lib()
        _bind_batch(L)
        h = _vp()
        s = _stream_handle(stream) if stream is not None else None
        check(L.jsvd_batch_create(int(max_rows), int(max_cols), s, ctypes.byref(h)))
        self._h = h
        self.max_rows, self.max_cols = int(max_rows), int(max_cols)

    def close(self):
        if self._h:
            check(lib().jsvd_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_device(self, d_items, count, d_a, d_u, d_s, d_v, d_sweeps):
        """All arguments are torch device tensors (items as uint8/int64 storage of BATCH_ITEM_DTYPE)."""
        import torch
        torch.cuda.current_stream().synchronize()
        buf = BatchBuffers(d_a.data_ptr(), d_u.data_ptr(), d_s.data_ptr(), d_v.data_ptr(), d_sweeps.data_ptr(),
                           d_a.numel(), d_u.numel(), d_s.numel(), d_v.numel())
        check(lib().jsvd_batch_solve_device(self._h, d_items.data_ptr(), int(count), ctypes.byref(buf)))

    def solve_host_packed(self, pk: "BatchPacked", order=None):
        """Host-pointer solve (jsvd_batch_solve_host) of a packed batch; items in `order` (default: as packed).
        Returns (u, sigma, v, sweeps) as flat host arrays (sweeps in queue order)."""
        items = pk.items if order is None else pk.items[order]
        items = np.ascontiguousarray(items)
        a = np.ascontiguousarray(pk.a, dtype=np.float64)
        u = np.zeros(max(pk.lens[1], 1))
        s = np.zeros(max(pk.lens[2], 1))
        v = np.zeros(max(pk.lens[3], 1))
        sw = np.zeros(len(items), dtype=np.int32)
        check(lib().jsvd_batch_solve_host(self._h, items.ctypes.data, len(items), a.ctypes.data, a.size,
                                          u.ctypes.data, u.size, s.ctypes.data, s.size, v.ctypes.data, v.size,
                                          sw.ctypes.data))
        return u, s, v, sw

    def set_shape(self, threads: int = 0, lds_bytes: int = 0):
        """Launch shape (threads per workgroup, LDS bytes per workgroup), one of BATCH_SHAPES; (0, 0) = default."""
        check(lib().jsvd_batch_set_shape(self._h, int(threads), int(lds_bytes)))

    def shape(self):
        """(threads, lds_bytes, resident workgroups) of the current launch shape."""
        t, l, w = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().jsvd_batch_get_shape(self._h, ctypes.byref(t), ctypes.byref(l), ctypes.byref(w)))
        return t.value, l.value, w.value

    def autotune(self, d_items, count, d_a, d_u, d_s, d_v, d_sweeps, repetitions: int = 1) -> BatchTuneReport:
        """Times every launch shape on this batch (device tensors as for solve_device, which it also fulfils),
        keeps the fastest and records it for batches with the same limits."""
        import torch
        torch.cuda.current_stream().synchronize()
        buf = BatchBuffers(d_a.data_ptr(), d_u.data_ptr(), d_s.data_ptr(), d_v.data_ptr(), d_sweeps.data_ptr(),
                           d_a.numel(), d_u.numel(), d_s.numel(), d_v.numel())
        rep = BatchTuneReport()
        check(lib().jsvd_batch_autotune(self._h, d_items.data_ptr(), int(count), ctypes.byref(buf),
                                        int(repetitions), ctypes.byref(rep)))
        return rep

    def solve(self, mats, precisions, tolerances, max_sweeps=60, order=None):
        """Host convenience: pack, upload, solve in one launch, download.  Returns (u list, s list, v list,
        sweeps array) in input order."""
        import torch
        pk = BatchPacked(mats, precisions, tolerances, max_sweeps)
        pk.check_limits(self.max_rows, self.max_cols)
        if order is None:
            order = pk.cost_order()
        items = pk.items[order]
        d_items = torch.from_numpy(items.view(np.uint8).copy()).to("cuda")
        d_a = torch.from_numpy(pk.a).to("cuda")
        d_u = torch.zeros(max(pk.lens[1], 1), dtype=torch.float64, device="cuda")
        d_s = torch.zeros(max(pk.lens[2], 1), dtype=torch.float64, device="cuda")
        d_v = torch.zeros(max(pk.lens[3], 1), dtype=torch.float64, device="cuda")
        d_sw = torch.zeros(len(mats), dtype=torch.int32, device="cuda")
        self.solve_device(d_items, len(mats), d_a, d_u, d_s, d_v, d_sw)
        u, s, v = d_u.cpu().numpy(), d_s.cpu().numpy(), d_v.cpu().numpy()
        sw_q = d_sw.cpu().numpy()
        sweeps = np.zeros(len(mats), dtype=np.int32)
        sweeps[order] = sw_q
        us, ss, vs = [], [], []
        for it in pk.items:
            m, n = int(it["rows"]), int(it["columns"])
            uo, so, vo = int(it["u_offset"]), int(it["sigma_offset"]), int(it["v_offset"])
            us.append(u[uo:uo + m * n].reshape(m, n))
            ss.append(s[so:so + n].copy())
            vs.append(v[vo:vo + n * n].reshape(n, n))
        return us, ss, vs, sweeps


# ------------------------------------------------------------------------------------------------
# request loop (jsvd_server_*)
# ------------------------------------------------------------------------------------------------
class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("lanes", ctypes.c_int32), ("small_max_rows", ctypes.c_int32), ("small_max_columns", ctypes.c_int32),
        ("max_batch", ctypes.c_int32), ("hbm_budget_bytes", ctypes.c_int64), ("max_sweeps", ctypes.c_int32),
        ("large_path", ctypes.c_int32),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("sweeps", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("batch_size", ctypes.c_int32), ("queued_s", ctypes.c_double), ("latency_s", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in ("submitted", "completed", "batches", "plans", "budget_in_use",
                                              "budget_peak", "max_in_flight")]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def _bind_server(L):
    if getattr(L, "_server_bound", False):
        return
    L.jsvd_server_config_default.argtypes = [ctypes.POINTER(ServerConfig)]
    L.jsvd_server_config_default.restype = None
    L.jsvd_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
    L.jsvd_server_destroy.argtypes = [_vp]
    L.jsvd_server_submit.argtypes = [_vp, _vp, _i64, _i64, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)]
    L.jsvd_server_wait.argtypes = [_vp, ctypes.c_uint64, _vp, _vp, _vp, ctypes.POINTER(RequestInfo)]
    L.jsvd_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    L._server_bound = True


class Server:
    """Continuous-batching request loop (jsvd_server_*).  submit() returns an id at once; wait(id) blocks."""

    def __init__(self, **kw):
        L = lib()
        _bind_server(L)
        cfg = ServerConfig()
        L.jsvd_server_config_default(ctypes.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, v)
        h = _vp()
        check(L.jsvd_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._shapes = {}

    def submit(self, a, relerr: float) -> int:
        arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        m, n = arr.shape
        rid = ctypes.c_uint64()
        check(lib().jsvd_server_submit(self._h, arr.ctypes.data, m, n, float(relerr), ctypes.byref(rid)))
        self._shapes[rid.value] = (m, n)
        return rid.value

    def wait(self, rid: int):
        m, n = self._shapes.pop(rid)
        u = np.zeros((m, n))
        s = np.zeros(n)
        v = np.zeros((n, n))
        info = RequestInfo()
        st = lib().jsvd_server_wait(self._h, rid, u.ctypes.data, s.ctypes.data, v.ctypes.data, ctypes.byref(info))
        check(st)
        return u, s, v, info

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().jsvd_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h:
            check(lib().jsvd_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------------------------
# *.mat files and the streamed file solve (jsvd_mat_stream_solve)
# ------------------------------------------------------------------------------------------------
def write_mat(path: str, matrices) -> None:
    """Writes row-major float64 matrices as *.mat records (u64_be rows | u64_be cols | f64_be payload)."""
    with open(path, "wb") as f:
        for m in matrices:
            arr = np.ascontiguousarray(np.asarray(m, dtype=np.float64))
            if arr.ndim != 2:
                raise ValueError("matrices must be 2-D")
            f.write(np.array(arr.shape, dtype=">u8").tobytes())
            f.write(arr.astype(">f8").tobytes())


def read_mat(path: str) -> list:
    """Reads every *.mat record of a file (memory-mapped) into float64 arrays."""
    buf = np.memmap(path, dtype=np.uint8, mode="r") if os.path.getsize(path) else np.zeros(0, np.uint8)
    out, pos = [], 0
    while pos < buf.size:
        if buf.size - pos < 16:
            raise RuntimeError("Truncated *.mat header.")
        rows, cols = (int(x) for x in np.frombuffer(buf[pos:pos + 16].tobytes(), dtype=">u8"))
        pos += 16
        nbytes = rows * cols * 8
        if buf.size - pos < nbytes:
            raise RuntimeError("Truncated *.mat payload.")
        out.append(np.frombuffer(buf[pos:pos + nbytes].tobytes(), dtype=">f8").astype(np.float64).reshape(rows, cols))
        pos += nbytes
    return out


class MatStreamReport(ctypes.Structure):
    """jsvd_mat_stream_report (include/jsvd.h)."""
    _fields_ = [
        ("records", ctypes.c_int64),
        ("payload_bytes", ctypes.c_int64),
        ("seconds_total", ctypes.c_double),
        ("seconds_load", ctypes.c_double),
        ("seconds_solve", ctypes.c_double),
        ("seconds_output", ctypes.c_double),
        ("sweeps_fp32_last", ctypes.c_int32),
        ("sweeps_fp64_last", ctypes.c_int32),
        ("total_sweeps", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def mat_stream_solve(input_path: str, output_path: Optional[str], config: Optional[JacobiSvdConfig] = None,
                     path: int = PATH_PLAIN_F64, mixed: Optional[MixedConfig] = None) -> MatStreamReport:
    """Solves every record of a *.mat file, overlapping the load of the next record with the current solve;
    writes U, sigma (1 x n), V per record to output_path (None: results are discarded)."""
    L = lib()
    L.jsvd_mat_stream_solve.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(JsvdConfig),
                                        ctypes.POINTER(MixedConfig), ctypes.POINTER(MatStreamReport)]
    rep = MatStreamReport()
    cfg = _as_c_config(config)
    mx = mixed if mixed is not None else mixed_config(path)
    check(L.jsvd_mat_stream_solve(input_path.encode(), output_path.encode() if output_path else None,
                                  ctypes.byref(cfg), ctypes.byref(mx), ctypes.byref(rep)))
    return rep


class TuneReport(ctypes.Structure):
    """jsvd_tune_report (include/jsvd.h)."""
    _fields_ = [
        ("candidates_tried", ctypes.c_int32),
        ("round_threads", ctypes.c_int32),
        ("round_ea", ctypes.c_int32),
        ("round_ev", ctypes.c_int32),
        ("apply_workgroups", ctypes.c_int32),
        ("block_width", ctypes.c_int32),
        ("rt_block_width", ctypes.c_int32),
        ("rt_apply_workgroups", ctypes.c_int32),
        ("round_us_default", ctypes.c_double),
        ("round_us_best", ctypes.c_double),
        ("block_sweep_us", ctypes.c_double * 3),
        ("block_round_us_default", ctypes.c_double),
        ("block_round_us_best", ctypes.c_double),
        ("span_waves", ctypes.c_int32),
        ("span_rounds", ctypes.c_int32),
        ("sweep_us_rounds", ctypes.c_double),
        ("sweep_us_span", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        out = {f: getattr(self, f) for f, _ in self._fields_}
        out["block_sweep_us"] = list(self.block_sweep_us)
        return out


def autotune(m: int, n: int, repetitions: int = 2, stream=None) -> TuneReport:
    """Measures the launch parameters for m x n on the device and caches the fastest for plans of that shape
    created or reshaped afterwards in this process (round-kernel launch shape, block apply grid size)."""
    L = lib()
    L.jsvd_autotune.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                ctypes.POINTER(TuneReport)]
    rep = TuneReport()
    s = _stream_handle(stream) if stream is not None else None
    check(L.jsvd_autotune(m, n, repetitions, s, ctypes.byref(rep)))
    return rep


def autotune_clear() -> None:
    """Forgets every cached autotune result."""
    check(lib().jsvd_autotune_clear())
