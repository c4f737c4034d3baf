"""ctypes front-end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this module, and
only to check results against.  The product library never calls it.

Two oracles live here:

* ``svd`` / ``rotation`` / ``round_robin`` / ``u_sigma`` call oracle/_build/liboracle.so, the C
  restatement of the reference solver (src/kernels.cu:485-678, 844-1004; see jacobi_oracle.c).
* ``dgesvj`` calls LAPACK ``dgesvj`` through the LAPACKE symbol exported by SciPy's bundled OpenBLAS
  (SciPy 1.15.3 in this image).  It is an independent FP64 one-sided Jacobi with its own ordering and
  tolerance (sqrt(m)*eps_mach), so it pins sigma, not the column order.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dptr = ctypes.POINTER(ctypes.c_double)
_iptr = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile jacobi_oracle.c with gcc (make) into oracle/_build/liboracle.so."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.oracle_round_robin.argtypes = [ctypes.c_int, _iptr, _iptr]
        lib.oracle_round_robin.restype = ctypes.c_int
        lib.oracle_rotation.argtypes = [ctypes.c_double] * 4 + [_dptr, _dptr]
        lib.oracle_rotation.restype = ctypes.c_int
        lib.oracle_pair_stats.argtypes = [_dptr, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                          _dptr, _dptr, _dptr]
        lib.oracle_u_sigma.argtypes = [_dptr, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                       _dptr, _dptr]
        lib.oracle_svd_ex.argtypes = [_dptr, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, _dptr, _dptr, _dptr, _iptr, _iptr, _dptr]
        lib.oracle_svd_ex.restype = ctypes.c_int
        lib.oracle_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dptr)


def round_robin(n: int) -> list[list[tuple[int, int]]]:
    """Round-robin schedule, src/kernels.cu:637-678 (see oracle_round_robin)."""
    lib = _load()
    if n < 2:
        return []
    even = n + (n % 2)
    half = even // 2
    pairs = np.zeros(2 * (even - 1) * half, dtype=np.int32)
    counts = np.zeros(even - 1, dtype=np.int32)
    rounds = lib.oracle_round_robin(n, pairs.ctypes.data_as(_iptr), counts.ctypes.data_as(_iptr))
    out = []
    for r in range(rounds):
        row = pairs[2 * r * half: 2 * (r * half + counts[r])].reshape(-1, 2)
        out.append([(int(p), int(q)) for p, q in row])
    return out


def rotation(app: float, aqq: float, apq: float, eps: float) -> tuple[int, float, float]:
    """Rotation parameters and skip test, src/kernels.cu:485-521.  Returns (rotates, c, s)."""
    lib = _load()
    c = ctypes.c_double()
    s = ctypes.c_double()
    r = lib.oracle_rotation(app, aqq, apq, eps, ctypes.byref(c), ctypes.byref(s))
    return int(r), c.value, s.value


def pair_stats(a_rowmajor: np.ndarray, p: int, q: int) -> tuple[float, float, float]:
    """(alpha, beta, gamma) of columns p, q, src/kernels.cu:412-472."""
    lib = _load()
    acol = np.ascontiguousarray(np.asarray(a_rowmajor, dtype=np.float64).T)
    m = a_rowmajor.shape[0]
    x, y, z = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib.oracle_pair_stats(_p(acol), m, m, p, q, ctypes.byref(x), ctypes.byref(y), ctypes.byref(z))
    return x.value, y.value, z.value


def u_sigma(a_rowmajor: np.ndarray, eps: float) -> tuple[np.ndarray, np.ndarray]:
    """sigma / U extraction, src/kernels.cu:586-629."""
    lib = _load()
    a = np.asarray(a_rowmajor, dtype=np.float64)
    m, n = a.shape
    acol = np.ascontiguousarray(a.T)
    u = np.zeros((m, n))
    s = np.zeros(n)
    lib.oracle_u_sigma(_p(acol), m, m, n, eps, _p(u), _p(s))
    return u, s


class OracleResult:
    def __init__(self, u, sigma, v, sweeps, rot_per_sweep, offmax_per_sweep):
        self.u = u
        self.sigma = sigma
        self.v = v
        self.sweeps = sweeps
        self.rot_per_sweep = rot_per_sweep
        self.offmax_per_sweep = offmax_per_sweep


def svd(a_rowmajor: np.ndarray, eps: float = 1e-9, max_sweeps: int = 128, threads: int = 0) -> OracleResult:
    """Full reference-semantics solve, src/kernels.cu:844-1004 (see oracle_svd_ex)."""
    lib = _load()
    a = np.ascontiguousarray(a_rowmajor, dtype=np.float64)
    m, n = a.shape
    u = np.zeros((m, n))
    s = np.zeros(n)
    v = np.zeros((n, n))
    sweeps = ctypes.c_int()
    rot = np.zeros(max(max_sweeps, 1), dtype=np.int32)
    off = np.zeros(max(max_sweeps, 1))
    rc = lib.oracle_svd_ex(_p(a), m, n, eps, max_sweeps, threads, _p(u), _p(s), _p(v), ctypes.byref(sweeps),
                           rot.ctypes.data_as(_iptr), _p(off))
    if rc != 0:
        raise ValueError("oracle_svd: invalid arguments")
    k = sweeps.value
    return OracleResult(u, s, v, k, rot[:k].copy(), off[:k].copy())


def max_threads() -> int:
    return int(_load().oracle_max_threads())


# --------------------------------------------------------------------------------------------
# LAPACK dgesvj (SciPy-bundled OpenBLAS, LAPACKE interface)
# --------------------------------------------------------------------------------------------
_lapacke = None


def _load_lapacke():
    global _lapacke
    if _lapacke is None:
        import scipy  # noqa: F401  (locates scipy.libs)
        libdir = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        cands = sorted(glob.glob(os.path.join(libdir, "libscipy_openblas*.so")))
        if not cands:
            raise RuntimeError("SciPy-bundled OpenBLAS not found; dgesvj oracle unavailable")
        lib = ctypes.CDLL(cands[0])
        fn = lib.scipy_LAPACKE_dgesvj
        fn.argtypes = [ctypes.c_int, ctypes.c_char, ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int,
                       _dptr, ctypes.c_int, _dptr, ctypes.c_int, _dptr, ctypes.c_int, _dptr]
        fn.restype = ctypes.c_int
        _lapacke = fn
    return _lapacke


def dgesvj(a_rowmajor: np.ndarray):
    """LAPACK dgesvj (JOBA='G', JOBU='U', JOBV='V').  Returns (U m x n, sigma n, V n x n, sweeps).

    sigma is in dgesvj's output order (descending); columns of U and V follow it.
    """
    fn = _load_lapacke()
    a = np.array(a_rowmajor, dtype=np.float64, order="F", copy=True)
    m, n = a.shape
    sva = np.zeros(n)
    v = np.zeros((n, n), order="F")
    stat = np.zeros(6)
    info = fn(102, b"G", b"U", b"V", m, n, _p(a), m, _p(sva), n, _p(v), n, _p(stat))
    if info < 0:
        raise ValueError(f"dgesvj: illegal argument {-info}")
    sigma = stat[0] * sva
    return np.ascontiguousarray(a), sigma, np.ascontiguousarray(v), int(stat[3])
