"""ctypes front end of the C oracle (oracle/lbm_oracle_c.c) -- TEST INFRASTRUCTURE ONLY.

The method tables are packed from :class:`oracle.lbm_oracle.OracleCollision` (moment matrices,
binomial-shift pairs, Leibniz terms of the cumulant recursion, equilibrium polynomials), so the C
code runs exactly the numpy oracle's generic algorithm, per cell and multi-threaded.
"""
import ctypes
import os

import numpy as np

from oracle.lbm_oracle import OracleCollision
from lbmpy_amd.equilibrium import EQ_VARS

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liblbm_oracle_c.so")
QMAX, TMAX = 27, 64
KIND_MOMENT_POP, KIND_MOMENT_RAW, KIND_CUMULANT = 0, 1, 2
FM = {None: 0, "Simple": 1, "Luo": 2, "Guo": 3, "CentralMoment": 4}
PATTERNS = {"pull": 0, "aa_even": 1, "aa_odd": 2}


class OracleMethod(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("Q", ctypes.c_int32), ("D", ctypes.c_int32),
                ("force_model", ctypes.c_int32), ("zero_centered", ctypes.c_int32),
                ("compressible", ctypes.c_int32), ("eq_shift", ctypes.c_int32), ("guo_scalar", ctypes.c_int32),
                ("c", (ctypes.c_int32 * 3) * QMAX), ("w", ctypes.c_double * QMAX), ("F", ctypes.c_double * 3),
                ("rates", ctypes.c_double * QMAX),
                ("M", (ctypes.c_double * QMAX) * QMAX), ("Minv", (ctypes.c_double * QMAX) * QMAX),
                ("bg", ctypes.c_double * QMAX), ("has_bg", ctypes.c_int32),
                ("n_terms", ctypes.c_int32 * QMAX), ("coef", (ctypes.c_double * TMAX) * QMAX),
                ("pow_", ((ctypes.c_int8 * 5) * TMAX) * QMAX),
                ("P", (ctypes.c_double * QMAX) * QMAX), ("Pinv", (ctypes.c_double * QMAX) * QMAX),
                ("mono", (ctypes.c_int32 * 3) * QMAX), ("idx000", ctypes.c_int32),
                ("first_order", ctypes.c_int32 * 3), ("n_shift", ctypes.c_int32),
                ("shift_i", ctypes.c_int32 * 256), ("shift_r", ctypes.c_int32 * 256),
                ("shift_pw", (ctypes.c_int32 * 3) * 256), ("shift_c", ctypes.c_double * 256),
                ("order", ctypes.c_int32 * QMAX), ("leib_start", ctypes.c_int32 * (QMAX + 1)),
                ("leib_rj", ctypes.c_int32 * 512), ("leib_rest", ctypes.c_int32 * 512),
                ("leib_b", ctypes.c_double * 512)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run python -m oracle.build")
        lib = ctypes.CDLL(LIB)
        lib.oracle_sweep.argtypes = [ctypes.POINTER(OracleMethod), ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        lib.oracle_sweep.restype = ctypes.c_int32
        lib.oracle_collide.argtypes = [ctypes.POINTER(OracleMethod), ctypes.c_void_p, ctypes.c_int64]
        lib.oracle_collide.restype = ctypes.c_int32
        lib.oracle_method_size.restype = ctypes.c_int64
        assert lib.oracle_method_size() == ctypes.sizeof(OracleMethod), "oracle_method layout mismatch"
        _lib = lib
    return _lib


def _set_poly(om, i, poly):
    if poly is None:
        om.n_terms[i] = -1
        return
    terms = list(poly.terms.items())
    if len(terms) > TMAX:
        raise ValueError("equilibrium polynomial too long for the C oracle")
    om.n_terms[i] = len(terms)
    for t, (e, c) in enumerate(terms):
        om.coef[i][t] = float(c)
        for k in range(len(EQ_VARS)):
            om.pow_[i][t][k] = e[k]


def pack(method, bindings=None):
    """OracleMethod for a method (moment family or cumulant family)."""
    oc = OracleCollision(method, bindings)
    om = OracleMethod()
    Q, D = oc.Q, oc.D
    om.Q, om.D = Q, D
    om.zero_centered = int(bool(method.zero_centered_pdfs))
    om.compressible = int(bool(method.compressible))
    fm = method.force_model
    name = type(fm).__name__ if fm is not None else None
    om.force_model = FM[name]
    om.eq_shift = int(bool(fm is not None and fm.equilibrium_velocity_shift))
    for i in range(Q):
        for a in range(D):
            om.c[i][a] = int(oc.c[i, a])
        om.w[i] = oc.w[i]
        om.rates[i] = oc.rates[i]
    for a in range(D):
        om.F[a] = oc.F[a]
    if method.kind == "moment":
        om.kind = KIND_MOMENT_RAW if oc.raw_space else KIND_MOMENT_POP
        om.guo_scalar = int(len(set(oc.rates.tolist())) == 1)
        M, Minv = oc.M, oc.Minv
        for i, p in enumerate(oc.meq_polys):
            _set_poly(om, i, p)
        if oc.bg is not None:
            om.has_bg = 1
            for i in range(Q):
                om.bg[i] = oc.bg[i]
    elif method.kind == "cumulant":
        om.kind = KIND_CUMULANT
        M, Minv = oc.Mmono, oc.Mmono_inv
        for i in range(Q):
            for j in range(Q):
                om.P[i][j] = oc.P[i, j]
                om.Pinv[i][j] = oc.Pinv[i, j]
        for i, p in enumerate(oc.ceq_polys):
            _set_poly(om, i, p)
        if oc.bg is not None:
            om.has_bg = 1
            for i in range(Q):
                om.bg[i] = oc.bg[i]
        monos = oc.monos
        for i, e in enumerate(monos):
            for a in range(3):
                om.mono[i][a] = e[a]
        om.idx000 = oc.midx[(0, 0, 0)]
        for a, i in enumerate(oc._first_order_indices()):
            om.first_order[a] = i
        if len(oc.shift_pairs) > 256:
            raise ValueError("too many shift pairs")
        om.n_shift = len(oc.shift_pairs)
        for k, (i, r, binom, pw) in enumerate(oc.shift_pairs):
            om.shift_i[k], om.shift_r[k] = i, r
            om.shift_c[k] = float(binom[0] * binom[1] * binom[2])
            for a in range(3):
                om.shift_pw[k][a] = pw[a]
        order = sorted(range(Q), key=lambda i: sum(monos[i]))
        for o, i in enumerate(order):
            om.order[o] = i
        pos = 0
        for i, e in enumerate(monos):
            om.leib_start[i] = pos
            if sum(e) == 0:
                continue
            for b, rj, rest in oc._leibniz_terms(e):
                om.leib_b[pos] = float(b)
                om.leib_rj[pos] = oc.midx[rj]
                om.leib_rest[pos] = oc.midx[rest]
                pos += 1
        om.leib_start[Q] = pos          # the zeroth monomial's range is empty
        if pos > 512:
            raise ValueError("too many Leibniz terms")
    else:
        raise NotImplementedError("the C oracle restates the moment and cumulant families")
    for i in range(Q):
        for j in range(Q):
            om.M[i][j] = M[i, j]
            om.Minv[i][j] = Minv[i, j]
    return om


class COracle:
    """Fully periodic domain without ghost layers, pdf array (Q, [nz,] ny, nx) float64."""

    def __init__(self, method, bindings=None):
        self.lib = load()
        self.om = pack(method, bindings)
        self.method = method

    def sweep(self, src, dst, pattern, nthreads=0):
        assert src.dtype == np.float64 and src.flags.c_contiguous
        shape = src.shape[1:]
        nx, ny, nz = shape[-1], shape[-2], (shape[-3] if len(shape) == 3 else 1)
        rc = self.lib.oracle_sweep(ctypes.byref(self.om), src.ctypes.data, dst.ctypes.data, nx, ny, nz,
                                   PATTERNS[pattern], nthreads)
        if rc:
            raise ValueError("oracle_sweep rejected its arguments")

    def collide(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).copy()
        self.lib.oracle_collide(ctypes.byref(self.om), f.ctypes.data, f.shape[1])
        return f
