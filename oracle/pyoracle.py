"""ctypes bindings for the CPU checkers (test infrastructure only).

Two libraries share one calling convention:

* ``oracle/liboracle.so`` — the C restatement of the reference DIM update
  (``dim_oracle.c``), prefix ``orc_``;
* ``oracle/_ref/libdimref.so`` — the reference's own functions compiled from
  its headers behind ``ref_harness.cpp``, prefix ``ref_``.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdimref.so")

EOS_IDEAL, EOS_STIFFENED, EOS_NOBLE_ABEL = 0, 1, 2
EOS_KIND = {"ideal_gas": EOS_IDEAL, "stiffened_gas": EOS_STIFFENED, "noble_abel": EOS_NOBLE_ABEL,
            "noble_abel_gas": EOS_NOBLE_ABEL, "real_gas": EOS_NOBLE_ABEL}
LAMBDA = {"kapila": 0, "allaire": 1, "equal_velocity": 2}
# boundary codes shared with include/dimhip.h; 3 = the name "reflection",
# which the reference accepts but does not flip (boundary.hpp:219-221)
BC = {"transmissive": 0, "reflective": 1, "nonreflective": 2, "reflection": 3}


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("nmat", C.c_int),
        ("kind", C.c_int * 4),
        ("gamma", C.c_double * 4),
        ("pinf", C.c_double * 4),
        ("lambda_model", C.c_int),
        ("alpha_source_floor", C.c_double),
        ("covolume", C.c_double * 4),
    ]


def make_params(dim, materials, lambda_model="kapila", alpha_source_floor=0.0):
    """materials: list of (kind, gamma, p_inf[, covolume]) with kind 'ideal_gas',
    'stiffened_gas' or 'noble_abel' (or its code)."""
    p = Params()
    p.dim = dim
    p.nmat = len(materials)
    for k, mat in enumerate(materials):
        kind, g, pinf = mat[:3]
        p.kind[k] = EOS_KIND[kind] if isinstance(kind, str) else int(kind)
        p.gamma[k] = g
        p.pinf[k] = pinf
        p.covolume[k] = mat[3] if len(mat) > 3 else 0.0
    p.lambda_model = LAMBDA[lambda_model] if isinstance(lambda_model, str) else lambda_model
    p.alpha_source_floor = alpha_source_floor
    return p


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Checker:
    """One of the two CPU checkers; ``kind`` is 'oracle' or 'ref'."""

    def __init__(self, kind="oracle"):
        self.kind = kind
        path = ORACLE_LIB if kind == "oracle" else REF_LIB
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle oracle ref`)")
        self.lib = C.CDLL(path)
        self.pre = "orc_" if kind == "oracle" else "ref_"
        self._fn("last_error").restype = C.c_char_p

    def _fn(self, name):
        return getattr(self.lib, self.pre + name)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._fn("last_error")().decode())

    @staticmethod
    def nv(p):
        return 2 * p.nmat + p.dim

    @staticmethod
    def np_(p):
        return 2 * p.nmat + p.dim + 1

    # --- function level ---------------------------------------------------
    def cons_to_prim(self, p, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, self.nv(p))
        out = np.zeros((U.shape[0], self.np_(p)))
        self._check(self._fn("cons_to_prim")(C.byref(p), U.shape[0], _dp(U), _dp(out)))
        return out

    def prim_to_cons(self, p, P):
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, self.np_(p))
        out = np.zeros((P.shape[0], self.nv(p)))
        self._check(self._fn("prim_to_cons")(C.byref(p), P.shape[0], _dp(P), _dp(out)))
        return out

    def repair_state(self, p, U):
        U = np.array(U, dtype=np.float64).reshape(-1, self.nv(p))
        self._check(self._fn("repair_state")(C.byref(p), U.shape[0], _dp(U)))
        return U

    def alpha_rhs(self, p, P):
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, self.np_(p))
        out = np.zeros((P.shape[0], p.nmat))
        self._check(self._fn("alpha_rhs")(C.byref(p), P.shape[0], _dp(P), _dp(out)))
        return out

    def pressure(self, p, rho_total, e, alpha, rho):
        rho_total = np.ascontiguousarray(rho_total, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        out = np.zeros(rho_total.shape[0])
        self._check(self._fn("pressure_from_internal_energy")(
            C.byref(p), rho_total.shape[0], _dp(rho_total), _dp(e), _dp(alpha), _dp(rho), _dp(out)))
        return out

    def sound_speed(self, p, rho_total, pr, alpha, rho):
        rho_total = np.ascontiguousarray(rho_total, dtype=np.float64)
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        out = np.zeros(rho_total.shape[0])
        self._check(self._fn("mixture_sound_speed")(
            C.byref(p), rho_total.shape[0], _dp(rho_total), _dp(pr), _dp(alpha), _dp(rho), _dp(out)))
        return out

    def hllc(self, p, d, UL, UR):
        UL = np.ascontiguousarray(UL, dtype=np.float64).reshape(-1, self.nv(p))
        UR = np.ascontiguousarray(UR, dtype=np.float64).reshape(-1, self.nv(p))
        F = np.zeros((UL.shape[0], p.nmat + p.dim + 1))
        fv = np.zeros(UL.shape[0])
        self._check(self._fn("hllc")(C.byref(p), d, UL.shape[0], _dp(UL), _dp(UR), _dp(F), _dp(fv)))
        return F, fv

    def reconstruct_line(self, p, d, U, dt, dx):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, self.nv(p))
        n = U.shape[0]
        L = np.zeros((max(n - 1, 0), self.nv(p)))
        R = np.zeros_like(L)
        self._check(self._fn("reconstruct_line")(C.byref(p), d, n, _dp(U), C.c_double(dt),
                                                  C.c_double(dx), _dp(L), _dp(R)))
        return L, R

    def solve_line(self, p, d, U, dx, dt):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, self.nv(p))
        out = np.zeros_like(U)
        self._check(self._fn("solve_line")(C.byref(p), d, U.shape[0], _dp(U), _dp(out),
                                            C.c_double(dx), C.c_double(dt)))
        return out

    # --- whole grid -------------------------------------------------------
    def dt_cfl(self, p, N, dx, U, cfl, dt_max, unsplit=False):
        N = np.ascontiguousarray(N, dtype=np.int32)
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        U = np.ascontiguousarray(U, dtype=np.float64)
        dt = C.c_double(0.0)
        self._check(self._fn("dt_cfl")(C.byref(p), _ip(N), _dp(dx), _dp(U), C.c_double(cfl),
                                        C.c_double(dt_max), int(unsplit), C.byref(dt)))
        return dt.value

    def apply_bc(self, p, N, bc_lo, bc_hi, U):
        N = np.ascontiguousarray(N, dtype=np.int32)
        lo = np.ascontiguousarray([self._bc(b) for b in bc_lo], dtype=np.int32)
        hi = np.ascontiguousarray([self._bc(b) for b in bc_hi], dtype=np.int32)
        U = np.array(U, dtype=np.float64)
        self._check(self._fn("apply_bc")(C.byref(p), _ip(N), _ip(lo), _ip(hi), _dp(U)))
        return U

    def _bc(self, b):
        code = BC[b] if isinstance(b, str) else int(b)
        # the restatement folds "reflection" into the copy behaviour
        if self.kind == "oracle" and code == 3:
            return 0
        return code

    def advance(self, p, N, dx, bc_lo, bc_hi, U, cfl, dt_max, nsteps=1, unsplit=False, threads=0):
        """nsteps x dim::advance_one_step.  Returns (U_new, dts, seconds)."""
        N = np.ascontiguousarray(N, dtype=np.int32)
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        lo = np.ascontiguousarray([self._bc(b) for b in bc_lo], dtype=np.int32)
        hi = np.ascontiguousarray([self._bc(b) for b in bc_hi], dtype=np.int32)
        U = np.array(U, dtype=np.float64)
        dts = np.zeros(nsteps)
        if self.kind == "ref":
            secs = C.c_double(0.0)
            self._check(self._fn("advance")(C.byref(p), _ip(N), _dp(dx), _ip(lo), _ip(hi),
                                             int(unsplit), C.c_double(cfl), C.c_double(dt_max),
                                             nsteps, threads, _dp(U), _dp(dts), C.byref(secs)))
            return U, dts, secs.value
        import time
        t0 = time.perf_counter()
        for s in range(nsteps):
            dt = C.c_double(0.0)
            self._check(self._fn("advance")(C.byref(p), _ip(N), _dp(dx), _ip(lo), _ip(hi),
                                             int(unsplit), C.c_double(cfl), C.c_double(dt_max),
                                             _dp(U), C.byref(dt)))
            dts[s] = dt.value
        return U, dts, time.perf_counter() - t0

    def conservation_totals(self, p, N, dx, U):
        if self.kind != "oracle":
            raise NotImplementedError
        N = np.ascontiguousarray(N, dtype=np.int32)
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        U = np.ascontiguousarray(U, dtype=np.float64)
        out = np.zeros(p.nmat + 1 + p.dim + 1)
        self._check(self._fn("conservation_totals")(C.byref(p), _ip(N), _dp(dx), _dp(U), _dp(out)))
        return out


def ref_initial_state(path, dim, n_index=0):
    """The reference's initial state of a case file (load_config +
    initialise_dim_from_config), packed (ncells, 2*nmat + dim)."""
    lib = C.CDLL(REF_LIB)
    fn = lib.ref_initial_state
    fn.restype = C.c_int
    fn.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                   C.POINTER(C.c_int)]
    lib.ref_last_error.restype = C.c_char_p
    n = C.c_longlong(0)
    m = C.c_int(0)
    if fn(path.encode(), dim, n_index, None, C.byref(n), C.byref(m)) != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    U = np.zeros((n.value, 2 * m.value + dim))
    if fn(path.encode(), dim, n_index, _dp(U), C.byref(n), C.byref(m)) != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    return U


def ref_available():
    return os.path.exists(REF_LIB)
