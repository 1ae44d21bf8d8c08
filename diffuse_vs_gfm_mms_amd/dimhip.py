"""Python host binding of the gfx950 DIM library (include/dimhip.h) via ctypes.

:class:`Solver` mirrors the reference's stepping interface
``dim::advance_one_step(U, N, dx, params, cfg, cfl, dt_max)``
(src/dim/solver/advance_step.hpp:20-149) over a device-resident state, plus
the multi-step device loop of ``run_dim_case`` (src/app/solver/run_dim.hpp:88-111).
Errors raise ``RuntimeError`` with the library's message, which keeps the
reference's ``std::runtime_error`` wording where the reference has one.

There is no CPU fallback: if ``_lib/libdimhip.so`` is missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libdimhip.so")

EOS = {"ideal_gas": 0, "stiffened_gas": 1, "noble_abel": 2, "noble_abel_gas": 2, "real_gas": 2}
BC = {"transmissive": 0, "zero_gradient": 0, "outflow": 0, "reflective": 1,
      "nonreflective": 2, "non_reflective": 2, "reflection": 3}
LAMBDA = {"kapila": 0, "": 0, "allaire": 1, "equal_velocity": 2}
TIME_UPDATE = {"split": 0, "unsplit": 1}
PRECISION = {"fp64": 0, "fp32_mixed": 1}

KAT_CONS_TO_PRIM, KAT_PRIM_TO_CONS, KAT_REPAIR, KAT_HLLC, KAT_ALPHA_RHS, KAT_IEEE, KAT_PRESSURE, KAT_COPY = range(8)


class Desc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("N", C.c_int * 3),
        ("dx", C.c_double * 3),
        ("nmat", C.c_int),
        ("eos_kind", C.c_int * 4),
        ("gamma", C.c_double * 4),
        ("p_inf", C.c_double * 4),
        ("bc_lo", C.c_int * 3),
        ("bc_hi", C.c_int * 3),
        ("lambda_model", C.c_int),
        ("alpha_source_floor", C.c_double),
        ("time_update", C.c_int),
        ("precision", C.c_int),
        ("device", C.c_int),
        ("covolume", C.c_double * 4),
    ]


IC_KIND = {"regions": 0, "planar_regions": 1, "shock_bubble": 2, "coated_shock_bubble": 3}
MAX_REGIONS = 16


class IcRegion(C.Structure):
    _fields_ = [("lower", C.c_double * 3), ("upper", C.c_double * 3), ("rho", C.c_double),
                ("vel", C.c_double * 3), ("p", C.c_double), ("material", C.c_int)]


class IcState(C.Structure):
    _fields_ = [("rho", C.c_double), ("vel", C.c_double * 3), ("p", C.c_double), ("material", C.c_int)]


class Ic(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("domain_min", C.c_double * 3),
        ("domain_max", C.c_double * 3),
        ("interface_thickness", C.c_double),
        ("sharpness_alpha", C.c_double),
        ("planar_normal", C.c_double * 3),
        ("nregions", C.c_int),
        ("regions", IcRegion * MAX_REGIONS),
        ("shock_axis", C.c_int),
        ("shock_position", C.c_double),
        ("left", IcState),
        ("right", IcState),
        ("bubble", IcState),
        ("film", IcState),
        ("bubble_center", C.c_double * 3),
        ("bubble_radius", C.c_double),
        ("film_radius", C.c_double),
    ]


def _fill_state(dst, st):
    rho, vel, p, slot = st
    dst.rho, dst.p, dst.material = rho, p, slot
    for a, v in enumerate(vel):
        dst.vel[a] = v


def make_ic(kind, domain_min, domain_max, thickness, sharpness=2.0, regions=(), planar_normal=None,
            shock_axis=0, shock_position=0.0, left=None, right=None, bubble=None, film=None,
            centre=(0.0, 0.0, 0.0), radius=0.0, film_radius=0.0):
    """An initial-condition description for Solver.initialise.  States are
    (rho, vel tuple, p, material slot); regions are (lower, upper, rho, vel, p, slot)."""
    ic = Ic()
    ic.kind = IC_KIND[kind]
    for a in range(len(domain_min)):
        ic.domain_min[a], ic.domain_max[a] = domain_min[a], domain_max[a]
    ic.interface_thickness, ic.sharpness_alpha = thickness, sharpness
    if planar_normal is not None:
        for a, v in enumerate(planar_normal):
            ic.planar_normal[a] = v
    if len(regions) > MAX_REGIONS:
        raise RuntimeError("dimhip_initialise: too many regions")
    ic.nregions = len(regions)
    for i, (lower, upper, rho, vel, p, slot) in enumerate(regions):
        r = ic.regions[i]
        for a in range(len(lower)):
            r.lower[a], r.upper[a] = lower[a], upper[a]
        for a, v in enumerate(vel):
            r.vel[a] = v
        r.rho, r.p, r.material = rho, p, slot
    ic.shock_axis, ic.shock_position = shock_axis, shock_position
    for name, st in (("left", left), ("right", right), ("bubble", bubble), ("film", film)):
        if st is not None:
            _fill_state(getattr(ic, name), st)
        else:
            getattr(ic, name).material = -1
    for a, v in enumerate(centre):
        ic.bubble_center[a] = v
    ic.bubble_radius, ic.film_radius = radius, film_radius
    return ic


class Timings(C.Structure):
    _fields_ = [("cfl_seconds", C.c_double), ("sweep_seconds", C.c_double),
                ("boundary_seconds", C.c_double)]


_lib = None


def load_library():
    """Load the in-tree HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    vp, dp, ip, llp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_longlong)
    sig = {
        "dimhip_version": (C.c_int, []),
        "dimhip_create": (C.c_int, [C.POINTER(Desc), C.POINTER(vp)]),
        "dimhip_destroy": (C.c_int, [vp]),
        "dimhip_upload": (C.c_int, [vp, dp, C.c_size_t]),
        "dimhip_download": (C.c_int, [vp, dp, C.c_size_t]),
        "dimhip_step": (C.c_int, [vp, C.c_double, C.c_double, dp, C.POINTER(Timings)]),
        "dimhip_advance": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_int, dp, llp]),
        "dimhip_compute_dt": (C.c_int, [vp, C.c_double, C.c_double, dp]),
        "dimhip_sweep": (C.c_int, [vp, C.c_int, C.c_double]),
        "dimhip_apply_bc": (C.c_int, [vp]),
        "dimhip_conservation": (C.c_int, [vp, dp]),
        "dimhip_boundary_flux": (C.c_int, [vp, dp]),
        "dimhip_track_boundary_flux": (C.c_int, [vp, C.c_int]),
        "dimhip_boundary_flux_integral": (C.c_int, [vp, dp]),
        "dimhip_upload_primitive": (C.c_int, [vp, dp, C.c_size_t]),
        "dimhip_initialise": (C.c_int, [vp, C.POINTER(Ic)]),
        "dimhip_output_rows": (C.c_int, [vp, dp, C.c_size_t]),
        "dimhip_output_rows_noalloc": (C.c_int, [vp, dp, C.c_size_t]),
        "dimhip_synchronize": (C.c_int, [vp]),
        "dimhip_set_timing": (C.c_int, [vp, C.c_int]),
        "dimhip_kernel_times": (C.c_int, [vp, dp, llp]),
        "dimhip_sweep_times": (C.c_int, [vp, dp, llp]),
        "dimhip_kat": (C.c_int, [C.POINTER(Desc), C.c_int, C.c_int, C.c_int, dp, dp, dp, dp]),
        "dimhip_last_error": (C.c_char_p, [vp]),
        "dimhip_batch_create": (C.c_int, [C.POINTER(Desc), C.c_int, C.POINTER(vp)]),
        "dimhip_batch_destroy": (C.c_int, [vp]),
        "dimhip_batch_upload": (C.c_int, [vp, C.c_int, dp, C.c_size_t]),
        "dimhip_batch_download": (C.c_int, [vp, C.c_int, dp, C.c_size_t]),
        "dimhip_batch_configure": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "dimhip_batch_autotune": (C.c_int, [vp, C.c_double, C.c_int, C.c_char_p, C.c_size_t]),
        "dimhip_batch_advance": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_int, dp, llp]),
        "dimhip_batch_set_timing": (C.c_int, [vp, C.c_int]),
        "dimhip_batch_kernel_times": (C.c_int, [vp, dp, llp]),
        "dimhip_batch_sweep_times": (C.c_int, [vp, dp, llp]),
        "dimhip_batch_last_error": (C.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return ["dimhip_version", "dimhip_create", "dimhip_destroy", "dimhip_upload", "dimhip_download",
            "dimhip_step", "dimhip_advance", "dimhip_compute_dt", "dimhip_sweep", "dimhip_apply_bc",
            "dimhip_conservation", "dimhip_boundary_flux", "dimhip_track_boundary_flux",
            "dimhip_boundary_flux_integral", "dimhip_upload_primitive", "dimhip_initialise", "dimhip_output_rows", "dimhip_output_rows_noalloc", "dimhip_synchronize", "dimhip_set_timing", "dimhip_kernel_times", "dimhip_sweep_times",
            "dimhip_kat", "dimhip_last_error", "dimhip_batch_create", "dimhip_batch_destroy",
            "dimhip_batch_upload", "dimhip_batch_download", "dimhip_batch_configure", "dimhip_batch_autotune",
            "dimhip_batch_advance", "dimhip_batch_set_timing", "dimhip_batch_kernel_times", "dimhip_batch_sweep_times",
            "dimhip_batch_last_error"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_desc(dim, N, dx, materials, bc_lo, bc_hi, lambda_model="kapila", alpha_source_floor=0.0,
              time_update="split", precision="fp64", device=0):
    d = Desc()
    d.dim = dim
    for a in range(3):
        d.N[a] = int(N[a]) if a < dim else 1
        d.dx[a] = float(dx[a]) if a < dim else 1.0
        d.bc_lo[a] = BC[bc_lo[a]] if a < dim else 0
        d.bc_hi[a] = BC[bc_hi[a]] if a < dim else 0
    d.nmat = len(materials)
    for k, mat in enumerate(materials):
        kind, g, pinf = mat[:3]
        d.eos_kind[k] = EOS[kind]
        d.gamma[k] = g
        d.p_inf[k] = pinf
        d.covolume[k] = mat[3] if len(mat) > 3 else 0.0
    d.lambda_model = LAMBDA[lambda_model]
    d.alpha_source_floor = alpha_source_floor
    d.time_update = TIME_UPDATE[time_update]
    d.precision = PRECISION[precision]
    d.device = device
    return d


class Solver:
    """Device-resident DIM state with the reference's stepping semantics."""

    def __init__(self, dim, N, dx, materials, bc_lo, bc_hi, lambda_model="kapila",
                 alpha_source_floor=0.0, time_update="split", precision="fp64", device=0):
        self.lib = load_library()
        self.dim, self.N = dim, tuple(int(n) for n in N)
        self.nmat = len(materials)
        self.nv = 2 * self.nmat + dim
        self.ncells = int(np.prod(self.N))
        self.desc = make_desc(dim, N, dx, materials, bc_lo, bc_hi, lambda_model, alpha_source_floor,
                              time_update, precision, device)
        h = C.c_void_p()
        self._check(self.lib.dimhip_create(C.byref(self.desc), C.byref(h)), None)
        self.h = h

    @classmethod
    def from_case(cls, case, precision="fp64", time_update="split", **kw):
        """A solver holding the case's initial state: its packed U when the case
        carries one, else its initial condition generated on the device."""
        s = cls(case.dim, case.N, case.dx, case.materials, case.bc_lo, case.bc_hi,
                case.lambda_model, case.alpha_source_floor, time_update, precision, **kw)
        if case.U is not None:
            s.upload(case.U)
        else:
            s.initialise(case.ic)
        return s

    def initialise(self, ic):
        """Initial condition generated on the device (dimhip_initialise); ic is
        an Ic from make_ic or a dict of make_ic's keyword arguments."""
        if isinstance(ic, dict):
            ic = make_ic(**ic)
        self._check(self.lib.dimhip_initialise(self.h, C.byref(ic)))

    def _check(self, rc, h="self"):
        if rc != 0:
            handle = self.h if h == "self" else None
            raise RuntimeError(self.lib.dimhip_last_error(handle).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.dimhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != self.ncells * self.nv:
            raise RuntimeError("dim::advance_one_step: grid size mismatch")
        self._check(self.lib.dimhip_upload(self.h, _dp(U), self.ncells))

    def download(self):
        out = np.empty((self.ncells, self.nv))
        self._check(self.lib.dimhip_download(self.h, _dp(out), self.ncells))
        return out

    def step(self, cfl, dt_max):
        """One dim::advance_one_step; returns dt."""
        dt = C.c_double(0.0)
        self._check(self.lib.dimhip_step(self.h, cfl, dt_max, C.byref(dt), None))
        return dt.value

    def advance(self, nsteps, cfl, t_end, time=0.0, use_graph=True):
        """Up to nsteps loop iterations of run_dim_case with dt_max = t_end - time."""
        t = C.c_double(time)
        n = C.c_longlong(0)
        self._check(self.lib.dimhip_advance(self.h, nsteps, cfl, t_end, int(use_graph),
                                            C.byref(t), C.byref(n)))
        return t.value, n.value

    def compute_dt(self, cfl, dt_max):
        dt = C.c_double(0.0)
        self._check(self.lib.dimhip_compute_dt(self.h, cfl, dt_max, C.byref(dt)))
        return dt.value

    def sweep(self, d, dt):
        self._check(self.lib.dimhip_sweep(self.h, d, dt))

    def apply_bc(self):
        self._check(self.lib.dimhip_apply_bc(self.h))

    def conservation(self):
        out = np.zeros(self.nmat + 1 + self.dim + 1)
        self._check(self.lib.dimhip_conservation(self.h, _dp(out)))
        return out

    def boundary_flux(self):
        """[flux mass_k, mass, mom_d, E] through the domain boundary of the current state."""
        out = np.zeros(self.nmat + 1 + self.dim + 1)
        self._check(self.lib.dimhip_boundary_flux(self.h, _dp(out)))
        return out

    def track_boundary_flux(self, enable):
        self._check(self.lib.dimhip_track_boundary_flux(self.h, int(bool(enable))))

    def boundary_flux_integral(self):
        out = np.zeros(self.nmat + 1 + self.dim + 1)
        self._check(self.lib.dimhip_boundary_flux_integral(self.h, _dp(out)))
        return out

    def upload_primitive(self, prim):
        """prim: (ncells, 2*nmat + dim + 1) rows rho[nmat], alpha[nmat], vel[dim], p."""
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        if prim.size != self.ncells * (2 * self.nmat + self.dim + 1):
            raise RuntimeError("dim::advance_one_step: grid size mismatch")
        self._check(self.lib.dimhip_upload_primitive(self.h, _dp(prim), self.ncells))

    def output_rows(self):
        """(ncells, 3 + dim + 3*nmat): rho, vel, p, e, alpha, rho_mat, mass per cell."""
        out = np.empty((self.ncells, 3 + self.dim + 3 * self.nmat))
        self._check(self.lib.dimhip_output_rows(self.h, _dp(out), self.ncells))
        return out

    def synchronize(self):
        self._check(self.lib.dimhip_synchronize(self.h))

    def set_timing(self, on, every=1):
        """Event timing of the kernels of one step in `every` (off when `on` is false)."""
        self._check(self.lib.dimhip_set_timing(self.h, int(every) if on else 0))

    def kernel_times(self):
        """Mean device ms per launch of (sweep, boundary, cfl) kernels and launch counts."""
        ms = np.zeros(3)
        cnt = (C.c_longlong * 3)()
        self._check(self.lib.dimhip_kernel_times(self.h, _dp(ms), cnt))
        return ms, list(cnt)

    def sweep_times(self):
        """Mean device ms per sweep launch along x, y, z and launch counts."""
        ms = np.zeros(3)
        cnt = (C.c_longlong * 3)()
        self._check(self.lib.dimhip_sweep_times(self.h, _dp(ms), cnt))
        return ms, list(cnt)


def kat(which, dim, materials, count, in0, in1=None, d=0, lambda_model="kapila",
        alpha_source_floor=0.0):
    """Run one device function of the path on `count` packed items (see include/dimhip.h)."""
    lib = load_library()
    desc = make_desc(dim, [3] * dim, [1.0] * dim, materials, ["transmissive"] * dim,
                     ["transmissive"] * dim, lambda_model, alpha_source_floor)
    nmat = len(materials)
    nv, np_, nf = 2 * nmat + dim, 2 * nmat + dim + 1, nmat + dim + 1
    in0 = np.ascontiguousarray(in0, dtype=np.float64)
    in1 = np.ascontiguousarray(in1 if in1 is not None else np.zeros(1), dtype=np.float64)
    size0 = {0: np_, 1: nv, 2: nv, 3: nf, 4: nmat, 5: 1, 6: 1, 7: 1}[which] * count
    out0 = np.zeros(max(size0, 1))
    out1 = np.zeros(max(count, 1))
    rc = lib.dimhip_kat(C.byref(desc), which, d, count, _dp(in0), _dp(in1), _dp(out0), _dp(out1))
    if rc != 0:
        raise RuntimeError(lib.dimhip_last_error(None).decode())
    return out0[:size0], out1[:count]


def batch_tuning_key(cases):
    """Identity of a batched instance set for the autotuner cache: every
    instance's grid, spacing, materials, boundaries and lambda model, and the
    ABI version (a new library may rank the configurations differently)."""
    import hashlib
    import json
    desc = [[list(c.N), [float(x) for x in c.dx], [list(m) for m in c.materials], list(c.bc_lo), list(c.bc_hi),
             c.lambda_model, float(c.alpha_source_floor)] for c in cases]
    blob = json.dumps({"abi": load_library().dimhip_version(), "instances": desc}, sort_keys=True)
    return hashlib.sha1(blob.encode()).hexdigest()[:16]


def load_tuned(path, key):
    """The stored autotuner choice for `key`, or None."""
    import json
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def store_tuned(path, key, entry):
    """Store an autotuner choice under `key` (other keys are kept)."""
    import json
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)


class Batch:
    """Many 2D two-material FP32-mixed instances advanced together
    (include/dimhip.h, batched form)."""

    def __init__(self, cases, device=0):
        self.lib = load_library()
        self.cases = list(cases)
        descs = (Desc * len(self.cases))()
        for i, c in enumerate(self.cases):
            descs[i] = make_desc(c.dim, c.N, c.dx, c.materials, c.bc_lo, c.bc_hi, c.lambda_model,
                                 c.alpha_source_floor, "split", "fp32_mixed", device)
        h = C.c_void_p()
        if self.lib.dimhip_batch_create(descs, len(self.cases), C.byref(h)) != 0:
            raise RuntimeError(self.lib.dimhip_batch_last_error(None).decode())
        self.h = h
        self.times = np.zeros(len(self.cases))
        for i, c in enumerate(self.cases):
            self.upload(i, c.U)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.dimhip_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.dimhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, i, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self._check(self.lib.dimhip_batch_upload(self.h, i, _dp(U), self.cases[i].ncells))

    def download(self, i):
        out = np.empty((self.cases[i].ncells, self.cases[i].nv))
        self._check(self.lib.dimhip_batch_download(self.h, i, _dp(out), self.cases[i].ncells))
        return out

    def configure(self, x_variant, y_variant, schedule, blocks_per_cu):
        self._check(self.lib.dimhip_batch_configure(self.h, x_variant, y_variant, schedule, blocks_per_cu))

    def autotune(self, cfl, trial_steps=3, repeats=1, cache=None):
        """Choose the tile shapes, schedule and blocks per CU.  With `cache` (a
        JSON file path) a choice stored for this instance set is applied
        without timing; otherwise the device autotuner times every
        configuration `repeats` times over `trial_steps` steps (each time from
        the current state, which it restores), the configuration with the
        smallest of its times is applied, and with `cache` it is stored."""
        import json
        key = batch_tuning_key(self.cases)
        if cache is not None:
            hit = load_tuned(cache, key)
            if hit is not None:
                b = hit["best"]
                self.configure(b["x"], b["y"], b["schedule"], b["blocks_per_cu"])
                return dict(hit, cached=True, key=key)
        best_ms = {}
        report = None
        for _ in range(max(1, repeats)):
            buf = C.create_string_buffer(16384)
            self._check(self.lib.dimhip_batch_autotune(self.h, cfl, trial_steps, buf, len(buf)))
            report = json.loads(buf.value.decode())
            for t in report["trials"]:
                k = (t["x"], t["y"], t["schedule"], t["blocks_per_cu"])
                best_ms[k] = min(best_ms.get(k, float("inf")), t["ms"])
        k = min(best_ms, key=best_ms.get)
        self.configure(*k)
        report["best"] = {"x": k[0], "y": k[1], "schedule": k[2], "blocks_per_cu": k[3], "ms": best_ms[k]}
        report.update(trial_steps=trial_steps, repeats=repeats, cached=False, key=key)
        if cache is not None:
            store_tuned(cache, key, {"best": report["best"], "trial_steps": trial_steps, "repeats": repeats,
                                     "instances": len(self.cases),
                                     "cells": int(sum(c.ncells for c in self.cases))})
        return report

    def advance(self, nsteps, cfl, t_end=1e300, use_graph=True):
        steps = (C.c_longlong * len(self.cases))()
        self._check(self.lib.dimhip_batch_advance(self.h, nsteps, cfl, t_end, int(use_graph),
                                                  _dp(self.times), steps))
        return self.times.copy(), list(steps)

    def set_timing(self, on):
        self._check(self.lib.dimhip_batch_set_timing(self.h, int(on)))

    def kernel_times(self):
        ms = np.zeros(3)
        cnt = (C.c_longlong * 3)()
        self._check(self.lib.dimhip_batch_kernel_times(self.h, _dp(ms), cnt))
        return ms, list(cnt)

    def sweep_times(self):
        ms = np.zeros(3)
        cnt = (C.c_longlong * 3)()
        self._check(self.lib.dimhip_batch_sweep_times(self.h, _dp(ms), cnt))
        return ms, list(cnt)
