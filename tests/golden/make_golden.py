"""Writes the golden fixtures from the reference's own DIM functions.

Runs in the container that holds the reference tree: it loads
oracle/_ref/libdimref.so (built by `make -C oracle ref` from the reference's
headers) and records inputs and outputs as .npz files next to this script.
The fixtures are data; nothing here is reference source.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle.pyoracle import Checker, make_params  # noqa: E402
from _states import MATSETS, random_line, random_states  # noqa: E402
from diffuse_vs_gfm_mms_amd import workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def functions(ref, rng):
    data = {}
    for dim in (1, 2, 3):
        for ms in ("water_air", "air_helium", "water_air_helium"):
            for lm in ("kapila", "allaire", "equal_velocity"):
                mats = MATSETS[ms]
                p = make_params(dim, mats, lm)
                key = f"d{dim}_{ms}_{lm}"
                U = random_states(mats, dim, 300, rng)
                V = random_states(mats, dim, 300, rng)
                P = ref.cons_to_prim(p, U)
                data[key + "_U"] = U
                data[key + "_V"] = V
                data[key + "_c2p"] = P
                data[key + "_p2c"] = ref.prim_to_cons(p, P)
                data[key + "_repair"] = ref.repair_state(p, U)
                data[key + "_rhs"] = ref.alpha_rhs(p, P)
                for d in range(dim):
                    F, fv = ref.hllc(p, d, U, V)
                    data[key + f"_hllc{d}_F"] = F
                    data[key + f"_hllc{d}_fv"] = fv
    np.savez_compressed(os.path.join(OUT, "functions.npz"), **data)


def lines(ref, rng):
    data = {}
    for dim in (1, 2, 3):
        for ms in ("water_air", "air_helium"):
            mats = MATSETS[ms]
            p = make_params(dim, mats)
            for n in (1, 3, 4, 7, 16, 64, 100):
                for d in range(dim):
                    key = f"d{dim}_{ms}_n{n}_dir{d}"
                    U = random_line(mats, dim, n, rng)
                    data[key + "_U"] = U
                    data[key + "_out"] = ref.solve_line(p, d, U, 1e-3, 2e-8)
    np.savez_compressed(os.path.join(OUT, "lines.npz"), **data)


def steps(ref):
    data = {}
    c = W.water_air_tube_1d(200)
    p = make_params(1, c.materials)
    U = c.U.copy()
    data["tube1d_U0"] = U.copy()
    done = 0
    for k in (1, 10, 50):
        U, dts, _ = ref.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, U, c.cfl, 1.0, nsteps=k - done)
        data[f"tube1d_U{k}"] = U.copy()
        data[f"tube1d_dt{k}"] = dts
        done = k
    c = W.helium_shock_bubble_2d(40, 20)
    data["bubble2d_U0"] = c.U.copy()
    combos = {
        "tr_refl": (c.bc_lo, c.bc_hi, False),
        "mixed": (("reflective", "reflective"), ("nonreflective", "reflection"), False),
        "unsplit": (c.bc_lo, c.bc_hi, True),
    }
    p = make_params(2, c.materials)
    for name, (lo, hi, unsplit) in combos.items():
        U, dts, _ = ref.advance(p, c.N, c.dx, lo, hi, c.U, c.cfl, 1.0, nsteps=5, unsplit=unsplit)
        data[f"bubble2d_{name}_U5"] = U
        data[f"bubble2d_{name}_dt"] = dts
    c = W.bubble_collapse_3d(10)
    p = make_params(3, c.materials)
    data["collapse3d_U0"] = c.U.copy()
    U, dts, _ = ref.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, c.U, c.cfl, 1.0, nsteps=2)
    data["collapse3d_U2"] = U
    data["collapse3d_dt"] = dts
    np.savez_compressed(os.path.join(OUT, "steps.npz"), **data)


def noble_abel(ref, rng):
    """Noble-Abel material sets (water + noble-Abel air, and with helium as a
    third material): functions, lines and a few steps of a coarse Gorsse TC9
    water/air bubble collapse and of a 1D water/noble-Abel-air tube."""
    data = {}
    for dim in (1, 2, 3):
        for ms in ("water_na_air", "water_na_air_helium", "na_air"):
            for lm in ("kapila", "allaire", "equal_velocity"):
                mats = MATSETS[ms]
                p = make_params(dim, mats, lm)
                key = f"d{dim}_{ms}_{lm}"
                U = random_states(mats, dim, 300, rng)
                V = random_states(mats, dim, 300, rng)
                P = ref.cons_to_prim(p, U)
                data[key + "_U"] = U
                data[key + "_V"] = V
                data[key + "_c2p"] = P
                data[key + "_p2c"] = ref.prim_to_cons(p, P)
                data[key + "_repair"] = ref.repair_state(p, U)
                data[key + "_rhs"] = ref.alpha_rhs(p, P)
                for d in range(dim):
                    F, fv = ref.hllc(p, d, U, V)
                    data[key + f"_hllc{d}_F"] = F
                    data[key + f"_hllc{d}_fv"] = fv
        mats = MATSETS["water_na_air"]
        p = make_params(dim, mats)
        for n in (3, 16, 64):
            for d in range(dim):
                key = f"line_d{dim}_n{n}_dir{d}"
                U = random_line(mats, dim, n, rng)
                data[key + "_U"] = U
                data[key + "_out"] = ref.solve_line(p, d, U, 1e-3, 2e-8)
    c = W.gorsse_tc9_2d(48, 40)
    p = make_params(2, c.materials)
    data["gorsse_U0"] = c.U.copy()
    U, dts, _ = ref.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, c.U, c.cfl, 1.0, nsteps=5)
    data["gorsse_U5"] = U
    data["gorsse_dt"] = dts
    c = W.water_na_air_tube_1d(200)
    p = make_params(1, c.materials)
    data["tube_U0"] = c.U.copy()
    U, dts, _ = ref.advance(p, c.N, c.dx, c.bc_lo, c.bc_hi, c.U, c.cfl, 1.0, nsteps=20)
    data["tube_U20"] = U
    data["tube_dt"] = dts
    np.savez_compressed(os.path.join(OUT, "noble_abel.npz"), **data)


def main():
    ref = Checker("ref")
    rng = np.random.default_rng(20261203)
    functions(ref, rng)
    lines(ref, rng)
    steps(ref)
    noble_abel(ref, np.random.default_rng(20261204))
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
