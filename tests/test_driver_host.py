"""Host side of the case driver (diffuse_vs_gfm_mms_amd/host/), on the CPU.

* Case files: the driver's loader accepts the reference's format and rejects
  what the reference rejects, with the reference's messages.
* Initial states: the driver builds primitive rows on the host and the device
  applies prim_to_cons.  Here the rows go through the oracle's prim_to_cons
  and must equal the reference's own initial state
  (load_config + initialise_dim_from_config, src/setup/dim_initial_conditions.hpp)
  bit for bit, for every initial-condition kind.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle.pyoracle import Checker, make_params, ref_available, ref_initial_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "diffuse_vs_gfm_mms_amd", "host")
CASES = os.path.join(ROOT, "tests", "cases")

CASE_FILES = {
    "tube_1d_diffuse.txt": 1,
    "water_air_1d_sharp.txt": 1,
    "helium_bubble_2d.txt": 2,
    "oblique_2d.txt": 2,
    "coated_bubble_2d.txt": 2,
    "box_regions_2d.txt": 2,
    "bubble_collapse_3d.txt": 3,
    "gorsse_tc9_2d.txt": 2,
    "three_material_1d.txt": 1,
}


def materials_of(path):
    mats = {}
    for line in open(path):
        line = line.split("#")[0].strip()
        if line.startswith("material ") or line.startswith("material="):
            parts = [t.strip() for t in line.split("=", 1)[1].split(",")]
            kv = dict(t.split("=") for t in parts[2:])
            kind = parts[1]
            # p_inf is read for stiffened gas only, the co-volume (key b, else
            # covolume) for noble-Abel only (material_builder.hpp:21-38)
            pinf = float(kv.get("p_inf", 0.0)) if kind == "stiffened_gas" else 0.0
            b = float(kv.get("b", kv.get("covolume", 0.0))) if kind in ("noble_abel", "noble_abel_gas",
                                                                         "real_gas") else 0.0
            mats[int(parts[0])] = (kind, float(kv["gamma"]), pinf, b)
    return [mats[k] for k in sorted(mats)]


@pytest.fixture(scope="module")
def dumpers(tmp_path_factory):
    out = {}
    d = tmp_path_factory.mktemp("icdump")
    for dim in (1, 2, 3):
        exe = str(d / f"ic_dump_{dim}d")
        subprocess.run(["g++", "-std=c++17", "-O2", f"-DDIM_={dim}", "-I", HOST,
                        os.path.join(ROOT, "tests", "host_ic_dump.cpp"), "-o", exe, "-pthread"], check=True)
        out[dim] = exe
    return out


def case_copy(tmp_path, name, **subst):
    text = open(os.path.join(CASES, name)).read().replace("OUTDIR", str(tmp_path / "out"))
    for k, v in subst.items():
        text = text.replace(k, v)
    p = tmp_path / name
    p.write_text(text)
    return str(p)


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", sorted(CASE_FILES))
def test_initial_state_matches_reference(tmp_path, dumpers, oracle, name):
    dim = CASE_FILES[name]
    path = case_copy(tmp_path, name)
    out = tmp_path / "prim.bin"
    subprocess.run([dumpers[dim], path, "0", str(out)], check=True)
    mats = materials_of(path)
    prim = np.fromfile(out, dtype=np.float64).reshape(-1, 2 * len(mats) + dim + 1)
    U = oracle.prim_to_cons(make_params(dim, mats), prim)
    U_ref = ref_initial_state(path, dim, 0)
    assert U.shape == U_ref.shape
    assert np.array_equal(U.view(np.uint64), U_ref.view(np.uint64)), name


@pytest.mark.parametrize("edit,message", [
    (("N = [128, 36]", "N = [128]"), "Wrong int array size"),
    (("cfl = 0.8", "cfl = 0.8\nbogus_key = 1"), "Unknown key: bogus_key"),
    (("bc_hi = [transmissive, reflective]", "bc_hi = [transmissive, mirror]"), "Invalid boundary condition"),
    (("output_times = [0.00002, 0.00004]", "output_times = [0.00004, 0.00002]"),
     "output_times must be sorted in ascending order"),
    (("material = 1, ideal_gas, gamma=1.648", "material = 2, ideal_gas, gamma=1.648"),
     "material ids must be contiguous and zero-based: expected ids 0..1"),
    (("bubble_radius = 0.025", "bubble_radius = 0.0"), "Invalid bubble radius"),
    (("interface_method = DIM", "interface_method = DIM\nuse_level_set = true"), "DIM should not use level set"),
    (("dimension = 2", "dimension = 3"), "Dimension mismatch"),
    (("shock_axis = 0", "shock_axis = 0\nmodel_type = plasma"), "Unknown model_type: plasma"),
])
def test_case_file_errors_match_reference(tmp_path, dumpers, edit, message):
    """The loader's error text is the reference's (config_loader.hpp:237-570);
    where the reference build is present, it fails with the same message."""
    path = case_copy(tmp_path, "helium_bubble_2d.txt", **{edit[0]: edit[1]})
    res = subprocess.run([dumpers[2], path, "0", str(tmp_path / "x.bin")], capture_output=True, text=True)
    assert res.returncode == 1
    assert res.stderr.strip() == "ERROR: " + message
    if ref_available():
        with pytest.raises(RuntimeError) as e:
            ref_initial_state(path, 2, 0)
        assert str(e.value) == message
