"""Input-deck forms of the drop-in driver that the default deck does not use.

- Legacy decks (-l, euler3d.cpp:243-336): the '.renumbered' dataset names,
  the per-node volumes read from 'areas' and the weights from
  'edge_weights.recalculated', with the cell-volume setup loop skipped
  (euler3d.cpp:422-431).  The deck's areas and recalculated weights are what
  calculate_cell_volumes (misc.h:40-76) makes of the plain deck, so both
  decks must give the plain deck's solution.
- 1-based maps: input.dat's base_array_index=1 (the reference's default,
  euler3d.cpp:92), forwarded to op_init as OP_MAPS_BASE_INDEX=1
  (euler3d.cpp:104-117), with every map value in the HDF5 files one higher.
- Output files (euler3d.cpp:720-779): --output-fluxes, --output-step-factors
  and --output-variables, compared with the oracle's arrays after the same
  cycles.

Each runs through this project's driver (euler3d_mi355x) and through the
reference's own euler3d.cpp built against this library (oracle/_ref/
euler3d_ref, Mach 1.2 at 0 degrees compiled in), with the reference's -v
check (1e-7 relative, N/5000 misses) and, for this project's driver, the
strict 1e-12 check.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import Oracle
from test_gpu_driver import BIN, h5, h5_read, read_deck

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFDRV = os.path.join(ROOT, "oracle", "_ref", "euler3d_ref")
needs_refdrv = pytest.mark.skipif(not os.path.exists(REFDRV), reason="oracle/_ref/euler3d_ref not built")
LEVELS, CYCLES = 3, 3
FLOWS = {"mi355x": (0.5, 3.0), "ref": (1.2, 0.0)}    # the reference driver's compiled-in freestream


def h5_write(L, f, ds, a, first):
    a = np.ascontiguousarray(a)
    typ = b"int" if a.dtype == np.int32 else b"double"
    rows = a.shape[0]
    cols = a.shape[1] if a.ndim == 2 else 1
    assert L.mgcfd_h5_write(f.encode(), ds.encode(), typ, a.ctypes.data, rows, cols, 1 if first else 0) == 0, \
        L.mgcfd_h5_last_error()


def write_deck(L, d, mesh, base=0, legacy=None):
    """A deck of 'mesh' in directory d: map values offset by 'base'; with
    legacy = per-level (areas, recalculated weights), the legacy names."""
    os.makedirs(d, exist_ok=True)
    rn = ".renumbered" if legacy is not None else ""
    for l, m in enumerate(mesh):
        f = os.path.join(d, f"level{l}.h5")
        h5_write(L, f, "node_coordinates" + rn, np.asarray(m["coords"], np.float64), True)
        h5_write(L, f, "edge-->node" + rn, np.asarray(m["edge_to_node"], np.int32) + base, False)
        h5_write(L, f, "bnd_node-->node" + rn, np.asarray(m["bnd_to_node"], np.int32).reshape(-1, 1) + base, False)
        h5_write(L, f, "bnd_node-->group", np.asarray(m["bnd_group"], np.int32).reshape(-1, 1), False)
        h5_write(L, f, "bnd_node_weights", np.asarray(m["bnd_weights"], np.float64), False)
        if legacy is not None:
            h5_write(L, f, "areas", legacy[l][0].reshape(-1, 1), False)
            h5_write(L, f, "edge_weights.recalculated", legacy[l][1], False)
        else:
            h5_write(L, f, "edge_weights", np.asarray(m["edge_weights"], np.float64), False)
        if l + 1 < len(mesh):
            h5_write(L, f, "node-->mg_node" + rn, np.asarray(m["node_to_mg_node"], np.int32).reshape(-1, 1) + base,
                     False)
    with open(os.path.join(d, "input.dat"), "w") as fh:
        fh.write(f"size={len(mesh[0]['coords'])}\nnum_levels={len(mesh)}\nbase_array_index={base}\n"
                 "mesh_name=m6wing\n[levels]\n")
        for l in range(len(mesh)):
            fh.write(f"{l}=level{l}.h5\n")


def write_solution(L, d, ref, cycles):
    for l, lv in enumerate(ref["levels"]):
        h5_write(L, os.path.join(d, f"solution.variables.L{l}.cycles={cycles}.h5"), f"p_variables_result_L{l}",
                 lv["variables"], True)


def run_driver(which, d, *args):
    exe = os.path.join(BIN, "euler3d_mi355x") if which == "mi355x" else REFDRV
    extra = []
    if which == "mi355x":
        mach, aoa = FLOWS["mi355x"]
        extra = ["--mach", str(mach), "--aoa", str(aoa)]
    p = subprocess.run([exe, "-i", "input.dat", "-d", d, "-g", str(CYCLES), *extra, *args],
                       capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout + p.stderr


@pytest.fixture(scope="module")
def plain(tmp_path_factory):
    """A generated deck, its mesh and the oracle's solves at both freestreams."""
    d = str(tmp_path_factory.mktemp("plain"))
    subprocess.run([os.path.join(BIN, "mgcfd_meshgen"), "--knn", "4000", "12400", "--levels", str(LEVELS),
                    "--seed", "21", "-o", d], check=True, capture_output=True)
    L = h5()
    mesh = read_deck(L, d, LEVELS)
    orc = Oracle("orc")
    refs = {k: orc.solve(mesh, CYCLES, mach, aoa) for k, (mach, aoa) in FLOWS.items()}
    assert all(r["status"] == 0 for r in refs.values())
    return mesh, refs, L


def legacy_arrays(mesh):
    """What calculate_cell_volumes (misc.h:40-76) makes of each level: its
    volumes and rewritten edge weights (before the damping loop)."""
    orc = Oracle("orc")
    out = []
    for m in mesh:
        vol = np.zeros(len(m["coords"]))
        ew = np.array(m["edge_weights"], dtype=np.float64, copy=True)
        orc.calculate_cell_volumes(np.ascontiguousarray(m["edge_to_node"], np.int32),
                                   np.ascontiguousarray(m["coords"], np.float64), ew, vol)
        out.append((vol, ew))
    return out


@pytest.mark.parametrize("which", ["mi355x", pytest.param("ref", marks=needs_refdrv)])
def test_legacy_deck(plain, tmp_path, which):
    mesh, refs, L = plain
    d = str(tmp_path / "legacy")
    write_deck(L, d, mesh, legacy=legacy_arrays(mesh))
    write_solution(L, d, refs[which], CYCLES)
    args = ["-l", "-v"] + (["--validate-strict"] if which == "mi355x" else [])
    rc, out = run_driver(which, d, *args)
    print(out[-2000:])
    assert rc == 0 and "Validation passed" in out, out
    # the plain names are absent from a legacy deck: without -l it cannot load
    rc, out = run_driver(which, d)
    assert rc != 0, out


@pytest.mark.parametrize("which", ["mi355x", pytest.param("ref", marks=needs_refdrv)])
def test_one_based_deck(plain, tmp_path, which):
    mesh, refs, L = plain
    d = str(tmp_path / "base1")
    write_deck(L, d, mesh, base=1)
    write_solution(L, d, refs[which], CYCLES)
    args = ["-v"] + (["--validate-strict"] if which == "mi355x" else [])
    rc, out = run_driver(which, d, *args)
    print(out[-2000:])
    assert rc == 0 and "Validation passed" in out, out


@pytest.mark.parametrize("which,device_loop", [("mi355x", False), ("mi355x", True),
                                               pytest.param("ref", False, marks=needs_refdrv)])
def test_output_files_match_oracle(plain, tmp_path, which, device_loop):
    mesh, refs, L = plain
    d = str(tmp_path / "deck")
    write_deck(L, d, mesh)
    prefix = str(tmp_path / "out.")
    # The reference sets these flags through int pointers to bool fields
    # (config.h:122-124 over config.h:90-98), so each flag's 4-byte store
    # also clears the next fields: --output-fluxes given after
    # --output-variables turns the latter off again.  This order survives.
    args = ["-o", prefix, "--output-step-factors", "--output-fluxes", "--output-variables"]
    if device_loop:
        args.append("--device-loop")
    rc, out = run_driver(which, d, *args)
    assert rc == 0, out
    ref = refs[which]
    for l in range(LEVELS):
        sfx = f".L{l}.cycles={CYCLES}.h5"
        var = h5_read(L, f"{prefix}variables{sfx}", f"p_variables_result_L{l}")
        flux = h5_read(L, f"{prefix}fluxes{sfx}", f"p_fluxes_result_L{l}")
        sf = h5_read(L, f"{prefix}step_factors{sfx}", f"p_step_factors_result_L{l}")
        o = ref["levels"][l]
        e_var = (np.abs(var - o["variables"]).max(axis=0) / np.abs(o["variables"]).max(axis=0)).max()
        e_sf = np.abs(sf[:, 0] - o["step_factors"]).max() / np.abs(o["step_factors"]).max()
        print(f"{which} L{l}: variables {e_var:.2e}, step factors {e_sf:.2e}, |flux| max {np.abs(flux).max():.2e}")
        assert e_var <= 1e-12 and e_sf <= 1e-12, (l, e_var, e_sf)
        # the last Runge-Kutta stage leaves the flux zeroed (time_stepping_kernels.h:81-85)
        assert np.array_equal(flux, o["fluxes"]) and not flux.any()
