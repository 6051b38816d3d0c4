"""End-to-end drop-in test: the euler3d_mi355x driver on an HDF5 input deck.

The deck comes from mgcfd_meshgen; the oracle solves the same mesh (read
back from the deck's HDF5 files) and writes the reference's solution files
(solution.variables.L<l>.cycles=<g>.h5, dataset p_variables_result_L<l>,
euler3d.cpp:314-335); the driver then runs with -v (the reference's
validation: 1e-7 relative, N/5000 misses) plus --validate-strict (1e-12
relative per variable), loop by loop and with the device-resident loop.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mgcfd
from oracle import Oracle

pytestmark = pytest.mark.gpu

BIN = os.path.join(mgcfd.PKG_DIR, "bin")
H5LIB = os.path.join(mgcfd.PKG_DIR, "lib", "libmgcfd_hdf5.so")


def h5():
    L = C.CDLL(H5LIB)
    L.mgcfd_h5_shape.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.mgcfd_h5_read.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_longlong, C.c_longlong]
    L.mgcfd_h5_write.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_longlong, C.c_longlong, C.c_int]
    L.mgcfd_h5_last_error.restype = C.c_char_p
    return L


def h5_read(L, f, ds):
    r, c = C.c_longlong(0), C.c_longlong(0)
    assert L.mgcfd_h5_shape(f.encode(), ds.encode(), C.byref(r), C.byref(c)) == 0, L.mgcfd_h5_last_error()
    is_int = ds in ("edge-->node", "bnd_node-->node", "bnd_node-->group", "node-->mg_node")
    a = np.empty((r.value, c.value), dtype=np.int32 if is_int else np.float64)
    assert L.mgcfd_h5_read(f.encode(), ds.encode(), b"int" if is_int else b"double", a.ctypes.data,
                           r.value, c.value) == 0
    return a


def read_deck(L, d, levels):
    mesh = []
    for l in range(levels):
        f = os.path.join(d, f"level{l}.h5")
        m = dict(coords=h5_read(L, f, "node_coordinates"), edge_to_node=h5_read(L, f, "edge-->node"),
                 edge_weights=h5_read(L, f, "edge_weights"), bnd_to_node=h5_read(L, f, "bnd_node-->node")[:, 0],
                 bnd_group=h5_read(L, f, "bnd_node-->group")[:, 0], bnd_weights=h5_read(L, f, "bnd_node_weights"),
                 node_to_mg_node=(h5_read(L, f, "node-->mg_node")[:, 0] if l + 1 < levels else None))
        mesh.append(m)
    return mesh


@pytest.fixture(scope="module")
def deck(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("deck"))
    subprocess.run([os.path.join(BIN, "mgcfd_meshgen"), "--knn", "4000", "12400", "--levels", "3", "--seed", "5",
                    "-o", d], check=True, capture_output=True)
    L = h5()
    mesh = read_deck(L, d, 3)
    cycles = 3
    ref = Oracle("orc").solve(mesh, cycles, 0.5, 3.0)
    assert ref["status"] == 0
    for l, lv in enumerate(ref["levels"]):
        v = np.ascontiguousarray(lv["variables"])
        f = os.path.join(d, f"solution.variables.L{l}.cycles={cycles}.h5")
        assert L.mgcfd_h5_write(f.encode(), f"p_variables_result_L{l}".encode(), b"double", v.ctypes.data,
                                v.shape[0], 5, 1) == 0
    return d, cycles, ref, L


def run(d, *args):
    p = subprocess.run([os.path.join(BIN, "euler3d_mi355x"), "-i", "input.dat", "-d", d, *args],
                       capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout + p.stderr


@pytest.mark.parametrize("mode", [[], ["--device-loop"]])
def test_driver_validates_against_oracle(deck, mode):
    d, cycles, _, _ = deck
    rc, out = run(d, "-g", str(cycles), "--mach", "0.5", "--aoa", "3", "-v", "--validate-strict", *mode)
    assert rc == 0, out
    assert "Validation passed" in out, out
    assert "Performing MG cycle 3 / 3" in out


def test_driver_output_files_match_oracle(deck, tmp_path):
    d, cycles, ref, L = deck
    prefix = str(tmp_path) + "/run."
    rc, out = run(d, "-g", str(cycles), "--mach", "0.5", "--aoa", "3", "--output-variables",
                  "--output-step-factors", "-o", prefix)
    assert rc == 0, out
    for l in range(3):
        g = h5_read(L, f"{prefix}variables.L{l}.cycles={cycles}.h5", f"p_variables_result_L{l}")
        o = ref["levels"][l]["variables"]
        err = (np.abs(g - o).max(axis=0) / np.abs(o).max(axis=0)).max()
        assert err <= 1e-12, (l, err)


def test_driver_detects_wrong_solution(deck, tmp_path):
    """The -v path must fail when the solution differs beyond 1e-7 in more than N/5000 values."""
    d, cycles, ref, L = deck
    bad = str(tmp_path)
    for f in os.listdir(d):
        if not f.startswith("solution"):
            os.symlink(os.path.join(d, f), os.path.join(bad, f))
    for l, lv in enumerate(ref["levels"]):
        v = np.ascontiguousarray(lv["variables"] * (1 + 1e-6))
        f = os.path.join(bad, f"solution.variables.L{l}.cycles={cycles}.h5")
        L.mgcfd_h5_write(f.encode(), f"p_variables_result_L{l}".encode(), b"double", v.ctypes.data, v.shape[0], 5, 1)
    rc, out = run(bad, "-g", str(cycles), "--mach", "0.5", "--aoa", "3", "-v")
    assert "Validation failed" in out, out


def test_driver_flow_interval_snapshots(deck, tmp_path):
    d, cycles, ref, L = deck
    prefix = str(tmp_path) + "/snap."
    rc, out = run(d, "-g", str(cycles), "--mach", "0.5", "--aoa", "3", "-I", "1", "-o", prefix, "--device-loop")
    assert rc == 0, out
    last = h5_read(L, f"{prefix}variables.L0.cycle={cycles}.h5", "p_variables")
    o = ref["levels"][0]["variables"]
    assert (np.abs(last - o).max(axis=0) / np.abs(o).max(axis=0)).max() <= 1e-12


def test_driver_down_v2_matches_oracle(deck, tmp_path):
    """--down-v2: the driver declares edge-->mg_nodes (euler3d.cpp:286-292)
    and prolongs with the inverse-distance loops (euler3d.cpp:603-623); the
    final state of every level matches the oracle's run of the same loops."""
    d, cycles, _, L = deck
    ref = Oracle("orc").solve(read_deck(L, d, 3), cycles, 0.5, 3.0, down_v2=True)
    assert ref["status"] == 0
    prefix = str(tmp_path) + "/v2."
    rc, out = run(d, "-g", str(cycles), "--mach", "0.5", "--aoa", "3", "--down-v2", "--output-variables",
                  "-o", prefix)
    assert rc == 0, out
    for l in range(3):
        g = h5_read(L, f"{prefix}variables.L{l}.cycles={cycles}.h5", f"p_variables_result_L{l}")
        o = ref["levels"][l]["variables"]
        err = (np.abs(g - o).max(axis=0) / np.abs(o).max(axis=0)).max()
        assert err <= 1e-12, (l, err)
    rc, out = run(d, "-g", "1", "--down-v2", "--device-loop")
    assert rc != 0 and "--down-v2 runs loop by loop" in out, out
