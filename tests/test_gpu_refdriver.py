"""The reference's own driver on the MI355X library: the drop-in check.

oracle/_ref/euler3d_ref is the reference's euler3d.cpp compiled where it lies
against this project's include/op_seq.h + include/op_hdf5.h and linked with
libmgcfd (oracle/Makefile, target refdriver).  Its constants are the
reference's compiled-in freestream (const.h:32-33: Mach 1.2, angle 0).  The
oracle solves the same deck and writes the solution files the reference's
-v path reads (euler3d.cpp:314-335); the reference's own validation
(euler3d.cpp:658-718: 1e-7 relative, at most N/5000 misses per level) must
pass, and its output files (-o prefix, variables) must match the oracle to
1e-12 relative per variable.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import Oracle
from test_gpu_driver import BIN, h5, h5_read, read_deck

pytestmark = pytest.mark.gpu

REFDRV = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "euler3d_ref")
MACH, AOA = 1.2, 0.0          # const.h:32-33, compiled into the reference driver

needs_refdrv = pytest.mark.skipif(not os.path.exists(REFDRV), reason="oracle/_ref/euler3d_ref not built")


@pytest.fixture(scope="module")
def refdeck(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("refdeck"))
    subprocess.run([os.path.join(BIN, "mgcfd_meshgen"), "--knn", "5000", "15500", "--levels", "3", "--seed", "9",
                    "-o", d], check=True, capture_output=True)
    L = h5()
    mesh = read_deck(L, d, 3)
    cycles = 4
    ref = Oracle("orc").solve(mesh, cycles, MACH, AOA)
    assert ref["status"] == 0
    for l, lv in enumerate(ref["levels"]):
        v = np.ascontiguousarray(lv["variables"])
        f = os.path.join(d, f"solution.variables.L{l}.cycles={cycles}.h5")
        assert L.mgcfd_h5_write(f.encode(), f"p_variables_result_L{l}".encode(), b"double", v.ctypes.data,
                                v.shape[0], 5, 1) == 0
    return d, cycles, ref, L


@needs_refdrv
def test_reference_driver_validates(refdeck, tmp_path):
    d, cycles, ref, L = refdeck
    prefix = str(tmp_path) + "/"
    p = subprocess.run([REFDRV, "-i", "input.dat", "-d", d, "-g", str(cycles), "-v", "-o", prefix,
                        "--output-variables"], capture_output=True, text=True, timeout=300)
    out = p.stdout + p.stderr
    print(out[-3000:])
    assert p.returncode == 0, out
    assert "Validation passed" in out, out
    for l in range(3):
        g = h5_read(L, f"{prefix}variables.L{l}.cycles={cycles}.h5", f"p_variables_result_L{l}")
        o = ref["levels"][l]["variables"]
        err = (np.abs(g - o).max(axis=0) / np.abs(o).max(axis=0)).max()
        assert err <= 1e-12, (l, err)
