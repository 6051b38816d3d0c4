"""Host-side logic without a GPU: the seeded mesh generators that build the
benchmark workloads, the deck writer, and the HDF5 layer of the OP2 surface.

The generators stand in for the reference's downloaded meshes (its decks are
an external download); the tests pin what the configurations state: node,
edge and boundary counts, level ratios, valid and complete maps, no
self-loops and seeded determinism.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mgcfd

BIN = os.path.join(mgcfd.PKG_DIR, "bin")
H5LIB = os.path.join(mgcfd.PKG_DIR, "lib", "libmgcfd_hdf5.so")


def check_levels(mesh):
    for l, m in enumerate(mesh):
        n = len(m["coords"])
        e2n = m["edge_to_node"]
        assert e2n.min() >= 0 and e2n.max() < n
        assert np.all(e2n[:, 0] != e2n[:, 1]), "self-loop"
        key = np.sort(e2n, axis=1)
        assert len(np.unique(key[:, 0] * n + key[:, 1])) == len(e2n), "duplicate edge"
        assert np.all(np.bincount(e2n.ravel(), minlength=n) > 0), "isolated node"
        b = m["bnd_to_node"]
        assert b.min() >= 0 and b.max() < n
        assert set(np.unique(m["bnd_group"])) <= {0, 3}
        assert np.isfinite(m["edge_weights"]).all() and np.isfinite(m["bnd_weights"]).all()
        if l + 1 < len(mesh):
            f2c = m["node_to_mg_node"]
            nc = len(mesh[l + 1]["coords"])
            assert len(f2c) == n and f2c.min() >= 0 and f2c.max() < nc
            assert len(np.unique(f2c)) == nc, "coarse node without children"
        else:
            assert m["node_to_mg_node"] is None


def test_box_counts_and_levels():
    mesh = mgcfd.mesh_box(40, 4, seed=1)
    sizes = [(len(m["coords"]), len(m["edge_to_node"]), len(m["bnd_to_node"])) for m in mesh]
    # 40^3 / 20^3 / 10^3 / 5^3 nodes, 3 n^2 (n-1) axis edges, 6 n^2 boundary entries
    assert sizes == [(64000, 187200, 9600), (8000, 22800, 2400), (1000, 2700, 600), (125, 300, 150)]
    assert (mesh[0]["bnd_group"] == 0).sum() == 1600          # the z = 0 wall face
    check_levels(mesh)


def test_knn_counts_ratios_and_maps():
    mesh = mgcfd.mesh_knn(30000, 93000, 4, seed=5)
    n = [len(m["coords"]) for m in mesh]
    assert n[0] == 30000 and len(mesh[0]["edge_to_node"]) == 93000
    for got, r in zip(n[1:], [0.55, 0.37, 0.27]):
        assert abs(got - 30000 * r) <= 0.02 * 30000 * r
    deg = np.bincount(mesh[0]["edge_to_node"].ravel(), minlength=n[0])
    assert deg.min() >= 3
    check_levels(mesh)


def test_generators_are_seeded():
    a = mgcfd.mesh_knn(2000, 6200, 3, seed=9)
    b = mgcfd.mesh_knn(2000, 6200, 3, seed=9)
    c = mgcfd.mesh_knn(2000, 6200, 3, seed=10)
    for x, y in zip(a, b):
        for k in x:
            if x[k] is not None:
                assert np.array_equal(x[k], y[k]), k
    assert not np.array_equal(a[0]["coords"], c[0]["coords"])


def test_bad_generator_arguments_fail_cleanly():
    with pytest.raises(mgcfd.MgcfdError):
        mgcfd.mesh_box(2, 4, seed=1)
    with pytest.raises(mgcfd.MgcfdError):
        mgcfd.mesh_knn(100, 10, 2, seed=1)


def h5():
    L = C.CDLL(H5LIB)
    L.mgcfd_h5_shape.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.mgcfd_h5_read.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_longlong, C.c_longlong]
    L.mgcfd_h5_write.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_longlong, C.c_longlong,
                                 C.c_int]
    L.mgcfd_h5_last_error.restype = C.c_char_p
    return L


def test_hdf5_round_trip_and_errors(tmp_path):
    L = h5()
    f = str(tmp_path / "a.h5").encode()
    x = np.arange(12, dtype=np.float64).reshape(4, 3)
    i = np.arange(8, dtype=np.int32).reshape(4, 2)
    assert L.mgcfd_h5_write(f, b"xs", b"double", x.ctypes.data, 4, 3, 1) == 0
    assert L.mgcfd_h5_write(f, b"edge-->node", b"int", i.ctypes.data, 4, 2, 0) == 0
    r, c = C.c_longlong(), C.c_longlong()
    assert L.mgcfd_h5_shape(f, b"xs", C.byref(r), C.byref(c)) == 0 and (r.value, c.value) == (4, 3)
    y = np.zeros((4, 3))
    assert L.mgcfd_h5_read(f, b"xs", b"double", y.ctypes.data, 4, 3) == 0 and np.array_equal(x, y)
    j = np.zeros((4, 2), np.int32)
    assert L.mgcfd_h5_read(f, b"edge-->node", b"int", j.ctypes.data, 4, 2) == 0 and np.array_equal(i, j)
    assert L.mgcfd_h5_read(f, b"xs", b"double", y.ctypes.data, 5, 3) != 0
    assert b"shape" in L.mgcfd_h5_last_error()
    assert L.mgcfd_h5_shape(f, b"missing", C.byref(r), C.byref(c)) != 0
    assert L.mgcfd_h5_shape(str(tmp_path / "none.h5").encode(), b"xs", C.byref(r), C.byref(c)) != 0


def test_meshgen_deck(tmp_path):
    d = str(tmp_path / "deck")
    subprocess.run([os.path.join(BIN, "mgcfd_meshgen"), "--box", "8", "--levels", "3", "--seed", "4", "-o", d],
                   check=True, capture_output=True)
    text = open(os.path.join(d, "input.dat")).read()
    assert "num_levels=3" in text and "mesh_name=" in text and "[levels]" in text
    L = h5()
    mesh = mgcfd.mesh_box(8, 3, seed=4)
    for l in range(3):
        f = os.path.join(d, f"level{l}.h5").encode()
        for ds, key, dt in [("node_coordinates", "coords", np.float64), ("edge-->node", "edge_to_node", np.int32),
                            ("edge_weights", "edge_weights", np.float64)]:
            a = mesh[l][key]
            got = np.zeros(a.shape, dt)
            assert L.mgcfd_h5_read(f, ds.encode(), b"int" if dt == np.int32 else b"double", got.ctypes.data,
                                   a.shape[0], a.shape[1]) == 0
            assert np.array_equal(got, a), (l, ds)
