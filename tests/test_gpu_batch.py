"""Batched solves: independent meshes of uneven size concatenated behind offset
tables and advanced together (mgcfd_batch_*), against the oracle's solve of
each mesh alone.

Bars: GATHER_EXACT and the part variants: every mesh's state within 1e-12
relative per variable of the oracle (the dt cbrt ulps are the only source of
difference for the exact variant); mixed precision: within 1e-6.  A mesh's
result must not depend on the other meshes in its batch: bitwise for the
deterministic variants.
"""
import numpy as np
import pytest

import mgcfd
from oracle import Oracle

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-12
MIXED_STATE_TOL = 1e-6


def rel(g, o):
    return float((np.abs(g - o).max(axis=0) / np.abs(o).max(axis=0)).max())


@pytest.fixture(scope="module")
def meshes():
    return [mgcfd.mesh_box(8, 3, seed=7), mgcfd.mesh_knn(3000, 9300, 3, seed=3),
            mgcfd.mesh_knn(1200, 3720, 3, seed=4)]


MACH = [0.5, 0.7, 0.35]
AOA = [3.0, -1.5, 5.0]


@pytest.fixture(scope="module")
def oracle_solves(meshes):
    o = Oracle("orc")
    return [o.solve(m, 3, MACH[i], AOA[i]) for i, m in enumerate(meshes)]


@pytest.fixture
def rt():
    r = mgcfd.Runtime()
    yield r
    r.close()


@pytest.mark.parametrize("variant,tol", [("gather_exact", STATE_TOL), ("part_coloured", STATE_TOL),
                                         ("mixed_coloured", MIXED_STATE_TOL)])
def test_batch_matches_each_solve_alone(rt, meshes, oracle_solves, variant, tol):
    b = mgcfd.Batch(rt, meshes, MACH, AOA, variant)
    st, rms = b.run(3)
    assert (st == 0).all()
    for i, m in enumerate(meshes):
        ref = oracle_solves[i]
        for l in range(len(m)):
            err = rel(b.fetch(i, l), ref["levels"][l]["variables"])
            print(f"{variant} mesh {i} L{l}: {err:.3e}")
            assert err <= tol, (i, l, err)
        rms_tol = 1e-10 if tol == STATE_TOL else 1e-3
        assert np.allclose(rms[i], ref["rms"], rtol=rms_tol, atol=0)
    b.close()


@pytest.mark.parametrize("variant", ["gather_exact", "part_coloured", "mixed_coloured"])
def test_batch_result_independent_of_mix(rt, meshes, variant):
    """Mesh 1 alone (a batch of one) and inside the three-mesh batch: bitwise equal."""
    alone = mgcfd.Batch(rt, [meshes[1]], [MACH[1]], [AOA[1]], variant)
    alone.run(3)
    a = [alone.fetch(0, l) for l in range(3)]
    alone.close()
    mix = mgcfd.Batch(rt, meshes, MACH, AOA, variant)
    mix.run(3)
    for l in range(3):
        g = mix.fetch(1, l)
        assert np.array_equal(g.view(np.uint64), a[l].view(np.uint64)), l
    mix.close()


@pytest.mark.parametrize("variant,tol", [("part_coloured", STATE_TOL), ("mixed_atomic", MIXED_STATE_TOL)])
def test_batch_autotune_keeps_results(rt, meshes, oracle_solves, variant, tol):
    """Tuning runs trial cycles on the batch and restores its state: the run
    after it matches the oracle from the setup state."""
    b = mgcfd.Batch(rt, meshes, MACH, AOA, variant)
    p = b.autotune(2)
    print(f"{variant}: autotuned {p}")
    assert all(x in (256, 192, 128) for x in p["part_nodes"])
    assert all(x in (256, 512) for x in p["part_threads"])
    assert all(x in (0, 2, 4) for x in p["persistent_per_cu"])
    st, _ = b.run(3)
    assert (st == 0).all()
    for i in range(len(meshes)):
        for l in range(3):
            assert rel(b.fetch(i, l), oracle_solves[i]["levels"][l]["variables"]) <= tol, (i, l)
    b.close()


@pytest.mark.parametrize("nodes,threads,persist", [(256, 512, 0), (192, 256, 0), (128, 512, 0), (256, 256, 4)])
@pytest.mark.parametrize("variant,tol", [("part_atomic", STATE_TOL), ("mixed_coloured", MIXED_STATE_TOL)])
def test_batch_part_configs_match_oracle(rt, meshes, oracle_solves, nodes, threads, persist, variant, tol):
    """Every part configuration the autotuner may choose, on every level."""
    b = mgcfd.Batch(rt, meshes, MACH, AOA, variant)
    for l in range(3):
        b.set_part_config(l, nodes, threads, persist)
    st, _ = b.run(3)
    assert (st == 0).all()
    for i in range(len(meshes)):
        for l in range(3):
            assert rel(b.fetch(i, l), oracle_solves[i]["levels"][l]["variables"]) <= tol, (i, l)
    b.close()


def test_batch_graph_and_eager_agree(rt, meshes):
    b1 = mgcfd.Batch(rt, meshes, MACH, AOA, "gather_exact")
    b1.run(2, use_graph=False)
    b2 = mgcfd.Batch(rt, meshes, MACH, AOA, "gather_exact")
    b2.run(2, use_graph=True)
    for i in range(len(meshes)):
        assert np.array_equal(b1.fetch(i, 0), b2.fetch(i, 0))
    b1.close()
    b2.close()


@pytest.mark.parametrize("variant,tol", [("part_coloured", STATE_TOL), ("mixed_atomic", MIXED_STATE_TOL)])
def test_batch_persistent_workgroups_walk_many_parts(rt, variant, tol):
    """Persistent workgroups (one per CU) with 64-node parts on a 40K-node
    mesh: every workgroup walks several parts of its XCD's slice."""
    ms = [mgcfd.mesh_knn(40000, 124000, 3, seed=21), mgcfd.mesh_knn(3000, 9300, 3, seed=22)]
    mach, aoa = [0.5, 0.6], [3.0, 1.0]
    o = Oracle("orc")
    refs = [o.solve(m, 3, mach[i], aoa[i]) for i, m in enumerate(ms)]
    b = mgcfd.Batch(rt, ms, mach, aoa, variant)
    for l in range(3):
        b.set_part_config(l, 64, 256, 1)
    st, _ = b.run(3)
    assert (st == 0).all()
    for i in range(2):
        for l in range(3):
            assert rel(b.fetch(i, l), refs[i]["levels"][l]["variables"]) <= tol, (i, l)
    b.close()
