"""Degenerate and small inputs, against the oracle.

The benchmark meshes are large and regular; these cases are not:
- a mesh smaller than one part (one workgroup, fewer owned nodes than the
  kernels' 256 and an odd count), with a duplicated edge, an edge given in
  both directions, and nodes that no edge touches: every flux variant and
  the boundary flux;
- a mesh with an empty boundary set, and global reductions (OP_MIN, OP_INC)
  over an empty set, which leave the caller's value as it was (OP2 combines
  the caller's value with none);
- a two-level hand mesh solved loop by loop, device-resident and
  graph-replayed, including a coarse node that no fine node restricts to
  (it keeps its own state, mg.h:29-88), against the oracle.
"""
import numpy as np
import pytest

import mgcfd
from conftest import perturbed_state
from oracle import Oracle
from test_gpu_parity import (FLUX_FAST_TOL, MIXED_FLUX_TOL, STATE_TOL, bitwise, oracle_setup, rel_err)

pytestmark = pytest.mark.gpu

VARIANTS = ["gather_exact", "gather_fast", "atomic", "nodal_exact", "part_coloured", "part_atomic",
            "mixed_coloured", "mixed_atomic"]


def hand_level(n, edges, bnd, groups, seed):
    rng = np.random.default_rng(seed)
    e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
    return dict(coords=rng.uniform(0.0, 2.0, size=(n, 3)),
                edge_to_node=e,
                edge_weights=rng.uniform(-0.6, 0.6, size=(len(e), 3)),
                bnd_to_node=np.asarray(bnd, dtype=np.int32).reshape(-1),
                bnd_group=np.asarray(groups, dtype=np.int32).reshape(-1),
                bnd_weights=rng.uniform(-0.4, 0.4, size=(len(bnd), 3)))


# 9 nodes: (0, 1) twice and (1, 0) once; nodes 6 and 8 have no edges
TINY_EDGES = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 2), (1, 0), (0, 1), (3, 5), (5, 7), (7, 3)]


@pytest.fixture
def rt():
    r = mgcfd.Runtime()
    r.set_flux_variant(mgcfd.FLUX_VARIANTS["gather_exact"])
    yield r
    r.close()


@pytest.fixture(scope="module")
def orc():
    return Oracle("orc")


def flux_case(rt, orc, mesh, variant):
    c = orc.freestream(0.5, 3.0)
    s = mgcfd.Solver(rt, mesh, 0.5, 3.0)
    s.setup()
    ref = oracle_setup(orc, mesh, c)
    m0 = mesh[0]
    n = len(m0["coords"])
    var = perturbed_state(n, c.ff_variable[:], seed=31)
    flux0 = perturbed_state(n, np.ones(5), seed=32, amp=1e-6) * 1e-9
    rt.upload(s.lv[0].variables, var)
    rt.upload(s.lv[0].fluxes, flux0)
    rt.set_flux_variant(mgcfd.FLUX_VARIANTS[variant])
    s.flux_edge(0)
    g = s.state(0, "fluxes")
    o = flux0.copy()
    orc.compute_flux_edge(np.ascontiguousarray(m0["edge_to_node"], np.int32), var, ref[0]["ew"], o, c)
    if variant in ("gather_exact", "nodal_exact"):
        assert bitwise(g, o), f"max rel {rel_err(g, o):.3e}"
    elif variant.startswith("mixed"):
        assert rel_err(g - flux0, o - flux0) <= MIXED_FLUX_TOL
    else:
        assert rel_err(g, o) <= FLUX_FAST_TOL
    return s, var, ref, c, g, flux0


@pytest.mark.parametrize("variant", VARIANTS)
def test_flux_on_a_mesh_smaller_than_one_part(rt, orc, variant):
    # boundary groups: wall (0), far field (3, 5), none (9) (flux.h:29-37)
    mesh = [hand_level(9, TINY_EDGES, [0, 3, 6, 8], [0, 3, 5, 9], seed=5)]
    s, var, ref, c, g, flux0 = flux_case(rt, orc, mesh, variant)
    # the nodes no edge touches keep their flux
    assert bitwise(g[[6, 8]], flux0[[6, 8]])
    # the boundary flux on top (bnd nodes include both edgeless nodes)
    s.flux_bnd(0)
    o = g.copy()
    m0 = mesh[0]
    orc.compute_bnd_node_flux(np.ascontiguousarray(m0["bnd_group"], np.int32),
                              np.ascontiguousarray(m0["bnd_to_node"], np.int32), ref[0]["bw"], var, o, c)
    assert bitwise(s.state(0, "fluxes"), o)


@pytest.mark.parametrize("variant", ["gather_exact", "part_atomic", "mixed_coloured"])
def test_mesh_with_no_boundary_nodes(rt, orc, variant):
    mesh = [hand_level(9, TINY_EDGES, [], [], seed=6)]
    s, *_ = flux_case(rt, orc, mesh, variant)
    before = s.state(0, "fluxes")
    s.flux_bnd(0)
    assert bitwise(s.state(0, "fluxes"), before)


def test_reductions_over_an_empty_set(rt):
    st = rt.decl_set(0, "empty_set")
    d1 = rt.decl_dat(st, 1, np.zeros(0), "empty_d1")
    d5 = rt.decl_dat(st, 5, np.zeros((0, 5)), "empty_d5")
    mn = np.array([123.25])
    rt.par_loop("get_min_dt_kernel", st, rt.arg(d1, -1, None, 1, acc=mgcfd.OP_READ), rt.gbl(mn, mgcfd.OP_MIN))
    assert mn[0] == 123.25
    r = np.array([7.5])
    rt.par_loop("calc_rms_kernel", st, rt.arg(d5, -1, None, 5, acc=mgcfd.OP_READ), rt.gbl(r, mgcfd.OP_INC))
    assert r[0] == 7.5
    cnt = np.array([4], dtype=np.int32)
    rt.par_loop("count_bad_vals", st, rt.arg(d5, -1, None, 5, acc=mgcfd.OP_READ), rt.gbl(cnt, mgcfd.OP_INC))
    assert cnt[0] == 4
    assert rt.fetch(d5).shape == (0, 5)


def two_level_mesh(childless):
    """10 fine nodes on 4 coarse ones; with 'childless', coarse node 3 has no
    fine node restricting to it."""
    fine = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2), (4, 5), (5, 6), (6, 4), (3, 4), (6, 7), (7, 8), (8, 9),
            (9, 7), (2, 5), (1, 8), (0, 9)]
    coarse = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)]
    l0 = hand_level(10, fine, [0, 4, 9], [0, 2, 3], seed=11)
    l1 = hand_level(4, coarse, [0, 3], [2, 0], seed=12)
    l0["node_to_mg_node"] = np.array([0, 0, 1, 1, 2, 2, 2, 0, 1, 2] if childless else
                                     [0, 0, 1, 1, 2, 2, 3, 3, 1, 2], dtype=np.int32)
    return [l0, l1]


@pytest.mark.parametrize("childless", [False, True])
@pytest.mark.parametrize("variant,mode", [("gather_exact", "loops"), ("gather_exact", "device"),
                                          ("gather_exact", "graph"), ("part_atomic", "graph"),
                                          ("part_coloured", "device"), ("mixed_atomic", "graph")])
def test_two_level_hand_mesh_solve(rt, orc, childless, variant, mode):
    mesh = two_level_mesh(childless)
    cycles = 3
    ref = orc.solve(mesh, cycles, 0.5, 3.0)
    s = mgcfd.Solver(rt, mesh, 0.5, 3.0)
    rt.set_flux_variant(mgcfd.FLUX_VARIANTS[variant])
    s.setup()
    if mode == "loops":
        st, rms, _ = s.run_loops(cycles)
    else:
        st, rms = s.run_device(cycles, use_graph=(mode == "graph"))
    assert st == ref["status"] == 0
    tol = 1e-6 if variant.startswith("mixed") else STATE_TOL
    for l in range(2):
        err = rel_err(s.state(l), ref["levels"][l]["variables"])
        print(f"{variant}/{mode} childless={childless} L{l}: {err:.3e}")
        assert err <= tol, f"level {l}"
    if not variant.startswith("mixed"):
        assert np.allclose(np.asarray(rms), ref["rms"], rtol=1e-10, atol=0)
