"""Pins the CPU oracle (oracle/mgcfd_oracle.c) -- CPU only.

1. Golden vectors: the reference's own kernel headers, compiled in place and
   driven in OP2 sequential order (tests/golden/make_golden.py), wrote the
   expected outputs; the oracle must reproduce every one bit for bit.
2. Where the compiled reference (oracle/_ref) is present, the oracle is also
   compared bit for bit on fresh seeded inputs, kernel by kernel and for a
   whole multigrid solve.
3. Properties that the mathematics fixes, each chosen so that a plausible
   mistake (dropped term, wrong sign or operand, swapped roles) fails it:
   freestream speed of sound, a<->b swap symmetry, zero weights, uniform-state
   dissipation, conservation, boundary branches, the closed forms of dt, the
   step factor and the RK coefficients, restriction as a mean, prolongation,
   the inverse-distance prolongation on a hand mesh and as a weighted mean,
   cell volumes of a regular lattice, and the -v tolerance rule.
"""
import os

import numpy as np
import pytest

from oracle import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KERNEL_FIXTURES = ["kernels_box6.npz", "kernels_knn400.npz"]
SOLVE_FIXTURES = ["solve_box8_g4.npz", "solve_knn400_g3.npz"]


def bits_equal(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.fixture(scope="module")
def orc():
    return Oracle("orc")


def consts_from(orc, arr):
    c = orc.freestream(0.5, 3.0)
    c.smoothing_coefficient = arr[0]
    for i in range(5):
        c.ff_variable[i] = arr[1 + i]
    for k, name in enumerate(["ff_flux_contribution_momentum_x", "ff_flux_contribution_momentum_y",
                              "ff_flux_contribution_momentum_z", "ff_flux_contribution_density_energy"]):
        for i in range(3):
            getattr(c, name)[i] = arr[6 + 3 * k + i]
    return c


# ------------------------------------------------------------ golden vectors

@pytest.mark.parametrize("fixture", KERNEL_FIXTURES)
def test_golden_kernels(orc, fixture):
    g = dict(np.load(os.path.join(GOLDEN, fixture)))
    c = consts_from(orc, g["consts"])
    fresh = orc.freestream(0.5, 3.0)
    assert bits_equal(np.array(fresh.ff_variable[:]), g["consts"][1:6])
    n = len(g["var"])
    ew = g["ew_in"].copy()
    vol = np.zeros(n)
    orc.calculate_cell_volumes(g["e2n"], g["coords"], ew, vol)
    assert bits_equal(vol, g["vol"]) and bits_equal(ew, g["ew_cv"])
    orc.dampen_ewt(ew)
    bw = g["bw_in"].copy()
    orc.dampen_ewt(bw)
    assert bits_equal(ew, g["ew"]) and bits_equal(bw, g["bw"])
    f = g["flux0"].copy()
    orc.compute_flux_edge(g["e2n"], g["var"], g["ew"], f, c)
    assert bits_equal(f, g["flux_edge"])
    orc.compute_bnd_node_flux(g["bnd_group"], g["b2n"], g["bw"], g["var"], f, c)
    assert bits_equal(f, g["flux_edge_bnd"])
    s = g["flux0"].copy()
    orc.unstructured_stream(g["e2n"], g["var"], g["ew"], s)
    assert bits_equal(s, g["stream"])
    dt = np.zeros(n)
    orc.calculate_dt(g["var"], g["vol"], dt)
    assert bits_equal(dt, g["dt"])
    assert bits_equal([orc.get_min_dt(dt)], g["min_dt"])
    sf = np.zeros(n)
    orc.compute_step_factor(g["var"], g["vol"], g["min_dt"][0], sf)
    assert bits_equal(sf, g["sf"])
    for rk in range(3):
        fl = g["flux_edge_bnd"].copy()
        v = g["var"].copy()
        orc.time_step(rk, g["sf"], fl, g["old"], v)
        assert bits_equal(v, g[f"ts_var_rk{rk}"]) and bits_equal(fl, g[f"ts_flux_rk{rk}"])
    res = np.zeros((n, 5))
    orc.residual(g["old"], g["var"], res)
    assert bits_equal(res, g["residual"])
    assert bits_equal([orc.calc_rms_sum(res)], g["rms_sum"])
    vc = g["var_c_in"].copy()
    scr = np.zeros(len(vc), np.int32)
    orc.up_pre(g["f2c"], vc, scr)
    orc.up(g["f2c"], g["var"], vc, scr)
    assert np.array_equal(scr, g["up_scratch"])
    orc.up_post(vc, scr)
    assert bits_equal(vc, g["var_c"])
    vf = g["var"].copy()
    orc.down(g["f2c"], vf, g["residual"], g["coords"], g["res_c"], g["coords_c"])
    assert bits_equal(vf, g["down_var"])
    rp, ws = np.zeros((n, 5)), np.zeros(n)
    orc.down_v2(g["e2n"], np.ascontiguousarray(g["f2c"][g["e2n"]]), g["coords"], g["coords_c"], g["res_c"], rp, ws)
    assert bits_equal(rp, g["dv2_rp"]) and bits_equal(ws, g["dv2_wsum"])
    vf = g["var"].copy()
    orc.down_v2_post(rp, ws, g["residual"], vf)
    assert bits_equal(vf, g["dv2_var"])
    diff = np.zeros((n, 5))
    orc.identify_differences(g["v_test"], g["var"], diff)
    assert bits_equal(diff, g["v_diff"])
    assert orc.count_non_zeros(diff) == g["v_count"][0] > 0


def mesh_from_fixture(g):
    mesh, l = [], 0
    while f"L{l}_coords" in g:
        mesh.append(dict(coords=g[f"L{l}_coords"], edge_to_node=g[f"L{l}_e2n"], edge_weights=g[f"L{l}_ew"],
                         bnd_to_node=g[f"L{l}_b2n"], bnd_group=g[f"L{l}_bg"], bnd_weights=g[f"L{l}_bw"],
                         node_to_mg_node=g.get(f"L{l}_f2c")))
        l += 1
    return mesh


@pytest.mark.parametrize("fixture", SOLVE_FIXTURES)
def test_golden_solve(orc, fixture):
    g = dict(np.load(os.path.join(GOLDEN, fixture)))
    mesh = mesh_from_fixture(g)
    r = orc.solve(mesh, int(g["cycles"][0]), 0.5, 3.0)
    assert r["status"] == 0
    assert bits_equal(r["min_dt"], g["min_dt"])
    assert bits_equal(r["rms"], g["rms"])
    for l in range(len(mesh)):
        assert bits_equal(r["levels"][l]["variables"], g[f"L{l}_variables"]), l
        assert bits_equal(r["levels"][l]["volumes"], g[f"L{l}_volumes"]), l


# ------------------------------------------- compiled reference, fresh inputs

needs_ref = pytest.mark.skipif(not Oracle.available("ref"), reason="oracle/_ref not built (no reference checkout)")


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_equals_reference_kernels(orc, knn_small, seed):
    ref = Oracle("ref")
    from conftest import perturbed_state
    m = knn_small[0]
    c, cr = orc.freestream(0.7, -2.0 + seed), ref.freestream(0.7, -2.0 + seed)
    assert bits_equal(c.ff_variable[:], cr.ff_variable[:])
    assert bits_equal(c.ff_flux_contribution_density_energy[:], cr.ff_flux_contribution_density_energy[:])
    e2n = np.ascontiguousarray(m["edge_to_node"], np.int32)
    n = len(m["coords"])
    var = perturbed_state(n, np.array(c.ff_variable[:]), seed, amp=0.05)
    w = np.array(m["edge_weights"])
    f1, f2 = np.zeros((n, 5)), np.zeros((n, 5))
    orc.compute_flux_edge(e2n, var, w, f1, c)
    ref.compute_flux_edge(e2n, var, w, f2, cr)
    assert bits_equal(f1, f2)
    grp = np.random.default_rng(seed).integers(-1, 10, size=len(m["bnd_to_node"])).astype(np.int32)
    b2n = np.ascontiguousarray(m["bnd_to_node"], np.int32)
    orc.compute_bnd_node_flux(grp, b2n, m["bnd_weights"], var, f1, c)
    ref.compute_bnd_node_flux(grp, b2n, m["bnd_weights"], var, f2, cr)
    assert bits_equal(f1, f2)


@needs_ref
def test_oracle_equals_reference_solve(orc, knn_small):
    ref = Oracle("ref")
    a = orc.solve(knn_small, 3, 0.6, 1.0)
    b = ref.solve(knn_small, 3, 0.6, 1.0)
    assert a["status"] == b["status"] == 0
    assert bits_equal(a["min_dt"], b["min_dt"]) and bits_equal(a["rms"], b["rms"])
    for la, lb in zip(a["levels"], b["levels"]):
        assert bits_equal(la["variables"], lb["variables"])


@needs_ref
@pytest.mark.parametrize("level", [0, 1])
def test_oracle_equals_reference_down_v2(orc, knn_small, level):
    """Inverse-distance prolongation loops (mg.h:90-204) in OP2 sequential
    order, on a kNN level whose fine nodes partly coincide with their coarse
    parents (both branches) and whose edges partly share one parent."""
    ref = Oracle("ref")
    f, c = knn_small[level], knn_small[level + 1]
    e2n = np.ascontiguousarray(f["edge_to_node"], np.int32)
    n2c = np.asarray(f["node_to_mg_node"], np.int32)
    e2mg = np.ascontiguousarray(n2c[e2n], np.int32)
    xf, xc = np.asarray(f["coords"], np.float64), np.asarray(c["coords"], np.float64)
    coinc = np.all(xf == xc[n2c], axis=1)
    assert 0 < coinc.sum() < len(xf)
    rng = np.random.default_rng(11 + level)
    rc = rng.uniform(-1e-3, 1e-3, (len(xc), 5))
    rf = rng.uniform(-1e-3, 1e-3, (len(xf), 5))
    outs = []
    for o in (orc, ref):
        rp, ws = np.full((len(xf), 5), 3.0), np.full(len(xf), 3.0)
        o.down_v2_pre(ws, rp)
        assert np.all(rp == 0.0) and np.all(ws == 0.0)
        o.down_v2(e2n, e2mg, xf, xc, rc, rp, ws)
        var = np.ones((len(xf), 5))
        o.down_v2_post(rp, ws, rf, var)
        outs.append((rp, ws, var))
    for x, y in zip(*outs):
        assert bits_equal(x, y)
    rp, ws, _ = outs[0]
    assert np.all(ws[coinc] == 1.0) and bits_equal(rp[coinc], rc[n2c[coinc]])


@needs_ref
def test_oracle_equals_reference_solve_down_v2(orc, knn_small):
    ref = Oracle("ref")
    a = orc.solve(knn_small, 3, 0.6, 1.0, down_v2=True)
    b = ref.solve(knn_small, 3, 0.6, 1.0, down_v2=True)
    assert a["status"] == b["status"] == 0
    assert bits_equal(a["min_dt"], b["min_dt"]) and bits_equal(a["rms"], b["rms"])
    for la, lb in zip(a["levels"], b["levels"]):
        assert bits_equal(la["variables"], lb["variables"])
    plain = orc.solve(knn_small, 3, 0.6, 1.0)
    assert not bits_equal(plain["levels"][0]["variables"], a["levels"][0]["variables"])


# ------------------------------------------------------------ closed forms

def thermo(U):
    rho, m, E = U[0], U[1:4], U[4]
    u = m / rho
    p = 0.4 * (E - 0.5 * rho * (u @ u))
    return u, p, np.sqrt(1.4 * p / rho)


def flux_tensor(U):
    u, p, _ = thermo(U)
    m, E = U[1:4], U[4]
    F = np.outer(u, m) + p * np.eye(3)
    return F, u * (E + p)


@pytest.mark.parametrize("mach,aoa", [(0.5, 3.0), (0.7, 2.0), (1.2, 0.0), (0.3, -2.0)])
def test_freestream(orc, mach, aoa):
    c = orc.freestream(mach, aoa)
    U = np.array(c.ff_variable[:])
    u, p, cs = thermo(U)
    assert U[0] == 1.4
    assert abs(cs - 1.0) < 1e-15 and abs(p - 1.0) < 1e-15
    a = np.deg2rad(aoa)
    assert np.allclose(u, [mach * np.cos(a), mach * np.sin(a), 0.0], rtol=0, atol=1e-15)
    F, fe = flux_tensor(U)
    assert np.allclose(np.array(c.ff_flux_contribution_momentum_x[:]), F[0], rtol=1e-14, atol=1e-15)
    assert np.allclose(np.array(c.ff_flux_contribution_momentum_y[:]), F[1], rtol=1e-14, atol=1e-15)
    assert np.allclose(np.array(c.ff_flux_contribution_density_energy[:]), fe, rtol=1e-14, atol=1e-15)
    assert c.smoothing_coefficient == float(np.float32(0.2))


def random_states(n, seed, amp=0.1):
    rng = np.random.default_rng(seed)
    U = np.zeros((n, 5))
    U[:, 0] = 1.4 * (1 + amp * rng.uniform(-1, 1, n))
    U[:, 1:4] = rng.uniform(-0.5, 0.5, (n, 3))
    u2 = (U[:, 1:4] ** 2).sum(1) / U[:, 0]
    U[:, 4] = 2.5 * (1 + amp * rng.uniform(-1, 1, n)) + 0.5 * u2
    return U


def test_flux_swap_symmetry(orc, knn_small):
    """Swapping the endpoints and negating w reproduces the reference's b-side
    formula (factor_b's operand order) exactly: bitwise equal fluxes."""
    m = knn_small[0]
    c = orc.freestream(0.5, 3.0)
    e2n = np.ascontiguousarray(m["edge_to_node"], np.int32)
    n = len(m["coords"])
    U = random_states(n, 5)
    w = np.array(m["edge_weights"])
    f1, f2 = np.zeros((n, 5)), np.zeros((n, 5))
    orc.compute_flux_edge(e2n, U, w, f1, c)
    orc.compute_flux_edge(np.ascontiguousarray(e2n[:, ::-1]), U, -w, f2, c)
    assert bits_equal(f1, f2)


def test_flux_zero_weights_and_conservation(orc, knn_small):
    m = knn_small[0]
    c = orc.freestream(0.5, 3.0)
    e2n = np.ascontiguousarray(m["edge_to_node"], np.int32)
    n = len(m["coords"])
    U = random_states(n, 6)
    f0 = np.random.default_rng(1).uniform(-1, 1, (n, 5))
    f = f0.copy()
    orc.compute_flux_edge(e2n, U, np.zeros((len(e2n), 3)), f, c)
    assert np.array_equal(f, f0)
    f = np.zeros((n, 5))
    orc.compute_flux_edge(e2n, U, np.array(m["edge_weights"]), f, c)
    # each edge adds equal and opposite amounts (factor_a == factor_b up to rounding)
    assert np.all(np.abs(f.sum(0)) <= 1e-13 * np.abs(f).sum(0))
    assert np.abs(f).max() > 0


def test_flux_uniform_state_single_edge(orc):
    """U_a == U_b: the dissipation term vanishes and flux_a = -w . F(U), flux_b = +w . F(U)."""
    c = orc.freestream(0.5, 3.0)
    U = random_states(1, 7)[0]
    var = np.stack([U, U])
    w = np.array([[0.3, -0.2, 0.7]])
    f = np.zeros((2, 5))
    orc.compute_flux_edge(np.array([[0, 1]], np.int32), var, w, f, c)
    F, fe = flux_tensor(U)
    m = U[1:4]
    expect_a = -np.concatenate([[w[0] @ m], F.T @ w[0], [w[0] @ fe]])
    assert np.allclose(f[0], expect_a, rtol=1e-14, atol=1e-16)
    assert np.allclose(f[1], -expect_a, rtol=1e-14, atol=1e-16)


def test_flux_dissipation_term(orc):
    """With w along x and states differing only in density, the density
    increment is factor*(rho_a - rho_b) - w0/2 (m_a + m_b)_x with
    factor = -|w| s/2 (|u_a| + |u_b| + c_a + c_b)."""
    c = orc.freestream(0.5, 3.0)
    Ua = np.array([1.4, 0.2, 0.1, 0.0, 2.6])
    Ub = np.array([1.3, 0.2, 0.1, 0.0, 2.6])
    w = np.array([[0.5, 0.0, 0.0]])
    f = np.zeros((2, 5))
    orc.compute_flux_edge(np.array([[0, 1]], np.int32), np.stack([Ua, Ub]), w, f, c)
    ua, _, ca = thermo(Ua)
    ub, _, cb = thermo(Ub)
    factor = -0.5 * c.smoothing_coefficient * 0.5 * (np.linalg.norm(ua) + np.linalg.norm(ub) + ca + cb)
    expect = factor * (Ua[0] - Ub[0]) - 0.25 * (Ua[1] + Ub[1])
    assert abs(f[0, 0] - expect) <= 1e-15 * abs(expect) + 1e-17
    assert abs(f[1, 0] + expect) <= 1e-14 * abs(expect)


def test_bnd_flux_branches(orc):
    c = orc.freestream(0.5, 3.0)
    ffv = np.array(c.ff_variable[:])
    U = random_states(4, 8)
    U[3] = ffv
    w = np.array([[0.2, -0.4, 0.1]] * 4)
    grp = np.array([-1, 2, 9, 5], np.int32)
    f = np.zeros((4, 5))
    orc.compute_bnd_node_flux(grp, np.arange(4, dtype=np.int32), w, U, f, c)
    for i in (0, 1):                       # pressure wall: momentum only, w * p
        _, p, _ = thermo(U[i])
        assert f[i, 0] == 0 and f[i, 4] == 0
        assert np.allclose(f[i, 1:4], w[i] * p, rtol=1e-15, atol=0)
    assert np.all(f[2] == 0)               # group > 7: untouched
    F, fe = flux_tensor(ffv)               # far field at U = U_inf: w . F_inf
    expect = np.concatenate([[w[3] @ ffv[1:4]], F.T @ w[3], [w[3] @ fe]])
    assert np.allclose(f[3], expect, rtol=1e-13, atol=1e-15)


def test_dt_step_factor_and_rk(orc):
    c = orc.freestream(0.5, 3.0)
    n = 16
    U = np.tile(np.array(c.ff_variable[:]), (n, 1))
    vol = np.random.default_rng(3).uniform(0.5, 2.0, n)
    dt = np.zeros(n)
    orc.calculate_dt(U, vol, dt)
    assert np.allclose(dt, 0.5 * np.cbrt(vol) / (0.5 + 1.0), rtol=1e-14, atol=0)
    mdt = orc.get_min_dt(dt)
    assert mdt == dt.min()
    sf = np.zeros(n)
    orc.compute_step_factor(U, vol, mdt, sf)
    assert bits_equal(sf, mdt / vol)
    for rk, coef in enumerate([1 / 4, 1 / 3, 1 / 2]):
        fl = np.ones((n, 5))
        v = np.zeros((n, 5))
        orc.time_step(rk, np.ones(n), fl, np.zeros((n, 5)), v)
        assert np.all(v == coef) and np.all(fl == 0)


def test_restrict_is_mean_and_keeps_childless(orc):
    f2c = np.array([0, 0, 0, 2, 2], np.int32)
    vf = np.arange(25, dtype=np.float64).reshape(5, 5)
    vc = np.full((3, 5), 7.5)
    scr = np.zeros(3, np.int32)
    orc.up_pre(f2c, vc, scr)
    orc.up(f2c, vf, vc, scr)
    orc.up_post(vc, scr)
    assert np.allclose(vc[0], vf[:3].mean(0), rtol=1e-15)
    assert np.allclose(vc[2], vf[3:].mean(0), rtol=1e-15)
    assert np.all(vc[1] == 7.5)            # no children: keeps its stale value


def test_prolong_formula(orc):
    f2c = np.array([0, 1], np.int32)
    xf = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 2.0]])
    xc = np.array([[3.0, 4.0, 0.0], [1.0, 2.0, 2.0]])
    rf = np.random.default_rng(4).uniform(-1, 1, (2, 5))
    rc = np.random.default_rng(5).uniform(-1, 1, (2, 5))
    vf = np.ones((2, 5))
    orc.down(f2c, vf, rf, xf, rc, xc)
    d = np.abs(xf[0] - xc[0])
    dm = np.linalg.norm(d)
    dr = rc[0] - rf[0]
    expect = np.array([1 - dm * dr[0], 1 - d[0] * dr[1], 1 - d[1] * dr[2], 1 - d[2] * dr[3], 1 - dm * dr[4]])
    assert np.allclose(vf[0], expect, rtol=1e-15, atol=1e-15)
    assert np.all(vf[1] == 1.0)            # coincident nodes: no correction


def test_down_v2_hand_mesh(orc):
    """Three fine nodes, two coarse nodes, every fine-to-coarse distance 5
    (weight 1/5).  n1 coincides with its parent C1, so its sums are set (not
    added) on each of its two edges.  n0 is side b of e1, whose b-side terms
    both take b's parent residual r0 (mg.h:178-189), so n0 gets 3 r0 + r1
    rather than 2 r0 + 2 r1."""
    xc = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, 5.0]])             # C0, C1
    xf = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 5.0], [3.0, 4.0, 5.0]])
    n2c = np.array([0, 1, 1], np.int32)
    e2n = np.array([[0, 1], [2, 0], [1, 2]], np.int32)
    e2mg = np.ascontiguousarray(n2c[e2n])
    rc = np.array([[1.0, 2.0, -3.0, 0.5, 4.0], [-2.0, 1.0, 1.0, 8.0, -1.0]])
    r0, r1 = rc
    rp, ws = np.zeros((3, 5)), np.zeros(3)
    orc.down_v2(e2n, e2mg, xf, xc, rc, rp, ws)
    assert np.allclose(rp[0], 0.6 * r0 + 0.2 * r1, rtol=1e-15, atol=1e-16) and np.isclose(ws[0], 0.8, rtol=1e-15)
    assert np.all(rp[1] == r1) and ws[1] == 1.0
    assert np.allclose(rp[2], 0.2 * r0 + 0.6 * r1, rtol=1e-15, atol=1e-16) and np.isclose(ws[2], 0.8, rtol=1e-15)
    rf = np.array([[0.5] * 5, [0.25] * 5, [-1.0] * 5])
    var = np.full((3, 5), 2.0)
    orc.down_v2_post(rp, ws, rf, var)
    assert np.allclose(var[0], 2.0 + 0.5 - (0.75 * r0 + 0.25 * r1), rtol=1e-15)
    assert np.all(var[1] == 2.0 + 0.25 - r1)
    assert np.allclose(var[2], 2.0 - 1.0 - (0.25 * r0 + 0.75 * r1), rtol=1e-15)


def test_down_v2_uniform_coarse_residual(orc, knn_small):
    """A weighted mean of equal values: with one residual on every coarse
    node the prolonged residual rp / wsum is that residual at every fine node
    with an edge."""
    f, c = knn_small[0], knn_small[1]
    e2n = np.ascontiguousarray(f["edge_to_node"], np.int32)
    n2c = np.asarray(f["node_to_mg_node"], np.int32)
    xf, xc = np.asarray(f["coords"], np.float64), np.asarray(c["coords"], np.float64)
    r = np.array([1e-3, -2e-3, 5e-4, 0.0, 3e-3])
    rc = np.tile(r, (len(xc), 1))
    rp, ws = np.zeros((len(xf), 5)), np.zeros(len(xf))
    orc.down_v2(e2n, np.ascontiguousarray(n2c[e2n]), xf, xc, rc, rp, ws)
    assert np.all(ws > 0)
    assert np.allclose(rp / ws[:, None], r[None, :], rtol=1e-13, atol=1e-18)


def test_cell_volumes_regular_lattice(orc):
    h = 0.5
    k = 3
    idx = np.arange(k ** 3).reshape(k, k, k)
    coords = np.stack(np.meshgrid(*[np.arange(k) * h] * 3, indexing="ij"), -1).reshape(-1, 3)
    edges, w = [], []
    for ax in range(3):
        a = np.take(idx, range(k - 1), axis=ax).ravel()
        b = np.take(idx, range(1, k), axis=ax).ravel()
        edges.append(np.stack([a, b], 1))
        ww = np.zeros((len(a), 3))
        ww[:, ax] = h * h
        w.append(ww)
    e2n = np.ascontiguousarray(np.concatenate(edges), np.int32)
    w = np.concatenate(w)
    vol = np.zeros(k ** 3)
    orc.calculate_cell_volumes(e2n, coords, w, vol)
    deg = np.bincount(e2n.ravel(), minlength=k ** 3)
    assert np.allclose(vol, deg * h ** 3 / 6, rtol=1e-15)
    assert np.isclose(vol[idx[1, 1, 1]], h ** 3)
    # rewritten weights are parallel to the edge, magnitude |w| / d
    dx = coords[e2n[:, 1]] - coords[e2n[:, 0]]
    assert np.allclose(w, dx * (h * h) / h / h, rtol=1e-15)


def test_v_tolerance_rule(orc):
    master = np.array([[1.0, -2.0, 1e-30, 0.0, 5.0]])
    test = master.copy()
    test[0, 0] *= 1 + 2e-7       # beyond 1e-7 relative: counted
    test[0, 1] *= 1 + 5e-8       # within: not counted
    test[0, 2] = 2e-19           # |t - r| < 3e-19 absolute floor: not counted
    test[0, 3] = 4e-19           # beyond the floor: counted
    diff = np.zeros((1, 5))
    orc.identify_differences(test, master, diff)
    assert orc.count_non_zeros(diff) == 2
    assert diff[0, 0] != 0 and diff[0, 3] != 0 and diff[0, 1] == 0 and diff[0, 2] == 0
