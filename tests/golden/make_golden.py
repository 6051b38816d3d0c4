#!/usr/bin/env python3
"""Writes the golden vectors of tests/test_oracle.py.

The expected outputs come from the reference's own kernel headers compiled in
place (oracle/_ref/libmgcfd_ref.so, built by `make -C oracle` from the
reference checkout), driven with OP2 sequential loop order.  Inputs are small
seeded meshes from this project's generator plus seeded states; every input
and output array is stored, so the fixtures stand alone.

    python tests/golden/make_golden.py      # rewrites tests/golden/*.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "mg-cfd-app-op2_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import mgcfd  # noqa: E402
from oracle import Oracle  # noqa: E402
from conftest import perturbed_state  # noqa: E402

MACH, AOA = 0.5, 3.0


def kernel_vectors(ref: Oracle, mesh: list[dict], seed: int) -> dict:
    """One call of every per-element loop on level 0 (and 0->1 transfer)."""
    m0, m1 = mesh[0], mesh[1]
    c = ref.freestream(MACH, AOA)
    ffv = np.array(c.ff_variable[:])
    n, n1 = len(m0["coords"]), len(m1["coords"])
    e2n = np.ascontiguousarray(m0["edge_to_node"], np.int32)
    b2n = np.ascontiguousarray(m0["bnd_to_node"], np.int32)
    rng = np.random.default_rng(seed)
    # every boundary branch: <=2 pressure wall (incl. negative), 3..7 far field, >7 none
    grp = rng.choice(np.array([-1, 0, 1, 2, 3, 5, 7, 8, 11], np.int32), size=len(b2n)).astype(np.int32)
    var = perturbed_state(n, ffv, seed)
    out = dict(e2n=e2n, b2n=b2n, bnd_group=grp, coords=np.ascontiguousarray(m0["coords"]),
               ew_in=np.array(m0["edge_weights"]), bw_in=np.array(m0["bnd_weights"]), var=var,
               consts=np.concatenate([[c.smoothing_coefficient], ffv, c.ff_flux_contribution_momentum_x[:],
                                      c.ff_flux_contribution_momentum_y[:], c.ff_flux_contribution_momentum_z[:],
                                      c.ff_flux_contribution_density_energy[:]]))
    # setup: cell volumes rewrite w and accumulate volumes; dampen both weight sets
    ew = out["ew_in"].copy()
    vol = np.zeros(n)
    ref.calculate_cell_volumes(e2n, out["coords"], ew, vol)
    out["vol"], out["ew_cv"] = vol.copy(), ew.copy()
    ref.dampen_ewt(ew)
    bw = out["bw_in"].copy()
    ref.dampen_ewt(bw)
    out["ew"], out["bw"] = ew, bw
    # flux loops, starting from a non-zero flux (INC semantics)
    f0 = rng.uniform(-1e-3, 1e-3, size=(n, 5))
    f = f0.copy()
    ref.compute_flux_edge(e2n, var, ew, f, c)
    out["flux0"], out["flux_edge"] = f0, f.copy()
    ref.compute_bnd_node_flux(grp, b2n, bw, var, f, c)
    out["flux_edge_bnd"] = f
    s = f0.copy()
    ref.unstructured_stream(e2n, var, ew, s)
    out["stream"] = s
    # time stepping
    dt = np.zeros(n)
    ref.calculate_dt(var, vol, dt)
    out["dt"] = dt
    mdt = ref.get_min_dt(dt)
    out["min_dt"] = np.array([mdt])
    sf = np.zeros(n)
    ref.compute_step_factor(var, vol, mdt, sf)
    out["sf"] = sf
    old = perturbed_state(n, ffv, seed + 1)
    out["old"] = old
    for rk in range(3):
        fl = out["flux_edge_bnd"].copy()
        v = var.copy()
        ref.time_step(rk, sf, fl, old, v)
        out[f"ts_var_rk{rk}"], out[f"ts_flux_rk{rk}"] = v, fl
    res = np.zeros((n, 5))
    ref.residual(old, var, res)
    out["residual"] = res
    out["rms_sum"] = np.array([ref.calc_rms_sum(res)])
    # multigrid transfer 0 -> 1 and back
    f2c = np.ascontiguousarray(m0["node_to_mg_node"], np.int32)
    vc = perturbed_state(n1, ffv, seed + 2)
    scr = np.zeros(n1, np.int32)
    out["f2c"], out["var_c_in"], out["coords_c"] = f2c, vc.copy(), np.ascontiguousarray(m1["coords"])
    ref.up_pre(f2c, vc, scr)
    ref.up(f2c, var, vc, scr)
    out["up_scratch"] = scr.copy()
    ref.up_post(vc, scr)
    out["var_c"] = vc
    rc = perturbed_state(n1, ffv, seed + 3) - vc
    vf = var.copy()
    ref.down(f2c, vf, res, out["coords"], rc, out["coords_c"])
    out["res_c"], out["down_var"] = rc, vf
    # inverse-distance prolongation over edge-->mg_nodes = node-->mg_node o edge-->node
    e2mg = np.ascontiguousarray(f2c[e2n], np.int32)
    rp, ws = np.full((n, 5), 9.0), np.full(n, 9.0)
    ref.down_v2_pre(ws, rp)
    ref.down_v2(e2n, e2mg, out["coords"], out["coords_c"], rc, rp, ws)
    vf2 = var.copy()
    ref.down_v2_post(rp, ws, res, vf2)
    out["dv2_rp"], out["dv2_wsum"], out["dv2_var"] = rp, ws, vf2
    # -v comparator on a perturbed copy
    test = var.copy()
    test[::7, 2] *= 1 + 3e-7
    test[::11, 4] *= 1 + 5e-8
    diff = np.zeros((n, 5))
    ref.identify_differences(test, var, diff)
    out["v_test"], out["v_diff"] = test, diff
    out["v_count"] = np.array([ref.count_non_zeros(diff)])
    return out


def solve_vectors(ref: Oracle, mesh: list[dict], cycles: int) -> dict:
    r = ref.solve(mesh, cycles, MACH, AOA)
    assert r["status"] == 0
    out = dict(status=np.array([r["status"]]), rms=r["rms"], min_dt=r["min_dt"], cycles=np.array([cycles]))
    for l, m in enumerate(mesh):
        out[f"L{l}_coords"] = np.ascontiguousarray(m["coords"])
        out[f"L{l}_e2n"] = np.ascontiguousarray(m["edge_to_node"], np.int32)
        out[f"L{l}_ew"] = np.ascontiguousarray(m["edge_weights"])
        out[f"L{l}_b2n"] = np.ascontiguousarray(m["bnd_to_node"], np.int32)
        out[f"L{l}_bg"] = np.ascontiguousarray(m["bnd_group"], np.int32)
        out[f"L{l}_bw"] = np.ascontiguousarray(m["bnd_weights"])
        if m.get("node_to_mg_node") is not None:
            out[f"L{l}_f2c"] = np.ascontiguousarray(m["node_to_mg_node"], np.int32)
        out[f"L{l}_variables"] = r["levels"][l]["variables"]
        out[f"L{l}_volumes"] = r["levels"][l]["volumes"]
    return out


def main():
    if not Oracle.available("ref"):
        sys.exit("oracle/_ref/libmgcfd_ref.so is missing: run `make -C oracle` with the reference checkout present")
    ref = Oracle("ref")
    box = mgcfd.mesh_box(6, 2, seed=21)
    knn = mgcfd.mesh_knn(400, 1240, 3, seed=22)
    np.savez_compressed(os.path.join(HERE, "kernels_box6.npz"), **kernel_vectors(ref, box, 31))
    np.savez_compressed(os.path.join(HERE, "kernels_knn400.npz"), **kernel_vectors(ref, knn, 32))
    np.savez_compressed(os.path.join(HERE, "solve_box8_g4.npz"), **solve_vectors(ref, mgcfd.mesh_box(8, 3, seed=23), 4))
    np.savez_compressed(os.path.join(HERE, "solve_knn400_g3.npz"), **solve_vectors(ref, knn, 3))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
