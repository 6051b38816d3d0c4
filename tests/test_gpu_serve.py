"""Served solves: continuous batching (mgcfd_server_*).

Requests of uneven size and precision arrive while others run, wait for a
slot or for memory, and finish independently.  Bars: an FP64 (GATHER_EXACT)
request's level-0 state within 1e-12 relative per variable of the oracle, a
mixed-precision one within 1e-6; every request bitwise equal to the same
request served alone; the device-side convergence stop gives exactly the
state of a run of that many cycles.
"""
import time

import numpy as np
import pytest

import mgcfd
from oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(g, o):
    return float((np.abs(g - o).max(axis=0) / np.abs(o).max(axis=0)).max())


@pytest.fixture(scope="module")
def meshes():
    return [mgcfd.mesh_knn(3000, 9300, 3, seed=11), mgcfd.mesh_box(8, 3, seed=12),
            mgcfd.mesh_knn(1500, 4650, 3, seed=13), mgcfd.mesh_knn(2200, 6820, 3, seed=14)]


REQS = [(0, 0.5, 3.0, "gather_exact"), (1, 0.7, -2.0, "mixed_coloured"), (2, 0.35, 5.0, "gather_exact"),
        (3, 0.6, 1.0, "mixed_coloured")]
CYCLES = 5


@pytest.fixture
def rt():
    r = mgcfd.Runtime()
    yield r
    r.close()


def drain(srv, ids, timeout=120.0, max_active=None):
    t0 = time.time()
    out = {}
    while len(out) < len(ids):
        a, q = srv.progress()
        if max_active is not None:
            assert a <= max_active
        for i in ids:
            if i not in out:
                r = srv.result(i)
                if r is not None:
                    out[i] = r
        assert time.time() - t0 < timeout, "served requests did not finish"
        time.sleep(0.001)
    return out


def test_served_results_match_oracle_and_alone(rt, meshes):
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=2, chunk=2)
    ids = []
    for k, (mi, mach, aoa, var) in enumerate(REQS):
        ids.append(srv.submit(meshes[mi], mach, aoa, var, max_cycles=CYCLES))
        srv.progress()                       # later requests arrive while earlier ones run
    res = drain(srv, ids, max_active=2)
    srv.close()
    orc = Oracle("orc")
    for k, (mi, mach, aoa, var) in enumerate(REQS):
        r = res[ids[k]]
        assert r["status"] == 0 and r["cycles"] == CYCLES
        ref = orc.solve(meshes[mi], CYCLES, mach, aoa)
        err = rel(r["variables"], ref["levels"][0]["variables"])
        print(f"request {k} ({var}): state error {err:.3e}, latency {r['latency'] * 1e3:.1f} ms")
        assert err <= (1e-12 if var == "gather_exact" else 1e-6)
        alone = mgcfd.Server(rt, budget_bytes=200e9, max_active=1, chunk=3)
        j = alone.submit(meshes[mi], mach, aoa, var, max_cycles=CYCLES)
        a = drain(alone, [j])[j]
        alone.close()
        assert np.array_equal(a["variables"].view(np.uint64), r["variables"].view(np.uint64)), k


def test_memory_budget_serialises(rt, meshes):
    est = [mgcfd.Server.estimate_bytes(meshes[i]) for i in range(3)]
    # room for the largest request, never for two at once
    budget = max(est) + 0.5 * min(est)
    srv = mgcfd.Server(rt, budget_bytes=budget, max_active=4, chunk=2)
    ids = [srv.submit(meshes[i], 0.5, 3.0, "gather_exact", max_cycles=3) for i in range(3)]
    res = drain(srv, ids, max_active=1)
    srv.close()
    assert all(res[i]["status"] == 0 for i in ids)


def test_request_over_budget_refused(rt, meshes):
    """A request whose estimate exceeds the whole budget is refused at submit
    (it could never be admitted), and the server keeps serving the others."""
    small = mgcfd.mesh_knn(300, 930, 3, seed=15)
    est, est_small = mgcfd.Server.estimate_bytes(meshes[0]), mgcfd.Server.estimate_bytes(small)
    assert est_small < est
    srv = mgcfd.Server(rt, budget_bytes=0.5 * (est + est_small), max_active=2, chunk=2)
    with pytest.raises(mgcfd.MgcfdError, match="more than the whole budget"):
        srv.submit(meshes[0], 0.5, 3.0, "gather_exact", max_cycles=3)
    j = srv.submit(small, 0.5, 3.0, "gather_exact", max_cycles=3)
    r = drain(srv, [j])[j]
    srv.close()
    assert r["status"] == 0 and r["cycles"] == 3


def test_admitted_requests_overlap_on_device(rt, meshes):
    """Requests admitted together run concurrently: their device intervals
    (HIP events on each request's stream) intersect."""
    big = [mgcfd.mesh_knn(20000, 62000, 3, seed=16 + k) for k in range(3)]
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=4, chunk=400)
    # all three are queued before the first progress call admits them
    ids = [srv.submit(m, 0.5, 3.0, "part_coloured", max_cycles=400) for m in big]
    res = drain(srv, ids, max_active=4)
    spans = [srv.timing(i) for i in ids]
    srv.close()
    print("device intervals (ms):", spans)
    assert all(res[i]["status"] == 0 for i in ids)
    assert all(0 <= a < b for a, b in spans)
    # the later requests started on the device before the first one ended
    assert spans[1][0] < spans[0][1] and spans[2][0] < spans[0][1]


def test_convergence_stop_is_exact(rt, meshes):
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=1, chunk=4)
    # this mesh's L0 RMS falls by ~3.6e-7 of its first value per cycle (oracle)
    drop = 1.0 - 1.5e-6
    j = srv.submit(meshes[0], 0.5, 3.0, "gather_exact", max_cycles=40, rms_drop=drop)
    r = drain(srv, [j])[j]
    srv.close()
    c = r["cycles"]
    print("converged after", c, "cycles; rms / rms[0]", r["rms"] / r["rms"][0])
    assert r["status"] == 3 and 2 <= c < 40
    assert r["rms"][c - 1] <= drop * r["rms"][0]
    assert all(r["rms"][k] > drop * r["rms"][0] for k in range(1, c - 1))
    ref = Oracle("orc").solve(meshes[0], c, 0.5, 3.0)
    assert rel(r["variables"], ref["levels"][0]["variables"]) <= 1e-12
