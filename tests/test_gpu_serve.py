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
    """Requests in contexts of their own (no grouping) run concurrently:
    their device intervals (HIP events on each context's stream) intersect."""
    big = [mgcfd.mesh_knn(20000, 62000, 3, seed=16 + k) for k in range(3)]
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=4, chunk=400, max_group=1)
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


def serve_alone(rt, mesh, mach, aoa, var, **kw):
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=1, chunk=3, max_group=1)
    j = srv.submit(mesh, mach, aoa, var, **kw)
    r = drain(srv, [j])[j]
    srv.close()
    return r


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def test_queued_requests_share_a_context(rt, meshes):
    """Requests queued behind a busy server are admitted as groups: the
    compatible ones (same variant, levels, cycle cap and drop) share one
    context, whose launches serve them all (identical device intervals), and
    each result is bitwise the request served alone."""
    reqs = [(0, 0.5, 3.0, "part_coloured"), (2, 0.4, 1.0, "mixed_coloured"), (3, 0.6, -1.0, "part_coloured"),
            (1, 0.7, 2.0, "part_coloured"), (0, 0.3, 4.0, "mixed_coloured")]
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=1, chunk=5)
    ids = [srv.submit(meshes[mi], mach, aoa, var, max_cycles=CYCLES) for mi, mach, aoa, var in reqs]
    res = drain(srv, ids, max_active=3)
    spans = [srv.timing(i) for i in ids]
    srv.close()
    print("device intervals (ms):", spans)
    assert spans[0] == spans[2] == spans[3]          # the three FP64 requests: one group
    assert spans[1] == spans[4] and spans[1] != spans[0]
    for k, (mi, mach, aoa, var) in enumerate(reqs):
        r = res[ids[k]]
        a = serve_alone(rt, meshes[mi], mach, aoa, var, max_cycles=CYCLES)
        assert r["status"] == a["status"] == 0 and r["cycles"] == a["cycles"] == CYCLES
        assert same(r["variables"], a["variables"]), k
        assert same(r["rms"], a["rms"]), k


def test_grouped_members_stop_at_their_own_convergence(rt, meshes):
    """Members of one group converge at different cycles; each one's state is
    copied on the device at its own convergence cycle while the others run
    on, so status, cycle count, RMS history and state equal the request
    served alone (whose convergence halts its context)."""
    drop = 1.0 - 1.5e-6
    reqs = [(0, 0.5, 3.0), (2, 0.5, 3.0), (3, 0.45, 2.0), (1, 0.5, 3.0)]
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=1, chunk=4)
    ids = [srv.submit(meshes[mi], mach, aoa, "gather_exact", max_cycles=40, rms_drop=drop) for mi, mach, aoa in reqs]
    res = drain(srv, ids)
    spans = [srv.timing(i) for i in ids]
    srv.close()
    cycles = [res[i]["cycles"] for i in ids]
    print("cycles per member:", cycles, "statuses:", [res[i]["status"] for i in ids])
    assert spans[0] == spans[1] == spans[2] == spans[3]     # all queued together: one group
    # the oracle's RMS histories put the stops at 6, 37 and 14 cycles, and
    # the third member at its 40-cycle cap
    assert len(set(cycles)) == 4
    for k, (mi, mach, aoa) in enumerate(reqs):
        r = res[ids[k]]
        a = serve_alone(rt, meshes[mi], mach, aoa, "gather_exact", max_cycles=40, rms_drop=drop)
        assert (r["status"], r["cycles"]) == (a["status"], a["cycles"]), k
        assert same(r["variables"], a["variables"]), k
        assert same(r["rms"], a["rms"]), k


def test_member_abort_is_served_alone(rt, meshes):
    """A member whose solve aborts (a NaN edge weight: bad values, status 2)
    is served again alone, so its result is the lone request's; the other
    members of its group are unaffected."""
    bad = [dict(l) for l in meshes[2]]
    w = np.array(bad[0]["edge_weights"], dtype=np.float64, copy=True)
    w[17, 1] = np.nan
    bad[0]["edge_weights"] = w
    reqs = [(meshes[0], 0.5, 3.0), (meshes[3], 0.5, 3.0), (bad, 0.5, 3.0), (meshes[1], 0.6, 1.0)]
    srv = mgcfd.Server(rt, budget_bytes=200e9, max_active=1, chunk=2)
    ids = [srv.submit(m, mach, aoa, "part_coloured", max_cycles=CYCLES) for m, mach, aoa in reqs]
    res = drain(srv, ids)
    spans = [srv.timing(i) for i in ids]
    srv.close()
    print("statuses:", [res[i]["status"] for i in ids], "intervals:", spans)
    assert res[ids[2]]["status"] == 2
    assert spans[0] == spans[1] == spans[3] and spans[2] != spans[1]    # re-served after its group
    for k, (m, mach, aoa) in enumerate(reqs):
        r = res[ids[k]]
        a = serve_alone(rt, m, mach, aoa, "part_coloured", max_cycles=CYCLES)
        assert (r["status"], r["cycles"]) == (a["status"], a["cycles"]), k
        assert same(r["variables"], a["variables"]), k
