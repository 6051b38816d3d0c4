"""mgcfd_get_min_dt_noalloc: the get_min_dt loop without any device
allocation (its per-workgroup minima in the library's fixed device array),
against op_par_loop_get_min_dt_kernel (time_stepping_kernels.h:34-41, site
euler3d.cpp:477-479) through the C ABI.  min is exact: results must be equal.
Two calls in a row, the second on a smaller set, so stale minima of the first
call left in the shared array must not leak into the second."""
import ctypes as C

import numpy as np
import pytest

import mgcfd

pytestmark = pytest.mark.gpu

DBL_MAX = np.finfo(np.float64).max


def noalloc(rt, st, dat, init):
    g = np.array([init])
    rc = rt.L.mgcfd_get_min_dt_noalloc(b"get_min_dt_kernel", st, rt.arg(dat, -1, None, 1),
                                       rt.gbl(g, mgcfd.OP_MIN))
    rt._keep.clear()
    if rc != 0:
        raise mgcfd.MgcfdError(mgcfd.last_error())
    return g[0]


def with_loop(rt, st, dat, init):
    g = np.array([init])
    rt.par_loop("get_min_dt_kernel", st, rt.arg(dat, -1, None, 1), rt.gbl(g, mgcfd.OP_MIN))
    return g[0]


def test_min_dt_noalloc_matches_the_loop():
    rng = np.random.default_rng(3)
    rt = mgcfd.Runtime()
    try:
        big_n, small_n = 1_500_000, 70_001
        big_v = rng.uniform(1e-3, 2.0, big_n)
        small_v = rng.uniform(0.5, 3.0, small_n)       # its minimum is above most of big's minima
        big_v[123_457] = 1e-9                           # a minimum in a late workgroup
        small_v[5] = 0.25
        big = rt.decl_set(big_n, "big")
        small = rt.decl_set(small_n, "small")
        dbig = rt.decl_dat(big, 1, big_v.reshape(-1, 1), "dt_big")
        dsmall = rt.decl_dat(small, 1, small_v.reshape(-1, 1), "dt_small")
        # the first loop renumbers and uploads the declared objects
        assert with_loop(rt, big, dbig, DBL_MAX) == big_v.min()
        for st, d, v in ((big, dbig, big_v), (small, dsmall, small_v)):
            a = with_loop(rt, st, d, DBL_MAX)
            b = noalloc(rt, st, d, DBL_MAX)
            print(f"set of {len(v)}: loop {a!r}, no-allocation entry {b!r}")
            assert a == b == v.min()
        # the gbl argument's initial value takes part in the minimum, as in the loop
        assert noalloc(rt, small, dsmall, 0.125) == with_loop(rt, small, dsmall, 0.125) == 0.125
        # negative values are kept (the driver's abort test reads them)
        small_v2 = small_v.copy()
        small_v2[9] = -4.0
        rt.upload(dsmall, small_v2.reshape(-1, 1))
        assert noalloc(rt, small, dsmall, DBL_MAX) == with_loop(rt, small, dsmall, DBL_MAX) == -4.0
        # a declaration since the last loop is refused, not served by allocating
        extra = rt.decl_set(10, "extra")
        rt.decl_dat(extra, 1, np.ones((10, 1)), "dt_extra")
        with pytest.raises(mgcfd.MgcfdError, match="declared since the last loop"):
            noalloc(rt, small, dsmall, DBL_MAX)
    finally:
        rt.close()


def test_min_dt_noalloc_refuses_sets_beyond_its_scratch():
    rt = mgcfd.Runtime()
    try:
        n = 65536 * 256 + 1
        st = rt.decl_set(n, "huge")
        d = rt.decl_dat(st, 1, np.ones((n, 1)), "dt_huge")
        assert with_loop(rt, st, d, DBL_MAX) == 1.0
        with pytest.raises(mgcfd.MgcfdError, match="more than the 16777216"):
            noalloc(rt, st, d, DBL_MAX)
    finally:
        rt.close()
