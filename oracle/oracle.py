"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/libmgcfd_oracle.so (this project's C restatement of the
reference kernels, oracle/mgcfd_oracle.c) and, where it was built,
oracle/_ref/libmgcfd_ref.so (the reference's own kernel headers compiled in
place, oracle/ref_harness.cpp).  Both expose the same functions (orc_* /
ref_*) on the reference's AoS host layout.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "libmgcfd_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libmgcfd_ref.so")


class Consts(C.Structure):
    _fields_ = [("smoothing_coefficient", C.c_double), ("ff_variable", C.c_double * 5),
                ("ff_flux_contribution_momentum_x", C.c_double * 3),
                ("ff_flux_contribution_momentum_y", C.c_double * 3),
                ("ff_flux_contribution_momentum_z", C.c_double * 3),
                ("ff_flux_contribution_density_energy", C.c_double * 3)]

    def as_dict(self) -> dict:
        return {k: (np.array(getattr(self, k)[:]) if k != "smoothing_coefficient" else self.smoothing_coefficient)
                for k, _ in self._fields_}


_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)


class Level(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_edges", C.c_int), ("n_bnd", C.c_int),
                ("edge_to_node", _PI), ("bnd_to_node", _PI), ("bnd_group", _PI), ("node_to_mg_node", _PI),
                ("coords", _PD), ("edge_weights", _PD), ("bnd_weights", _PD), ("variables", _PD),
                ("old_variables", _PD), ("residuals", _PD), ("fluxes", _PD), ("step_factors", _PD),
                ("volumes", _PD), ("up_scratch", _PI),
                ("edge_to_mg_nodes", _PI), ("prolonged", _PD), ("prolonged_wsum", _PD)]


def _p(a: np.ndarray):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_PD)
    return a.ctypes.data_as(_PI)


class Oracle:
    """prefix 'orc' = C restatement, 'ref' = reference kernels compiled in place."""

    def __init__(self, which: str = "orc"):
        path = ORACLE_LIB if which == "orc" else REF_LIB
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.p = which
        f = self.fn
        f("freestream").argtypes = [C.c_double, C.c_double, C.POINTER(Consts)]
        f("get_min_dt").restype = C.c_double
        f("calc_rms_sum").restype = C.c_double
        f("count_bad_vals").restype = C.c_int
        f("count_non_zeros").restype = C.c_int
        f("mg_run").argtypes = [C.POINTER(Level), C.c_int, C.c_int, C.POINTER(Consts), _PD, _PD, _PI]
        f("setup").argtypes = [C.POINTER(Level), C.c_int, C.POINTER(Consts)]

    def fn(self, name):
        return getattr(self.lib, f"{self.p}_{name}")

    @staticmethod
    def available(which: str) -> bool:
        return os.path.exists(ORACLE_LIB if which == "orc" else REF_LIB)

    def freestream(self, mach=1.2, aoa=0.0) -> Consts:
        c = Consts()
        self.fn("freestream")(C.c_double(mach), C.c_double(aoa), C.byref(c))
        return c

    # --- per-loop calls (arrays are modified in place, AoS layout)
    def initialize_variables(self, var, c):
        self.fn("initialize_variables")(len(var), C.byref(c), _p(var))

    def calculate_cell_volumes(self, e2n, coords, ew, vol):
        self.fn("calculate_cell_volumes")(len(e2n), _p(e2n), _p(coords), _p(ew), _p(vol))

    def dampen_ewt(self, w):
        self.fn("dampen_ewt")(len(w), _p(w))

    def calculate_dt(self, var, vol, dt):
        self.fn("calculate_dt")(len(var), _p(var), _p(vol), _p(dt))

    def get_min_dt(self, dt) -> float:
        return self.fn("get_min_dt")(len(dt), _p(dt))

    def compute_step_factor(self, var, vol, min_dt, sf):
        self.fn("compute_step_factor")(len(vol), _p(var), _p(vol), C.c_double(min_dt), _p(sf))

    def compute_flux_edge(self, e2n, var, ew, flux, c):
        self.fn("compute_flux_edge")(len(e2n), _p(e2n), _p(var), _p(ew), _p(flux), C.byref(c))

    def compute_bnd_node_flux(self, group, b2n, bw, var, flux, c):
        self.fn("compute_bnd_node_flux")(len(b2n), _p(group), _p(b2n), _p(bw), _p(var), _p(flux), C.byref(c))

    def time_step(self, rk, sf, flux, old, var):
        self.fn("time_step")(len(var), rk, _p(sf), _p(flux), _p(old), _p(var))

    def unstructured_stream(self, e2n, var, ew, flux):
        self.fn("unstructured_stream")(len(e2n), _p(e2n), _p(var), _p(ew), _p(flux))

    def copy_double(self, var, old):
        self.fn("copy_double")(len(var), _p(var), _p(old))

    def residual(self, old, var, res):
        self.fn("residual")(len(var), _p(old), _p(var), _p(res))

    def calc_rms_sum(self, res) -> float:
        return self.fn("calc_rms_sum")(len(res), _p(res))

    def count_bad_vals(self, var) -> int:
        return self.fn("count_bad_vals")(len(var), _p(var))

    def up_pre(self, f2c, var_c, scratch_c):
        self.fn("up_pre")(len(f2c), _p(f2c), _p(var_c), _p(scratch_c))

    def up(self, f2c, var_f, var_c, scratch_c):
        self.fn("up")(len(f2c), _p(f2c), _p(var_f), _p(var_c), _p(scratch_c))

    def up_post(self, var_c, scratch_c):
        self.fn("up_post")(len(var_c), _p(var_c), _p(scratch_c))

    def down(self, f2c, var_f, res_f, xyz_f, res_c, xyz_c):
        self.fn("down")(len(f2c), _p(f2c), _p(var_f), _p(res_f), _p(xyz_f), _p(res_c), _p(xyz_c))

    def down_v2_pre(self, wsum, rp):
        self.fn("down_v2_pre")(len(wsum), _p(wsum), _p(rp))

    def down_v2(self, e2n, e2mg, xyz_f, xyz_c, res_c, rp, wsum):
        self.fn("down_v2")(len(e2n), _p(e2n), _p(e2mg), _p(xyz_f), _p(xyz_c), _p(res_c), _p(rp), _p(wsum))

    def down_v2_post(self, rp, wsum, res_f, var_f):
        self.fn("down_v2_post")(len(wsum), _p(rp), _p(wsum), _p(res_f), _p(var_f))

    def identify_differences(self, test, master, diff):
        self.fn("identify_differences")(len(test), _p(test), _p(master), _p(diff))

    def count_non_zeros(self, diff) -> int:
        return self.fn("count_non_zeros")(len(diff), _p(diff))

    # --- multi-threaded CPU baseline (reference kernel, coloured edges; 'ref' only)
    def colour_edges(self, e2n, n_nodes, max_colours=64):
        order = np.zeros(len(e2n), np.int32)
        cptr = np.zeros(max_colours + 1, np.int32)
        nc = self.fn("colour_edges")(len(e2n), n_nodes, _p(e2n), max_colours, _p(order), _p(cptr))
        if nc < 0:
            raise RuntimeError("edge colouring needs more than %d colours" % max_colours)
        return order, cptr[:nc + 1].copy()

    def compute_flux_edge_omp(self, colouring, e2n, var, ew, flux, c, threads):
        order, cptr = colouring
        self.fn("compute_flux_edge_omp")(len(cptr) - 1, _p(cptr), _p(order), _p(e2n), _p(var), _p(ew), _p(flux),
                                         C.byref(c), threads)

    # --- whole solve
    @staticmethod
    def _levels(mesh: list[dict], down_v2: bool = False):
        """Working copies of every level's arrays and the C level table over them."""
        nl = len(mesh)
        arrs = []
        levels = (Level * nl)()
        for i, m in enumerate(mesh):
            n, e, b = len(m["coords"]), len(m["edge_to_node"]), len(m["bnd_to_node"])
            a = dict(
                e2n=np.ascontiguousarray(m["edge_to_node"], dtype=np.int32),
                b2n=np.ascontiguousarray(m["bnd_to_node"], dtype=np.int32),
                bg=np.ascontiguousarray(m["bnd_group"], dtype=np.int32),
                n2c=(np.ascontiguousarray(m["node_to_mg_node"], dtype=np.int32)
                     if m.get("node_to_mg_node") is not None else None),
                xyz=np.ascontiguousarray(m["coords"], dtype=np.float64),
                ew=np.array(m["edge_weights"], dtype=np.float64, copy=True),
                bw=np.array(m["bnd_weights"], dtype=np.float64, copy=True),
                var=np.zeros((n, 5)), old=np.zeros((n, 5)), res=np.zeros((n, 5)), flux=np.zeros((n, 5)),
                sf=np.zeros(n), vol=np.zeros(n), scr=np.zeros(n, dtype=np.int32), e2mg=None,
                rp=np.zeros((n, 5)), wsum=np.zeros(n))
            if down_v2 and a["n2c"] is not None:
                a["e2mg"] = np.ascontiguousarray(a["n2c"][a["e2n"]], dtype=np.int32)
            arrs.append(a)
            L = levels[i]
            L.n_nodes, L.n_edges, L.n_bnd = n, e, b
            L.edge_to_node, L.bnd_to_node, L.bnd_group = _p(a["e2n"]), _p(a["b2n"]), _p(a["bg"])
            L.node_to_mg_node = _p(a["n2c"]) if a["n2c"] is not None else None
            L.coords, L.edge_weights, L.bnd_weights = _p(a["xyz"]), _p(a["ew"]), _p(a["bw"])
            L.variables, L.old_variables, L.residuals = _p(a["var"]), _p(a["old"]), _p(a["res"])
            L.fluxes, L.step_factors, L.volumes, L.up_scratch = _p(a["flux"]), _p(a["sf"]), _p(a["vol"]), _p(a["scr"])
            L.edge_to_mg_nodes = _p(a["e2mg"]) if a["e2mg"] is not None else None
            L.prolonged, L.prolonged_wsum = _p(a["rp"]), _p(a["wsum"])
        return arrs, levels

    def solve(self, mesh: list[dict], cycles: int, mach=1.2, aoa=0.0, consts: Consts | None = None,
              after_setup=None, down_v2=False, checkpoints=()):
        """Setup (euler3d.cpp:412-441) + multigrid loop (:452-641) on copies of
        the mesh arrays.  after_setup(list of per-level array dicts) may modify
        the state between the two (failure-path tests).  down_v2 prolongs with
        the inverse-distance loops (euler3d.cpp:603-623) over the edge-->mg_nodes
        map e2mg[e][k] = node_to_mg_node[e2n[e][k]].  checkpoints: cycle
        counts below 'cycles' after which every level's variables are copied
        (each multigrid cycle starts and ends at level 0, so the run in pieces
        is the run in one).  Returns dict(status, rms, min_dt, steps,
        levels=[...], snapshots={cycle: [variables per level]})."""
        c = consts if consts is not None else self.freestream(mach, aoa)
        nl = len(mesh)
        arrs, levels = self._levels(mesh, down_v2)
        self.fn("setup")(levels, nl, C.byref(c))
        if after_setup is not None:
            after_setup(arrs)
        rms = np.zeros(max(cycles, 1))
        mdt = np.zeros(max(1, cycles * max(1, 2 * nl - 2) + 8))
        steps = C.c_int(0)
        snaps, done, nsteps, st = {}, 0, 0, 0
        for k in sorted({int(x) for x in checkpoints if 0 < x < cycles}) + [cycles]:
            st = self.fn("mg_run")(levels, nl, k - done, C.byref(c), _p(rms[done:]), _p(mdt[nsteps:]),
                                   C.byref(steps))
            nsteps += steps.value
            done = k
            if st != 0:
                break
            if k < cycles:
                snaps[k] = [a["var"].copy() for a in arrs]
        steps.value = nsteps
        out = []
        for a in arrs:
            out.append(dict(variables=a["var"], old_variables=a["old"], residuals=a["res"], fluxes=a["flux"],
                            step_factors=a["sf"], volumes=a["vol"], edge_weights=a["ew"], bnd_weights=a["bw"],
                            up_scratch=a["scr"], prolonged=a["rp"], prolonged_wsum=a["wsum"]))
        return dict(status=st, rms=rms[:cycles], min_dt=mdt[:steps.value], steps=steps.value, levels=out,
                    snapshots=snaps)

    @staticmethod
    def renumber(mesh: list[dict]):
        """The mesh with every level's nodes in Z-order of their coordinates,
        edges sorted by their renumbered first node and boundary elements by
        node -- what op_renumber does for OP2's CPU backends (euler3d.cpp:
        340-377) so that a block of consecutive edges touches a compact patch
        of nodes.  Returns (mesh, perms) with perms[l][new] = old."""
        perms = []
        for m in mesh:
            x = np.asarray(m["coords"], dtype=np.float64)
            lo, hi = x.min(axis=0), x.max(axis=0)
            q = ((x - lo) / np.maximum(hi - lo, 1e-300) * (2 ** 21 - 1)).astype(np.uint64)
            key = np.zeros(len(x), dtype=np.uint64)
            for bit in range(21):
                for ax in range(3):
                    key |= ((q[:, ax] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + ax)
            perms.append(np.argsort(key, kind="stable"))
        out = []
        for l, m in enumerate(mesh):
            p = perms[l]
            inv = np.empty_like(p)
            inv[p] = np.arange(len(p))
            e2n = inv[np.asarray(m["edge_to_node"])]
            eo = np.argsort(e2n[:, 0], kind="stable")
            b2n = inv[np.asarray(m["bnd_to_node"]).reshape(-1)]
            bo = np.argsort(b2n, kind="stable")
            r = dict(coords=np.asarray(m["coords"])[p], edge_to_node=e2n[eo],
                     edge_weights=np.asarray(m["edge_weights"])[eo], bnd_to_node=b2n[bo],
                     bnd_group=np.asarray(m["bnd_group"]).reshape(-1)[bo],
                     bnd_weights=np.asarray(m["bnd_weights"])[bo])
            if m.get("node_to_mg_node") is not None and l + 1 < len(mesh):
                pc = perms[l + 1]
                invc = np.empty_like(pc)
                invc[pc] = np.arange(len(pc))
                r["node_to_mg_node"] = invc[np.asarray(m["node_to_mg_node"]).reshape(-1)[p]]
            out.append(r)
        return out, perms

    def solve_omp(self, mesh: list[dict], cycles: int, mach=1.2, aoa=0.0, threads: int = 1):
        """The reference kernels' whole solve with every loop spread over
        'threads' host threads as OP2's OpenMP backend runs it, on the mesh
        renumbered as op_renumber would (ref library only).  A timing
        baseline: the block colours change the order of each node's flux
        increments.  Setup, renumbering and the thread plan are outside the
        timed region.  Returns dict(status, rms, seconds, levels=[variables
        per level, original node order])."""
        if self.p != "ref":
            raise RuntimeError("solve_omp needs the reference-kernel library")
        import time
        c = self.freestream(mach, aoa)
        nl = len(mesh)
        mesh, perms = self.renumber(mesh)
        arrs, levels = self._levels(mesh)
        self.fn("setup")(levels, nl, C.byref(c))
        self.lib.ref_omp_plan_create.restype = C.c_void_p
        self.lib.ref_omp_plan_create.argtypes = [C.POINTER(Level), C.c_int]
        self.lib.ref_omp_plan_free.argtypes = [C.c_void_p]
        self.lib.ref_mg_run_omp.argtypes = [C.c_void_p, C.POINTER(Level), C.c_int, C.c_int, C.POINTER(Consts), _PD,
                                            C.c_int]
        plan = self.lib.ref_omp_plan_create(levels, nl)
        if not plan:
            raise RuntimeError("solve_omp: edge-block colouring needs more than 256 colours")
        rms = np.zeros(max(cycles, 1))
        t = time.perf_counter()
        st = self.lib.ref_mg_run_omp(plan, levels, nl, cycles, C.byref(c), _p(rms), threads)
        el = time.perf_counter() - t
        self.lib.ref_omp_plan_free(plan)
        states = []
        for a, p in zip(arrs, perms):
            v = np.empty_like(a["var"])
            v[p] = a["var"]
            states.append(v)
        return dict(status=st, rms=rms[:cycles], seconds=el, levels=states)
