"""Python binding of libmgcfd (the MI355X MG-CFD hot path) over its C ABI.

The library exposes the OP2 API surface that the reference's euler3d.cpp
uses (sets, maps, dats, ``op_par_loop_<kernel>`` per loop site) with
hand-written HIP kernels behind it.  This module is a thin ctypes mirror of
that ABI plus helpers that declare a multigrid mesh the way euler3d.cpp does
(euler3d.cpp:243-441) and run its loops.

Nothing here computes on the host: every loop goes to the HIP library, and a
missing library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libmgcfd.so")

OP_READ, OP_WRITE, OP_RW, OP_INC, OP_MIN, OP_MAX = range(6)
FLUX_GATHER_EXACT, FLUX_GATHER_FAST, FLUX_ATOMIC, FLUX_NODAL_EXACT = 0, 1, 2, 3
FLUX_VARIANTS = {"gather_exact": 0, "gather_fast": 1, "atomic": 2, "nodal_exact": 3, "part_coloured": 4,
                 "part_atomic": 5, "mixed_coloured": 6, "mixed_atomic": 7}
FP64_FLUX_VARIANTS = {k: v for k, v in FLUX_VARIANTS.items() if not k.startswith("mixed")}


class OpSetCore(C.Structure):
    _fields_ = [("index", C.c_int), ("size", C.c_int), ("name", C.c_char_p), ("impl", C.c_void_p)]


class OpMapCore(C.Structure):
    _fields_ = [("index", C.c_int), ("from_", C.POINTER(OpSetCore)), ("to", C.POINTER(OpSetCore)),
                ("dim", C.c_int), ("name", C.c_char_p), ("impl", C.c_void_p)]


class OpDatCore(C.Structure):
    _fields_ = [("index", C.c_int), ("set", C.POINTER(OpSetCore)), ("dim", C.c_int), ("size", C.c_int),
                ("type", C.c_char_p), ("name", C.c_char_p), ("impl", C.c_void_p)]


OpSet = C.POINTER(OpSetCore)
OpMap = C.POINTER(OpMapCore)
OpDat = C.POINTER(OpDatCore)


class OpArg(C.Structure):
    _fields_ = [("argtype", C.c_int), ("dat", OpDat), ("map", OpMap), ("idx", C.c_int), ("dim", C.c_int),
                ("size", C.c_int), ("data", C.c_void_p), ("type", C.c_char_p), ("acc", C.c_int)]


class LevelBind(C.Structure):
    _fields_ = [("nodes", OpSet), ("edges", OpSet), ("bnd_nodes", OpSet),
                ("edge_to_node", OpMap), ("bnd_to_node", OpMap), ("node_to_mg_node", OpMap),
                ("variables", OpDat), ("old_variables", OpDat), ("residuals", OpDat), ("fluxes", OpDat),
                ("step_factors", OpDat), ("volumes", OpDat), ("coords", OpDat), ("edge_weights", OpDat),
                ("bnd_weights", OpDat), ("bnd_groups", OpDat), ("up_scratch", OpDat)]


class MeshLevelC(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_edges", C.c_int), ("n_bnd", C.c_int),
                ("coords", C.POINTER(C.c_double)), ("edge_to_node", C.POINTER(C.c_int)),
                ("edge_weights", C.POINTER(C.c_double)), ("bnd_to_node", C.POINTER(C.c_int)),
                ("bnd_group", C.POINTER(C.c_int)), ("bnd_weights", C.POINTER(C.c_double)),
                ("node_to_mg_node", C.POINTER(C.c_int))]


class MeshC(C.Structure):
    _fields_ = [("levels", C.c_int), ("L", MeshLevelC * 8)]


class MgcfdError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    """Loads libmgcfd.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MgcfdError(f"libmgcfd.so not built: {LIB_PATH} missing (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.mgcfd_last_error.restype = C.c_char_p
    L.mgcfd_mesh_last_error.restype = C.c_char_p
    L.op_init.argtypes = [C.c_int, C.c_void_p, C.c_int]
    L.op_decl_set.restype = OpSet
    L.op_decl_set.argtypes = [C.c_int, C.c_char_p]
    L.op_decl_map.restype = OpMap
    L.mgcfd_decl_map_compose.restype = OpMap
    L.mgcfd_decl_map_compose.argtypes = [OpMap, OpMap, C.c_char_p]
    L.op_decl_map.argtypes = [OpSet, OpSet, C.c_int, C.c_void_p, C.c_char_p]
    L.op_decl_dat_char.restype = OpDat
    L.op_decl_dat_char.argtypes = [OpSet, C.c_int, C.c_char_p, C.c_int, C.c_void_p, C.c_char_p]
    L.op_decl_dat_temp_char.restype = OpDat
    L.op_decl_dat_temp_char.argtypes = [OpSet, C.c_int, C.c_char_p, C.c_int, C.c_char_p]
    L.op_decl_const_char.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_void_p, C.c_char_p]
    L.op_arg_dat.restype = OpArg
    L.op_arg_dat.argtypes = [OpDat, C.c_int, OpMap, C.c_int, C.c_char_p, C.c_int]
    L.op_arg_gbl_char.restype = OpArg
    L.op_arg_gbl_char.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int, C.c_int]
    L.op_fetch_data_char.argtypes = [OpDat, C.c_void_p]
    L.op_upload_data_char.argtypes = [OpDat, C.c_void_p]
    L.op_partition.argtypes = [C.c_char_p, C.c_char_p, OpSet, OpMap, OpDat]
    L.op_get_size.argtypes = [OpSet]
    L.mgcfd_mg_bind.argtypes = [C.c_int, C.POINTER(LevelBind)]
    L.mgcfd_mg_run.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.mgcfd_time_flux.argtypes = [C.c_int, C.c_int, C.POINTER(OpMap), C.POINTER(OpDat), C.POINTER(OpDat),
                                  C.POINTER(OpDat), C.c_int, C.POINTER(C.c_double)]
    L.mgcfd_device_info.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    L.mgcfd_plan_stats.argtypes = [OpMap, C.POINTER(C.c_int)]
    L.mgcfd_part_plan_stats.argtypes = [OpMap, C.POINTER(C.c_int)]
    L.mgcfd_mesh_box.argtypes = [C.c_int, C.c_int, C.c_uint64, C.POINTER(C.POINTER(MeshC))]
    L.mgcfd_mesh_knn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.POINTER(MeshC))]
    L.mgcfd_mesh_free.argtypes = [C.POINTER(MeshC)]
    for name in LOOPS:
        fn = getattr(L, "op_par_loop_" + name)
        fn.restype = C.c_int
    _lib = L
    return L


# op_par_loop entry points and their argument counts (include/mgcfd_op2.h).
LOOPS = {
    "initialize_variables_kernel": 1, "zero_5d_array_kernel": 1, "zero_1d_array_kernel": 1,
    "calculate_cell_volumes": 5, "dampen_ewt": 1, "copy_double_kernel": 2, "calculate_dt_kernel": 3,
    "get_min_dt_kernel": 2, "compute_step_factor_kernel": 4, "compute_flux_edge_kernel": 5,
    "compute_bnd_node_flux_kernel": 4, "time_step_kernel": 5, "unstructured_stream_kernel": 5,
    "residual_kernel": 3, "calc_rms_kernel": 2, "count_bad_vals": 2, "up_pre_kernel": 2, "up_kernel": 3,
    "up_post_kernel": 2, "down_kernel": 5, "down_v2_kernel_pre": 2, "down_v2_kernel": 10,
    "down_v2_kernel_post": 4, "identify_differences": 3, "count_non_zeros": 2,
}


def _check(rc: int, what: str):
    if rc != 0:
        raise MgcfdError(f"{what}: {lib().mgcfd_last_error().decode()}")


def last_error() -> str:
    return lib().mgcfd_last_error().decode()


# --------------------------------------------------------------- meshes --

def _mesh_to_numpy(p) -> list[dict]:
    m = p.contents
    out = []
    for l in range(m.levels):
        L = m.L[l]
        n, e, b = L.n_nodes, L.n_edges, L.n_bnd

        def arr(ptr, count, dtype, shape):
            if count == 0:
                return np.zeros(shape, dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True).reshape(shape)

        lv = dict(
            coords=arr(L.coords, 3 * n, np.float64, (n, 3)),
            edge_to_node=arr(L.edge_to_node, 2 * e, np.int32, (e, 2)),
            edge_weights=arr(L.edge_weights, 3 * e, np.float64, (e, 3)),
            bnd_to_node=arr(L.bnd_to_node, b, np.int32, (b,)),
            bnd_group=arr(L.bnd_group, b, np.int32, (b,)),
            bnd_weights=arr(L.bnd_weights, 3 * b, np.float64, (b, 3)),
            node_to_mg_node=(arr(L.node_to_mg_node, n, np.int32, (n,)) if bool(L.node_to_mg_node) else None),
        )
        out.append(lv)
    lib().mgcfd_mesh_free(p)
    return out


def mesh_box(n: int, levels: int = 4, seed: int = 1) -> list[dict]:
    """Seeded jittered n^3 box with 2x-per-axis coarsening (include/mgcfd_mesh.h)."""
    p = C.POINTER(MeshC)()
    if lib().mgcfd_mesh_box(n, levels, seed, C.byref(p)) != 0:
        raise MgcfdError(lib().mgcfd_mesh_last_error().decode())
    return _mesh_to_numpy(p)


def mesh_knn(n_nodes: int, n_edges: int, levels: int = 4, ratios=None, seed: int = 1) -> list[dict]:
    """Seeded jittered-point nearest-neighbour graph with greedy agglomeration."""
    p = C.POINTER(MeshC)()
    rp = None
    if ratios is not None:
        ra = (C.c_double * len(ratios))(*ratios)
        rp = C.cast(ra, C.c_void_p)
    if lib().mgcfd_mesh_knn(n_nodes, n_edges, levels, rp, seed, C.byref(p)) != 0:
        raise MgcfdError(lib().mgcfd_mesh_last_error().decode())
    return _mesh_to_numpy(p)


def _mesh_c(mesh: list[dict], keep: list) -> MeshC:
    """MeshC view of a list-of-levels mesh (arrays kept alive in 'keep')."""
    m = MeshC()
    m.levels = len(mesh)
    for l, lv in enumerate(mesh):
        L = m.L[l]
        a = dict(coords=np.ascontiguousarray(lv["coords"], np.float64),
                 e2n=np.ascontiguousarray(lv["edge_to_node"], np.int32),
                 ew=np.ascontiguousarray(lv["edge_weights"], np.float64),
                 b2n=np.ascontiguousarray(lv["bnd_to_node"], np.int32),
                 bg=np.ascontiguousarray(lv["bnd_group"], np.int32),
                 bw=np.ascontiguousarray(lv["bnd_weights"], np.float64),
                 f2c=(np.ascontiguousarray(lv["node_to_mg_node"], np.int32)
                      if lv.get("node_to_mg_node") is not None else None))
        keep.append(a)
        L.n_nodes, L.n_edges, L.n_bnd = len(a["coords"]), len(a["e2n"]), len(a["b2n"])
        pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.coords = a["coords"].ctypes.data_as(pd)
        L.edge_to_node = a["e2n"].ctypes.data_as(pi)
        L.edge_weights = a["ew"].ctypes.data_as(pd)
        L.bnd_to_node = a["b2n"].ctypes.data_as(pi)
        L.bnd_group = a["bg"].ctypes.data_as(pi)
        L.bnd_weights = a["bw"].ctypes.data_as(pd)
        L.node_to_mg_node = a["f2c"].ctypes.data_as(pi) if a["f2c"] is not None else None
    return m


# ---------------------------------------------------------- freestream --

def freestream(mach: float = 1.2, aoa_deg: float = 0.0) -> dict:
    """Freestream state and flux contributions (euler3d.cpp:156-189)."""
    aoa = (3.1415926535897931 / 180.0) * aoa_deg
    rho, p = 1.4, 1.0
    c = np.sqrt(1.4 * p / rho)
    speed = mach * c
    u = np.array([speed * np.cos(aoa), speed * np.sin(aoa), 0.0])
    m = rho * u
    E = rho * (0.5 * (speed * speed)) + (p / (1.4 - 1.0))
    fmx = np.array([u[0] * m[0] + p, u[0] * m[1], u[0] * m[2]])
    fmy = np.array([fmx[1], u[1] * m[1] + p, u[1] * m[2]])
    fmz = np.array([fmx[2], fmy[2], u[2] * m[2] + p])
    dep = E + p
    fde = np.array([u[0] * dep, u[1] * dep, u[2] * dep])
    return dict(smoothing_coefficient=float(np.float32(0.2)), ff_variable=np.array([rho, m[0], m[1], m[2], E]),
                ff_flux_contribution_momentum_x=fmx, ff_flux_contribution_momentum_y=fmy,
                ff_flux_contribution_momentum_z=fmz, ff_flux_contribution_density_energy=fde)


# ------------------------------------------------------------- runtime --

class Runtime:
    """One library instance: op_init on entry, op_exit on close."""

    def __init__(self, base_index: int = 0):
        L = lib()
        args = [b"mgcfd"]
        if base_index:
            args.append(f"OP_MAPS_BASE_INDEX={base_index}".encode())
        argv = (C.c_char_p * len(args))(*args)
        self._argv = argv
        _check(L.op_init(len(args), C.cast(argv, C.c_void_p), 0), "op_init")
        self.L = L
        self._keep = []
        self.open = True

    def close(self):
        if self.open:
            self.L.op_exit()
            self.open = False

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- declarations
    def decl_set(self, size: int, name: str):
        s = self.L.op_decl_set(size, name.encode())
        if not s:
            raise MgcfdError(last_error())
        return s

    def decl_map(self, frm, to, dim: int, data: np.ndarray, name: str):
        a = np.ascontiguousarray(data, dtype=np.int32)
        m = self.L.op_decl_map(frm, to, dim, a.ctypes.data, name.encode())
        if not m:
            raise MgcfdError(last_error())
        return m

    def compose_map(self, outer, inner, name: str):
        """C[e][k] = inner[outer[e][k]] (mgcfd_decl_map_compose)."""
        m = self.L.mgcfd_decl_map_compose(outer, inner, name.encode())
        if not m:
            raise MgcfdError(last_error())
        return m

    def decl_dat(self, st, dim: int, data: np.ndarray | None, name: str, typ: str = "double"):
        if data is None:
            size = 8 if typ == "double" else 4
            d = self.L.op_decl_dat_temp_char(st, dim, typ.encode(), size, name.encode())
        else:
            dt = np.float64 if typ == "double" else np.int32
            a = np.ascontiguousarray(data, dtype=dt)
            d = self.L.op_decl_dat_char(st, dim, typ.encode(), a.itemsize, a.ctypes.data, name.encode())
        if not d:
            raise MgcfdError(last_error())
        return d

    def decl_const(self, name: str, value):
        if name == "mesh_name":
            a = np.array([int(value)], dtype=np.int32)
            _check(self.L.op_decl_const_char(1, b"int", 4, a.ctypes.data, name.encode()), "op_decl_const")
            return
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(value, dtype=np.float64)))
        _check(self.L.op_decl_const_char(a.size, b"double", 8, a.ctypes.data, name.encode()), "op_decl_const")

    def decl_consts(self, fs: dict):
        for k, v in fs.items():
            self.decl_const(k, v)

    # -- args
    def arg(self, dat, idx: int = -1, mp=None, dim: int = 1, typ: str = "double", acc: int = OP_READ) -> OpArg:
        return self.L.op_arg_dat(dat, idx, mp, dim, typ.encode(), acc)

    def gbl(self, arr: np.ndarray, acc: int) -> OpArg:
        typ = b"double" if arr.dtype == np.float64 else b"int"
        self._keep.append(arr)
        return self.L.op_arg_gbl_char(arr.ctypes.data, arr.size, typ, arr.itemsize, acc)

    def par_loop(self, kernel: str, st, *args):
        fn = getattr(self.L, "op_par_loop_" + kernel)
        if len(args) != LOOPS[kernel]:
            raise MgcfdError(f"{kernel}: expected {LOOPS[kernel]} args, got {len(args)}")
        rc = fn(kernel.encode(), st, *args)
        self._keep.clear()
        if rc != 0:
            raise MgcfdError(f"op_par_loop {kernel}: {last_error()}")

    # -- data
    def fetch(self, dat) -> np.ndarray:
        d = dat.contents
        n = d.set.contents.size
        typ = d.type.decode()
        dt = np.float64 if typ == "double" else np.int32
        out = np.empty((n, d.dim), dtype=dt)
        _check(self.L.op_fetch_data_char(dat, out.ctypes.data), "op_fetch_data")
        return out

    def upload(self, dat, data: np.ndarray):
        d = dat.contents
        dt = np.float64 if d.type.decode() == "double" else np.int32
        a = np.ascontiguousarray(data, dtype=dt)
        if a.size != d.set.contents.size * d.dim:
            raise MgcfdError("upload: size mismatch")
        _check(self.L.op_upload_data_char(dat, a.ctypes.data), "op_upload_data")

    def partition(self, nodes=None, coords=None):
        _check(self.L.op_partition(b"PARMETIS", b"KWAY", nodes, None, coords), "op_partition")

    def map_part_config(self, m, part_nodes: int, part_threads: int = 0, record_bytes: int = 0):
        """Part plan configuration of an edge map (include/mgcfd_op2.h, mgcfd_map_part_config)."""
        _check(self.L.mgcfd_map_part_config(m, part_nodes, part_threads, record_bytes), "mgcfd_map_part_config")

    def set_flux_variant(self, v: int):
        _check(self.L.mgcfd_set_flux_variant(v), "mgcfd_set_flux_variant")

    def sync(self):
        _check(self.L.mgcfd_sync(), "mgcfd_sync")

    def device_info(self):
        buf = C.create_string_buffer(256)
        ncu = C.c_int(0)
        _check(self.L.mgcfd_device_info(buf, 256, C.byref(ncu)), "mgcfd_device_info")
        return buf.value.decode(), ncu.value


@dataclass
class Batch:
    """n independent multigrid solves concatenated behind offset tables and
    advanced together (include/mgcfd_op2.h, mgcfd_batch_*)."""

    def __init__(self, rt: "Runtime", meshes: list[list[dict]], mach, aoa_deg, variant: str = "gather_exact"):
        self.rt, self.L = rt, rt.L
        self.n = len(meshes)
        self.sizes = [[len(lv["coords"]) for lv in m] for m in meshes]
        keep = []
        arr = (MeshC * self.n)(*[_mesh_c(m, keep) for m in meshes])
        ma = (C.c_double * self.n)(*[float(x) for x in np.broadcast_to(mach, (self.n,))])
        aa = (C.c_double * self.n)(*[float(x) for x in np.broadcast_to(aoa_deg, (self.n,))])
        self.h = C.c_void_p()
        _check(self.L.mgcfd_batch_create(self.n, arr, ma, aa, FLUX_VARIANTS[variant], C.byref(self.h)),
               "mgcfd_batch_create")

    def autotune(self, iters: int = 5) -> dict:
        """Per level, the nodes per part, threads per part and persistent
        workgroups per CU (0: one workgroup per part) the autotuner chose."""
        L = len(self.sizes[0])
        nodes, threads, persist = (C.c_int * L)(), (C.c_int * L)(), (C.c_int * L)()
        _check(self.L.mgcfd_batch_autotune_levels(self.h, iters, nodes, threads, persist),
               "mgcfd_batch_autotune_levels")
        return dict(part_nodes=list(nodes), part_threads=list(threads), persistent_per_cu=list(persist))

    def set_part_config(self, level: int, part_nodes: int, part_threads: int, persistent_per_cu: int = 0):
        _check(self.L.mgcfd_batch_set_part_config(self.h, level, part_nodes, part_threads, persistent_per_cu),
               "mgcfd_batch_set_part_config")

    def run(self, cycles: int, use_graph: bool = True):
        rms = np.zeros((self.n, max(cycles, 1)))
        st = np.zeros(self.n, np.int32)
        _check(self.L.mgcfd_batch_run(self.h, cycles, int(use_graph), rms.ctypes.data_as(C.c_void_p),
                                      st.ctypes.data_as(C.c_void_p)), "mgcfd_batch_run")
        return st, rms[:, :cycles]

    def fetch(self, mesh: int, level: int, what: str = "variables") -> np.ndarray:
        dim = 5 if what in ("variables", "fluxes") else 1
        out = np.zeros((self.sizes[mesh][level], dim))
        _check(self.L.mgcfd_batch_fetch(self.h, mesh, level, what.encode(), out.ctypes.data_as(C.c_void_p)),
               "mgcfd_batch_fetch")
        return out

    def close(self):
        if self.h:
            _check(self.L.mgcfd_batch_destroy(self.h), "mgcfd_batch_destroy")
            self.h = C.c_void_p()


class Server:
    """Continuous-batching request loop (include/mgcfd_op2.h, mgcfd_server_*)."""

    def __init__(self, rt: "Runtime", budget_bytes: float = 200e9, max_active: int = 8, chunk: int = 10,
                 max_group: int = 8, max_group_nodes: int = 4_000_000):
        self.rt, self.L = rt, rt.L
        self.h = C.c_void_p()
        self.n0 = {}
        _check(self.L.mgcfd_server_create(C.c_double(budget_bytes), max_active, chunk, C.byref(self.h)),
               "mgcfd_server_create")
        _check(self.L.mgcfd_server_set_grouping(self.h, max_group, C.c_longlong(max_group_nodes)),
               "mgcfd_server_set_grouping")

    def submit(self, mesh: list[dict], mach: float, aoa_deg: float, variant: str = "gather_exact",
               max_cycles: int = 200, rms_drop: float = 0.0) -> int:
        keep = []
        m = _mesh_c(mesh, keep)
        rid = C.c_int(0)
        _check(self.L.mgcfd_server_submit(self.h, C.byref(m), C.c_double(mach), C.c_double(aoa_deg),
                                          FLUX_VARIANTS[variant], max_cycles, C.c_double(rms_drop), C.byref(rid)),
               "mgcfd_server_submit")
        self.n0[rid.value] = (len(mesh[0]["coords"]), max_cycles)
        return rid.value

    def progress(self):
        a, q = C.c_int(0), C.c_int(0)
        _check(self.L.mgcfd_server_progress(self.h, C.byref(a), C.byref(q)), "mgcfd_server_progress")
        return a.value, q.value

    def result(self, rid: int):
        """None while queued or running, else dict(status, cycles, latency, variables, rms)."""
        n0, cap = self.n0[rid]
        state, st, cyc, lat = C.c_int(0), C.c_int(0), C.c_int(0), C.c_double(0)
        var = np.zeros((n0, 5))
        rms = np.zeros(cap)
        _check(self.L.mgcfd_server_result(self.h, rid, C.byref(state), C.byref(st), C.byref(cyc), C.byref(lat),
                                          var.ctypes.data_as(C.c_void_p), rms.ctypes.data_as(C.c_void_p)),
               "mgcfd_server_result")
        if state.value != 2:
            return None
        return dict(status=st.value, cycles=cyc.value, latency=lat.value, variables=var, rms=rms[:cyc.value])

    def timing(self, rid: int):
        """(device start, device end) of a finished request, ms since the server was created."""
        a, b = C.c_double(0), C.c_double(0)
        _check(self.L.mgcfd_server_timing(self.h, rid, C.byref(a), C.byref(b)), "mgcfd_server_timing")
        return a.value, b.value

    @staticmethod
    def estimate_bytes(mesh: list[dict]) -> float:
        """The device-memory estimate admission charges for this mesh."""
        keep = []
        m = _mesh_c(mesh, keep)
        b = C.c_double(0)
        _check(lib().mgcfd_server_estimate_bytes(C.byref(m), C.byref(b)), "mgcfd_server_estimate_bytes")
        return b.value

    def close(self):
        if self.h:
            _check(self.L.mgcfd_server_destroy(self.h), "mgcfd_server_destroy")
            self.h = C.c_void_p()


class LevelHandles:
    """Sets, maps and dats of one multigrid level, named as euler3d.cpp names them."""
    nodes: object = None
    edges: object = None
    bnd_nodes: object = None
    edge_to_node: object = None
    bnd_to_node: object = None
    node_to_mg_node: object = None
    edge_to_mg_nodes: object = None
    prolonged: object = None
    prolonged_wsum: object = None
    coords: object = None
    edge_weights: object = None
    bnd_weights: object = None
    bnd_groups: object = None
    variables: object = None
    old_variables: object = None
    residuals: object = None
    volumes: object = None
    step_factors: object = None
    fluxes: object = None
    dummy_fluxes: object = None
    up_scratch: object = None
    sizes: dict = field(default_factory=dict)


class Solver:
    """Declares a multigrid mesh as euler3d.cpp:243-409 does and runs its loops.

    ``mesh`` is a list of per-level dicts (see mesh_box / mesh_knn).
    ``down_v2`` declares the edge-->mg_nodes maps (euler3d.cpp:286-292) and
    prolongs with the inverse-distance loops (euler3d.cpp:603-623).
    """

    def __init__(self, rt: Runtime, mesh: list[dict], mach: float = 1.2, aoa_deg: float = 0.0,
                 measure_mem_bound: bool = False, down_v2: bool = False):
        self.rt = rt
        self.down_v2 = down_v2
        self.nlev = len(mesh)
        self.fs = freestream(mach, aoa_deg)
        rt.decl_consts(self.fs)
        rt.decl_const("mesh_name", 3)
        self.lv = []
        for i, m in enumerate(mesh):
            h = LevelHandles()
            n, e, b = len(m["coords"]), len(m["edge_to_node"]), len(m["bnd_to_node"])
            h.sizes = dict(nodes=n, edges=e, bnd=b)
            h.nodes = rt.decl_set(n, f"op_nodes_L{i}")
            h.edges = rt.decl_set(e, f"op_edges_L{i}")
            h.bnd_nodes = rt.decl_set(b, f"op_bnd_nodes_L{i}")
            h.edge_to_node = rt.decl_map(h.edges, h.nodes, 2, m["edge_to_node"], "edge-->node")
            h.bnd_to_node = rt.decl_map(h.bnd_nodes, h.nodes, 1, m["bnd_to_node"], "bnd_node-->node")
            h.bnd_groups = rt.decl_dat(h.bnd_nodes, 1, m["bnd_group"], "bnd_node-->group", "int")
            h.edge_weights = rt.decl_dat(h.edges, 3, m["edge_weights"], "edge_weights")
            h.bnd_weights = rt.decl_dat(h.bnd_nodes, 3, m["bnd_weights"], "bnd_node_weights")
            h.coords = rt.decl_dat(h.nodes, 3, m["coords"], "node_coordinates")
            self.lv.append(h)
        for h in self.lv:       # geometric node order (euler3d.cpp:340-377 passes the coordinates)
            rt.partition(h.nodes, h.coords)
        for i in range(1, self.nlev):
            self.lv[i - 1].node_to_mg_node = rt.decl_map(self.lv[i - 1].nodes, self.lv[i].nodes, 1,
                                                         mesh[i - 1]["node_to_mg_node"], f"op_node-->mg_node_L{i}")
            if down_v2:
                f = self.lv[i - 1]
                f.edge_to_mg_nodes = rt.compose_map(f.edge_to_node, f.node_to_mg_node, f"op_edge-->mg_nodes_L{i}")
                f.prolonged = rt.decl_dat(f.nodes, 5, None, f"p_residuals_prolonged_L{i - 1}")
                f.prolonged_wsum = rt.decl_dat(f.nodes, 1, None, f"p_residuals_prolonged_wsum_L{i - 1}")
        for i, h in enumerate(self.lv):
            h.variables = rt.decl_dat(h.nodes, 5, None, f"p_variables_L{i}")
            h.old_variables = rt.decl_dat(h.nodes, 5, None, f"p_old_variables_L{i}")
            h.residuals = rt.decl_dat(h.nodes, 5, None, f"p_residuals_L{i}")
            h.volumes = rt.decl_dat(h.nodes, 1, None, f"p_volumes_L{i}")
            h.step_factors = rt.decl_dat(h.nodes, 1, None, f"p_step_factors_L{i}")
            h.fluxes = rt.decl_dat(h.nodes, 5, None, f"p_fluxes_L{i}")
            if measure_mem_bound:
                h.dummy_fluxes = rt.decl_dat(h.nodes, 5, None, f"p_dummy_fluxes_L{i}")
            if i > 0:
                h.up_scratch = rt.decl_dat(h.nodes, 1, None, f"p_up_scratch_L{i}", "int")

    # ---- loop sites of euler3d.cpp (one method per op_par_loop)
    def setup(self):
        """euler3d.cpp:412-441: initialise variables, zero fluxes, volumes, damp weights."""
        rt = self.rt
        for h in self.lv:
            rt.par_loop("initialize_variables_kernel", h.nodes, rt.arg(h.variables, -1, None, 5, acc=OP_WRITE))
            rt.par_loop("zero_5d_array_kernel", h.nodes, rt.arg(h.fluxes, -1, None, 5, acc=OP_WRITE))
            if h.dummy_fluxes:
                rt.par_loop("zero_5d_array_kernel", h.nodes, rt.arg(h.dummy_fluxes, -1, None, 5, acc=OP_WRITE))
            rt.par_loop("zero_1d_array_kernel", h.nodes, rt.arg(h.volumes, -1, None, 1, acc=OP_WRITE))
            rt.par_loop("calculate_cell_volumes", h.edges,
                        rt.arg(h.coords, 0, h.edge_to_node, 3, acc=OP_READ),
                        rt.arg(h.coords, 1, h.edge_to_node, 3, acc=OP_READ),
                        rt.arg(h.edge_weights, -1, None, 3, acc=OP_RW),
                        rt.arg(h.volumes, 0, h.edge_to_node, 1, acc=OP_INC),
                        rt.arg(h.volumes, 1, h.edge_to_node, 1, acc=OP_INC))
        for h in self.lv:
            rt.par_loop("dampen_ewt", h.edges, rt.arg(h.edge_weights, -1, None, 3, acc=OP_RW))
            rt.par_loop("dampen_ewt", h.bnd_nodes, rt.arg(h.bnd_weights, -1, None, 3, acc=OP_RW))

    def flux_edge(self, l: int, fluxes=None):
        rt, h = self.rt, self.lv[l]
        f = fluxes if fluxes is not None else h.fluxes
        rt.par_loop("compute_flux_edge_kernel", h.edges,
                    rt.arg(h.variables, 0, h.edge_to_node, 5, acc=OP_READ),
                    rt.arg(h.variables, 1, h.edge_to_node, 5, acc=OP_READ),
                    rt.arg(h.edge_weights, -1, None, 3, acc=OP_READ),
                    rt.arg(f, 0, h.edge_to_node, 5, acc=OP_INC),
                    rt.arg(f, 1, h.edge_to_node, 5, acc=OP_INC))

    def flux_bnd(self, l: int):
        rt, h = self.rt, self.lv[l]
        rt.par_loop("compute_bnd_node_flux_kernel", h.bnd_nodes,
                    rt.arg(h.bnd_groups, -1, None, 1, "int", OP_READ),
                    rt.arg(h.bnd_weights, -1, None, 3, acc=OP_READ),
                    rt.arg(h.variables, 0, h.bnd_to_node, 5, acc=OP_READ),
                    rt.arg(h.fluxes, 0, h.bnd_to_node, 5, acc=OP_INC))

    def stream(self, l: int):
        rt, h = self.rt, self.lv[l]
        rt.par_loop("unstructured_stream_kernel", h.edges,
                    rt.arg(h.variables, 0, h.edge_to_node, 5, acc=OP_READ),
                    rt.arg(h.variables, 1, h.edge_to_node, 5, acc=OP_READ),
                    rt.arg(h.edge_weights, -1, None, 3, acc=OP_READ),
                    rt.arg(h.dummy_fluxes, 0, h.edge_to_node, 5, acc=OP_INC),
                    rt.arg(h.dummy_fluxes, 1, h.edge_to_node, 5, acc=OP_INC))

    def smooth(self, l: int) -> tuple[float, float | None, int]:
        """One smoothing step, euler3d.cpp:467-550, loop by loop (host syncs at
        the global reductions as in the reference).  Returns (min_dt, rms, bad)."""
        rt, h = self.rt, self.lv[l]
        rt.par_loop("copy_double_kernel", h.nodes, rt.arg(h.variables, -1, None, 5, acc=OP_READ),
                    rt.arg(h.old_variables, -1, None, 5, acc=OP_WRITE))
        rt.par_loop("calculate_dt_kernel", h.nodes, rt.arg(h.variables, -1, None, 5, acc=OP_READ),
                    rt.arg(h.volumes, -1, None, 1, acc=OP_READ), rt.arg(h.step_factors, -1, None, 1, acc=OP_WRITE))
        min_dt = np.array([np.finfo(np.float64).max])
        rt.par_loop("get_min_dt_kernel", h.nodes, rt.arg(h.step_factors, -1, None, 1, acc=OP_READ),
                    rt.gbl(min_dt, OP_MIN))
        if min_dt[0] < 0.0:
            return float(min_dt[0]), None, -1
        rt.par_loop("compute_step_factor_kernel", h.nodes, rt.arg(h.variables, -1, None, 5, acc=OP_READ),
                    rt.arg(h.volumes, -1, None, 1, acc=OP_READ), rt.gbl(min_dt, OP_READ),
                    rt.arg(h.step_factors, -1, None, 1, acc=OP_WRITE))
        for rk in range(3):
            self.flux_edge(l)
            self.flux_bnd(l)
            rkc = np.array([rk], dtype=np.int32)
            rt.par_loop("time_step_kernel", h.nodes, rt.gbl(rkc, OP_READ),
                        rt.arg(h.step_factors, -1, None, 1, acc=OP_READ),
                        rt.arg(h.fluxes, -1, None, 5, acc=OP_RW),
                        rt.arg(h.old_variables, -1, None, 5, acc=OP_READ),
                        rt.arg(h.variables, -1, None, 5, acc=OP_WRITE))
            if h.dummy_fluxes:
                self.stream(l)
        rt.par_loop("residual_kernel", h.nodes, rt.arg(h.old_variables, -1, None, 5, acc=OP_READ),
                    rt.arg(h.variables, -1, None, 5, acc=OP_READ), rt.arg(h.residuals, -1, None, 5, acc=OP_WRITE))
        rms = None
        bad = 0
        if l == 0:
            r = np.zeros(1)
            rt.par_loop("calc_rms_kernel", h.nodes, rt.arg(h.residuals, -1, None, 5, acc=OP_READ), rt.gbl(r, OP_INC))
            rms = float(np.sqrt(r[0] / h.sizes["nodes"]))
            cnt = np.zeros(1, dtype=np.int32)
            rt.par_loop("count_bad_vals", h.nodes, rt.arg(h.variables, -1, None, 5, acc=OP_READ), rt.gbl(cnt, OP_INC))
            bad = int(cnt[0])
        return float(min_dt[0]), rms, bad

    def restrict(self, l: int):
        """MG_UP to level l (euler3d.cpp:579-592)."""
        rt, f, c = self.rt, self.lv[l - 1], self.lv[l]
        m = f.node_to_mg_node
        rt.par_loop("up_pre_kernel", f.nodes, rt.arg(c.variables, 0, m, 5, acc=OP_WRITE),
                    rt.arg(c.up_scratch, 0, m, 1, "int", OP_WRITE))
        rt.par_loop("up_kernel", f.nodes, rt.arg(f.variables, -1, None, 5, acc=OP_READ),
                    rt.arg(c.variables, 0, m, 5, acc=OP_INC), rt.arg(c.up_scratch, 0, m, 1, "int", OP_INC))
        rt.par_loop("up_post_kernel", c.nodes, rt.arg(c.variables, -1, None, 5, acc=OP_RW),
                    rt.arg(c.up_scratch, -1, None, 1, "int", OP_READ))

    def prolong(self, l: int):
        """MG_DOWN to level l (euler3d.cpp:601-631): the inverse-distance loops
        when the edge-->mg_nodes map is declared, else down_kernel."""
        rt, f, c = self.rt, self.lv[l], self.lv[l + 1]
        if f.edge_to_mg_nodes:
            en, em, R, I = f.edge_to_node, f.edge_to_mg_nodes, OP_READ, OP_INC
            rt.par_loop("down_v2_kernel_pre", f.nodes, rt.arg(f.prolonged_wsum, -1, None, 1, acc=OP_WRITE),
                        rt.arg(f.prolonged, -1, None, 5, acc=OP_WRITE))
            rt.par_loop("down_v2_kernel", f.edges, rt.arg(f.coords, 0, en, 3, acc=R), rt.arg(f.coords, 1, en, 3, acc=R),
                        rt.arg(c.coords, 0, em, 3, acc=R), rt.arg(c.coords, 1, em, 3, acc=R),
                        rt.arg(c.residuals, 0, em, 5, acc=R), rt.arg(c.residuals, 1, em, 5, acc=R),
                        rt.arg(f.prolonged, 0, en, 5, acc=I), rt.arg(f.prolonged, 1, en, 5, acc=I),
                        rt.arg(f.prolonged_wsum, 0, en, 1, acc=I), rt.arg(f.prolonged_wsum, 1, en, 1, acc=I))
            rt.par_loop("down_v2_kernel_post", f.nodes, rt.arg(f.prolonged, -1, None, 5, acc=R),
                        rt.arg(f.prolonged_wsum, -1, None, 1, acc=R), rt.arg(f.residuals, -1, None, 5, acc=R),
                        rt.arg(f.variables, -1, None, 5, acc=I))
            return
        m = f.node_to_mg_node
        rt.par_loop("down_kernel", f.nodes, rt.arg(f.variables, -1, None, 5, acc=OP_INC),
                    rt.arg(f.residuals, -1, None, 5, acc=OP_READ), rt.arg(f.coords, -1, None, 3, acc=OP_READ),
                    rt.arg(c.residuals, 0, m, 5, acc=OP_READ), rt.arg(c.coords, 0, m, 3, acc=OP_READ))

    def run_loops(self, cycles: int):
        """The multigrid loop of euler3d.cpp:452-641 driven loop by loop.
        Returns (status, rms_history, min_dt_history)."""
        level, up, i = 0, True, 0
        rms_hist, mdt = [], []
        while i < cycles:
            min_dt, rms, bad = self.smooth(level)
            mdt.append(min_dt)
            if bad < 0:
                return 1, rms_hist, mdt
            if level == 0:
                rms_hist.append(rms)
                if bad > 0:
                    return 2, rms_hist, mdt
            if self.nlev <= 1:
                i += 1
            elif up:
                level += 1
                self.restrict(level)
                if level == self.nlev - 1:
                    up = False
            else:
                level -= 1
                self.prolong(level)
                if level == 0:
                    up, i = True, i + 1
        return 0, rms_hist, mdt

    def bind(self):
        if self.down_v2:
            raise MgcfdError("the device-resident multigrid prolongs with down_kernel; "
                             "a down_v2 solver runs loop by loop (run_loops)")
        binds = (LevelBind * self.nlev)()
        for i, h in enumerate(self.lv):
            b = binds[i]
            b.nodes, b.edges, b.bnd_nodes = h.nodes, h.edges, h.bnd_nodes
            b.edge_to_node, b.bnd_to_node = h.edge_to_node, h.bnd_to_node
            if h.node_to_mg_node:
                b.node_to_mg_node = h.node_to_mg_node
            b.variables, b.old_variables, b.residuals = h.variables, h.old_variables, h.residuals
            b.fluxes, b.step_factors, b.volumes = h.fluxes, h.step_factors, h.volumes
            b.coords, b.edge_weights, b.bnd_weights, b.bnd_groups = h.coords, h.edge_weights, h.bnd_weights, h.bnd_groups
            if h.up_scratch:
                b.up_scratch = h.up_scratch
        _check(self.rt.L.mgcfd_mg_bind(self.nlev, binds), "mgcfd_mg_bind")

    def run_device(self, cycles: int, use_graph: bool = True):
        """Device-resident multigrid (mgcfd_mg_run).  Returns (status, rms_history)."""
        self.bind()
        rms = np.zeros(max(cycles, 1))
        st = C.c_int(0)
        _check(self.rt.L.mgcfd_mg_run(cycles, 1 if use_graph else 0,
                                      rms.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st)), "mgcfd_mg_run")
        return st.value, rms[:cycles]

    def state(self, l: int, what: str = "variables") -> np.ndarray:
        return self.rt.fetch(getattr(self.lv[l], what))
