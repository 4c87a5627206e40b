"""Python bindings of libmosp_hip.so (the MI355X DynaMOSP update engine).

Mirrors the reference's in-memory pipeline interface (src/mospUpdate.cu,
src/sospUpdateGpu.cu, src/combinedGraphGpu.cu) over the C ABI declared in
include/mosp_hip.h.  Device buffers come from the library itself
(mosp_dev_alloc / mosp_memcpy), so no other GPU framework is needed; every
computation runs in the HIP kernels of the library, and the module raises if
the library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import pathlib
import weakref
from dataclasses import dataclass, field

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
# MOSP_LIB_PATH selects another build of the same library (e.g. a variant
# compiled with a different workgroup size); default: the in-tree build.
LIB_PATH = pathlib.Path(os.environ.get("MOSP_LIB_PATH", HERE.parent / "lib" / "libmosp_hip.so"))
DISTANCE_INF = (2**63 - 1) // 4

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I, _L, _VP = C.c_int32, C.c_int64, C.c_void_p


class Csr(C.Structure):
    """mosp_csr (include/mosp_hip.h) -- DeviceCsr of the reference."""
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("row_ptr", _VP),
                ("col_ind", _VP), ("weights", _VP)]


class Changes(C.Structure):
    """mosp_changes -- DeviceChanges of the reference."""
    _fields_ = [("changed_from", _VP), ("changed_to", _VP), ("n_changed", C.c_int32),
                ("insert_heads", _VP), ("n_insert_heads", C.c_int32)]


class SospStats(C.Structure):
    """mosp_sosp_stats -- SospStats of the reference (+ kernel_ms)."""
    _fields_ = [("invalidated", C.c_int32), ("jump_rounds", C.c_int32),
                ("iterations", C.c_int32), ("epochs", C.c_int32), ("pushes", C.c_int64),
                ("packed_parents", C.c_int32), ("kernel_ms", C.c_float)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class CombineStats(C.Structure):
    """mosp_combine_stats -- CombineStats of the reference (+ phase split)."""
    _fields_ = [("n_edges", C.c_int32), ("scale", C.c_int64), ("delta", C.c_int64),
                ("search", SospStats), ("sosp_phase_ms", C.c_float),
                ("combined_phase_ms", C.c_float)]


class BatchInstance(C.Structure):
    """mosp_batch_instance -- one update instance of a batched launch."""
    _fields_ = [("graph", C.c_void_p), ("k", C.c_int32), ("changes", C.POINTER(Changes)),
                ("source", C.c_int32), ("delta", C.POINTER(C.c_int64)),
                ("max_weight", C.POINTER(C.c_int64)), ("pref", C.POINTER(C.c_int32)),
                ("n_pref", C.c_int32), ("d_dist_in", C.c_void_p), ("d_parents_in", C.c_void_p),
                ("d_dist", C.c_void_p), ("d_parents", C.c_void_p), ("d_comb_dist", C.c_void_p),
                ("d_comb_parent", C.c_void_p)]


EXPORTS = (
    "mosp_last_error", "mosp_ws_create", "mosp_ws_reserve", "mosp_ws_destroy",
    "mosp_ws_set_grid", "mosp_ws_grid", "mosp_ws_set_combined_delta_scale", "mosp_ws_last_profile", "mosp_ws_async_events", "mosp_default_delta", "mosp_sosp_update",
    "mosp_sosp_update_multi", "mosp_sosp_from_scratch", "mosp_combined_sosp",
    "mosp_update_fused", "mosp_device_graph_upload", "mosp_device_graph_free",
    "mosp_device_graph_view", "mosp_device_graph_objectives", "mosp_path_costs_gpu",
    "mosp_preference_scale", "mosp_combined_edge_weight", "mosp_choose_packing",
    "mosp_apply_batch", "mosp_update_host", "mosp_gen_road_grid", "mosp_gen_rmat",
    "mosp_gen_batch", "mosp_set_device", "mosp_device_count", "mosp_dev_alloc",
    "mosp_dev_free", "mosp_memcpy", "mosp_device_synchronize", "mosp_batch_create",
    "mosp_batch_destroy", "mosp_batch_configure", "mosp_batch_run", "mosp_batch_check",
    "mosp_batch_stats", "mosp_batch_schedule", "mosp_batch_set_cost_model",
    "mosp_device_graph_create", "mosp_device_graph_apply", "mosp_device_graph_edges",
    "mosp_device_graph_download", "mosp_device_graph_weight_stats", "mosp_batch_create_capacity",
    "mosp_batch_reset", "mosp_batch_instance_bytes", "mosp_batch_stream", "mosp_device_cus",
    "mosp_server_create", "mosp_server_add_graph", "mosp_server_start", "mosp_server_submit",
    "mosp_server_wait", "mosp_server_drain", "mosp_server_request", "mosp_server_request_trees",
    "mosp_server_release", "mosp_server_stats", "mosp_server_reset_stats", "mosp_server_clock",
    "mosp_server_destroy", "mosp_parallel_sosp_update", "mosp_parallel_combined_graph",
    "mosp_canonicalize_tree", "mosp_path_costs_host", "mosp_ws_set_barrier_timeout",
    "mosp_batch_set_barrier_timeout", "mosp_sosp_update_noalloc", "mosp_ws_set_shape",
    "mosp_batch_set_shape",
)

_LIB: C.CDLL | None = None

# Every live object that owns library or device resources, released at
# interpreter exit in dependency order (servers and their worker threads
# first, plain device buffers last) while the HIP runtime is still intact,
# instead of by finalizers that may run after it has been torn down.
_LIVE: "weakref.WeakSet" = weakref.WeakSet()
_RELEASE_ORDER = ("Server", "BatchedUpdate", "Session", "Workspace", "DeviceGraph",
                  "DeviceBuffer")


def _track(obj) -> None:
    _LIVE.add(obj)


@atexit.register
def _release_all() -> None:
    if _LIB is None:
        return
    rank = {name: i for i, name in enumerate(_RELEASE_ORDER)}
    for obj in sorted(list(_LIVE), key=lambda o: rank.get(type(o).__name__, len(rank))):
        try:
            (getattr(obj, "close", None) or getattr(obj, "free"))()
        except Exception:
            pass
    try:
        _LIB.mosp_device_synchronize()
    except Exception:
        pass


def lib() -> C.CDLL:
    """Load libmosp_hip.so (built by `make -C mosp-cuda_amd`); raise if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C mosp-cuda_amd` "
                           "(python -c 'import __graft_entry__; __graft_entry__.build()')")
    # A process must hold ONE HIP runtime.  If torch was imported before, its
    # bundled runtime (soname libamdhip64.so.7) is already loaded and the
    # library binds to it; otherwise the library loads the system runtime,
    # and importing torch afterwards would bring in a second copy.
    L = C.CDLL(str(LIB_PATH))
    L.mosp_set_device.argtypes = [_I]
    L.mosp_device_count.restype = _I
    L.mosp_device_cus.restype = _I
    L.mosp_dev_alloc.restype = _VP
    L.mosp_dev_alloc.argtypes = [C.c_uint64]
    L.mosp_dev_free.argtypes = [_VP]
    L.mosp_memcpy.argtypes = [_VP, _VP, C.c_uint64, _I]
    L.mosp_last_error.restype = C.c_char_p
    L.mosp_ws_create.restype = _VP
    L.mosp_ws_reserve.argtypes = [_VP, _I, _I]
    L.mosp_ws_destroy.argtypes = [_VP]
    L.mosp_ws_set_grid.argtypes = [_VP, _I]
    L.mosp_ws_set_combined_delta_scale.argtypes = [_VP, _I]
    L.mosp_ws_grid.argtypes = [_VP]
    L.mosp_ws_grid.restype = _I
    L.mosp_ws_last_profile.argtypes = [_VP, _f64p, _I, C.POINTER(C.c_int32)]
    L.mosp_ws_async_events.argtypes = [_VP, C.POINTER(C.c_int32)]
    L.mosp_default_delta.argtypes = [_L, _I, _L]
    L.mosp_default_delta.restype = _L
    L.mosp_sosp_update.argtypes = [C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Changes), _I, _L,
                                   _L, _VP, _VP, _VP, C.POINTER(SospStats), _VP]
    L.mosp_sosp_update_multi.argtypes = [_I, C.POINTER(Csr), C.POINTER(Csr),
                                         C.POINTER(Changes), _I, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), _VP, C.POINTER(_VP),
                                         C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP),
                                         C.POINTER(SospStats), _VP]
    L.mosp_sosp_from_scratch.argtypes = [C.POINTER(Csr), _I, _L, _L, _VP, _VP, _VP,
                                         C.POINTER(SospStats), _VP]
    L.mosp_combined_sosp.argtypes = [_VP, _I, _I, _i32p, _I, _I, _L, _VP, _VP, _VP,
                                     C.POINTER(CombineStats), _VP]
    L.mosp_update_fused.argtypes = [_VP, _I, C.POINTER(Changes), _I, C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), _i32p, _I, _VP, _VP, _VP, _VP, _VP,
                                    _VP, _VP, C.POINTER(SospStats), C.POINTER(CombineStats),
                                    _VP]
    L.mosp_device_graph_upload.restype = _VP
    L.mosp_device_graph_upload.argtypes = [_I, _I, _i32p, _i32p, _i32p]
    L.mosp_device_graph_free.argtypes = [_VP]
    L.mosp_device_graph_create.restype = _VP
    L.mosp_device_graph_apply.argtypes = [_VP, _I, _i32p, _i32p, _i32p, _I, _i32p, _i32p, _VP,
                                          _u32p, _VP]
    L.mosp_device_graph_edges.argtypes = [_VP]
    L.mosp_device_graph_edges.restype = _L
    L.mosp_device_graph_download.argtypes = [_VP, _i32p, _i32p, _i32p]
    L.mosp_device_graph_weight_stats.argtypes = [_VP, _i64p, _i64p, _VP]
    L.mosp_device_graph_view.argtypes = [_VP, _I, C.POINTER(Csr), C.POINTER(Csr)]
    L.mosp_device_graph_objectives.argtypes = [_VP]
    L.mosp_device_graph_objectives.restype = _I
    L.mosp_path_costs_gpu.argtypes = [_VP, _VP, _I, _VP, _VP]
    L.mosp_preference_scale.argtypes = [_i32p, _I, _I]
    L.mosp_preference_scale.restype = _L
    L.mosp_combined_edge_weight.argtypes = [C.c_uint32, _i32p, _I, _I, _L]
    L.mosp_combined_edge_weight.restype = _L
    L.mosp_choose_packing.argtypes = [_I, _L, C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
    L.mosp_apply_batch.restype = _L
    L.mosp_apply_batch.argtypes = [_I, _I, _i32p, _i32p, _i32p, _I, _i32p, _i32p, _i32p, _I,
                                   _i32p, _i32p, _i32p, _i32p, _i32p, _u32p]
    L.mosp_sosp_update_noalloc.argtypes = [C.POINTER(Csr), C.POINTER(Csr), C.POINTER(Changes),
                                           _I, _L, _L, _VP, _VP, C.POINTER(SospStats), _VP]
    L.mosp_ws_set_barrier_timeout.argtypes = [_VP, C.c_double]
    L.mosp_ws_set_shape.argtypes = [_VP, _I, _I]
    L.mosp_batch_set_shape.argtypes = [_VP, _I, _I]
    L.mosp_batch_set_barrier_timeout.argtypes = [_VP, C.c_double]
    L.mosp_canonicalize_tree.restype = None
    L.mosp_canonicalize_tree.argtypes = [_I, _I, _i32p, _i32p, _i32p, _I, _I, _i64p, _i32p]
    L.mosp_path_costs_host.argtypes = [_I, _I, _i32p, _i32p, _i32p, _i32p, _I, _i64p]
    L.mosp_parallel_sosp_update.argtypes = [C.c_char_p] * 5 + [_I, _I, C.c_char_p, C.c_char_p]
    L.mosp_parallel_combined_graph.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), _I, _I,
                                               C.c_char_p, C.c_char_p, _i32p, _I]
    L.mosp_update_host.argtypes = [_I, _I, _i32p, _i32p, _i32p, _I, _i32p, _i32p, _i32p, _I,
                                   _i32p, _i32p, _I, _I, _i32p, _I, _L, _I, _i64p, _i32p,
                                   C.POINTER(C.c_int64), _i32p, _i32p, _i32p, _i64p, _i32p,
                                   _i64p, _i32p, C.POINTER(SospStats),
                                   C.POINTER(CombineStats), _f64p]
    gen_tail = [C.POINTER(C.c_int32), C.POINTER(C.c_int64), _VP, _VP, _VP, _L]
    L.mosp_gen_road_grid.argtypes = [_I, _I, _I, _I, _I, C.c_double, C.c_double, C.c_uint64,
                                     *gen_tail]
    L.mosp_gen_rmat.argtypes = [_I, _I, C.c_double, C.c_double, C.c_double, _I, _I, _I,
                                C.c_uint64, *gen_tail]
    L.mosp_gen_batch.argtypes = [_I, _I, _i32p, _i32p, _i32p, _I, C.c_double, _I, _I, _I,
                                 C.c_uint64, _I, C.POINTER(C.c_int32), _i32p, _i32p, _i32p,
                                 C.POINTER(C.c_int32), _i32p, _i32p]
    L.mosp_batch_create.restype = _VP
    L.mosp_batch_create.argtypes = [C.POINTER(BatchInstance), _I]
    L.mosp_batch_destroy.argtypes = [_VP]
    L.mosp_batch_configure.argtypes = [_VP, _I, _I, C.c_double, _I]
    L.mosp_batch_run.argtypes = [_VP, _I, C.POINTER(C.c_float)]
    L.mosp_batch_check.argtypes = [_VP]
    L.mosp_batch_stats.argtypes = [_VP, _I, C.POINTER(SospStats), C.POINTER(CombineStats)]
    L.mosp_batch_schedule.argtypes = [_VP, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.mosp_batch_set_cost_model.argtypes = [_VP, C.c_double]
    L.mosp_batch_create_capacity.restype = _VP
    L.mosp_batch_create_capacity.argtypes = [_I, _I, _L]
    L.mosp_batch_reset.argtypes = [_VP, C.POINTER(BatchInstance), _I]
    L.mosp_batch_instance_bytes.restype = _L
    L.mosp_batch_instance_bytes.argtypes = [_I, _I]
    L.mosp_batch_stream.restype = _VP
    L.mosp_batch_stream.argtypes = [_VP]
    L.mosp_server_create.restype = _VP
    L.mosp_server_create.argtypes = [_VP]
    L.mosp_server_add_graph.restype = _I
    L.mosp_server_add_graph.argtypes = [_VP, _I, _I, _i32p, _i32p, _i32p, _I, _i64p, _i32p,
                                        _i32p, _I, _I]
    L.mosp_server_start.argtypes = [_VP]
    L.mosp_server_submit.restype = _L
    L.mosp_server_submit.argtypes = [_VP, _I, _I, _i32p, _i32p, _i32p, _I, _i32p, _i32p]
    L.mosp_server_wait.argtypes = [_VP, _L, C.c_double]
    L.mosp_server_drain.argtypes = [_VP, C.c_double]
    L.mosp_server_request.argtypes = [_VP, _L, _VP]
    L.mosp_server_request_trees.argtypes = [_VP, _L, _i64p, _i32p, _i64p, _i32p]
    L.mosp_server_release.argtypes = [_VP, _L]
    L.mosp_server_stats.argtypes = [_VP, _VP, _VP, _I]
    L.mosp_server_reset_stats.argtypes = [_VP]
    L.mosp_server_clock.restype = C.c_double
    L.mosp_server_clock.argtypes = [_VP]
    L.mosp_server_destroy.argtypes = [_VP]
    _LIB = L
    return L


def last_error() -> str:
    return (lib().mosp_last_error() or b"").decode()


def _check(ok: int, what: str) -> None:
    if not ok:
        raise RuntimeError(f"{what} failed: {last_error()}")


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


def _nz(a: np.ndarray) -> np.ndarray:
    """numpy arrays passed through ndpointer must not be empty-with-None."""
    return a if a.size else np.zeros(1, a.dtype)


# --------------------------------------------------------------- host data --

@dataclass
class HostGraph:
    """The reference's CsrGraph: rowPtr, colInd, edge-major weights."""
    n: int
    K: int
    row: np.ndarray
    col: np.ndarray
    w: np.ndarray

    @property
    def m(self) -> int:
        return int(self.row[-1])


@dataclass
class Batch:
    """The reference's ChangeBatch (insertion-major weights)."""
    K: int
    ins_from: np.ndarray
    ins_to: np.ndarray
    ins_w: np.ndarray
    del_from: np.ndarray
    del_to: np.ndarray
    mask: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))

    @property
    def n_ins(self) -> int:
        return len(self.ins_from)

    @property
    def n_del(self) -> int:
        return len(self.del_from)


def gen_road_grid(width: int, height: int, K: int, seed: int, drop: float = 0.3,
                  diag: float = 0.02, wmin: int = 1, wmax: int = 100) -> HostGraph:
    """One call of the generator into buffers of an upper bound of its arcs
    (both directions of every grid edge and diagonal)."""
    L = lib()
    n = C.c_int32(0)
    m = C.c_int64(0)
    cells = width * height
    cap = 2 * (width * (height - 1) + height * (width - 1)) + 2 * int(diag * cells + 2) + 16
    row = np.zeros(cells + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    w = np.zeros(max(cap, 1) * K, np.int32)
    _check(L.mosp_gen_road_grid(width, height, K, wmin, wmax, drop, diag, seed, C.byref(n),
                                C.byref(m), row.ctypes.data, col.ctypes.data, w.ctypes.data,
                                cap), "road grid")
    return HostGraph(n.value, K, row[: n.value + 1], col[: m.value].copy(),
                     w[: m.value * K].copy())


def gen_rmat(scale: int, edge_factor: int, K: int, seed: int, wmin: int = 1, wmax: int = 100,
             a: float = 0.57, b: float = 0.19, c: float = 0.19) -> HostGraph:
    """One call of the generator into buffers for every draw (an upper
    bound: self-loops and duplicates are removed)."""
    L = lib()
    n = C.c_int32(0)
    m = C.c_int64(0)
    cap = edge_factor << scale
    row = np.zeros((1 << scale) + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    w = np.zeros(max(cap, 1) * K, np.int32)
    _check(L.mosp_gen_rmat(scale, edge_factor, a, b, c, K, wmin, wmax, seed, C.byref(n),
                           C.byref(m), row.ctypes.data, col.ctypes.data, w.ctypes.data, cap),
           "rmat")
    return HostGraph(n.value, K, row, col[: m.value].copy(), w[: m.value * K].copy())


def gen_batch(g: HostGraph, changes: int, ins_pct: float, seed: int, safe: bool = False,
              wmin: int = 1, wmax: int = 100, source: int = 0) -> Batch:
    cap = max(changes, 1)
    f, t, dfm, dto = (np.zeros(cap, np.int32) for _ in range(4))
    w = np.zeros(cap * g.K, np.int32)
    ni, nd = C.c_int32(0), C.c_int32(0)
    _check(lib().mosp_gen_batch(g.n, g.K, g.row, _nz(g.col), _nz(g.w), changes, ins_pct,
                                int(safe), wmin, wmax, seed, source, C.byref(ni), f, t, w,
                                C.byref(nd), dfm, dto), "batch generator")
    a, b = ni.value, nd.value
    return Batch(g.K, f[:a].copy(), t[:a].copy(), w[: a * g.K].copy(), dfm[:b].copy(),
                 dto[:b].copy())


def apply_batch(g: HostGraph, b: Batch) -> tuple[HostGraph, np.ndarray]:
    cap = g.m + b.n_ins
    row = np.zeros(g.n + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    w = np.zeros(max(cap, 1) * g.K, np.int32)
    mask = np.zeros(max(b.n_ins, 1), np.uint32)
    m = lib().mosp_apply_batch(g.n, g.K, g.row, _nz(g.col), _nz(g.w), b.n_ins,
                               _nz(_i32(b.ins_from)), _nz(_i32(b.ins_to)), _nz(_i32(b.ins_w)),
                               b.n_del, _nz(_i32(b.del_from)), _nz(_i32(b.del_to)), row, col, w,
                               mask)
    if m < 0:
        raise RuntimeError("applyChangeBatch failed")
    return HostGraph(g.n, g.K, row, col[:m].copy(), w[: m * g.K].copy()), mask[: b.n_ins].copy()


def canonicalize_tree(g: HostGraph, k: int, source: int, dist, parent) -> np.ndarray:
    """The product's lowest-id tie rule (canonicalizeTree, src/mospUpdate.cu:40-55)
    on a copy of `parent`."""
    p = _i32(parent).copy()
    lib().mosp_canonicalize_tree(g.n, g.K, g.row, _nz(g.col), _nz(g.w), k, source,
                                 np.ascontiguousarray(dist, dtype=np.int64), p)
    return p


def path_costs_host(g: HostGraph, parent, source: int) -> np.ndarray | None:
    """The product's host mospPathCosts: n*K costs, or None if a tree edge is
    missing (src/combinedGraphGpu.cu:161-208)."""
    costs = np.zeros(g.n * g.K, np.int64)
    ok = lib().mosp_path_costs_host(g.n, g.K, g.row, _nz(g.col), _nz(g.w), _i32(parent), source,
                                    costs)
    return costs if ok else None


def changed_lists(b: Batch, mask: np.ndarray, k: int) -> tuple[np.ndarray, np.ndarray]:
    """changed_k = deletions + insertions that raised objective k
    (src/mospUpdate.cu:124-136)."""
    sel = ((mask.astype(np.int64) >> k) & 1) == 1
    return (np.concatenate([_i32(b.del_from), _i32(b.ins_from)[sel]]).astype(np.int32),
            np.concatenate([_i32(b.del_to), _i32(b.ins_to)[sel]]).astype(np.int32))


def preference_scale(pref, K: int) -> int:
    a = _i32(pref)
    return lib().mosp_preference_scale(_nz(a), len(a), K)


def choose_packing(n: int, max_w: int):
    bits = C.c_int32(0)
    bound = C.c_uint64(0)
    ok = lib().mosp_choose_packing(n, max_w, C.byref(bits), C.byref(bound))
    return (bits.value, bound.value) if ok else None


def default_delta(m: int, n: int, wsum: int) -> int:
    return lib().mosp_default_delta(m, n, wsum)


def objective_bounds(g: HostGraph, b: Batch, K: int):
    """maxWeight (original + inserted weights) and weightSum (original edges)
    per objective (src/mospUpdate.cu:104-122)."""
    w = g.w.reshape(-1, g.K)[:, :K].astype(np.int64) if g.m else np.zeros((0, K), np.int64)
    maxw = np.maximum(w.max(axis=0) if g.m else np.ones(K, np.int64), 1)
    if b.n_ins:
        maxw = np.maximum(maxw, _i32(b.ins_w).reshape(-1, g.K)[:, :K].max(axis=0))
    return maxw.astype(np.int64), w.sum(axis=0).astype(np.int64)


def update_host(g: HostGraph, b: Batch, init_dist: np.ndarray, init_parents: np.ndarray,
                K: int = 0, source: int = 0, pref=(), delta: int = 0, fused: bool = True) -> dict:
    """The in-memory MOSP update (src/mospUpdate.cu:57-225) on host arrays:
    apply, upload, K updates + combined step on the GPU, download."""
    Ku = K if K > 0 else g.K
    Ku = min(Ku, g.K)
    n = g.n
    cap = g.m + b.n_ins
    row = np.zeros(n + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    w = np.zeros(max(cap, 1) * g.K, np.int32)
    dist = np.zeros(Ku * n, np.int64)
    par = np.zeros(Ku * n, np.int32)
    cd = np.zeros(n, np.int64)
    cp = np.zeros(n, np.int32)
    st = (SospStats * Ku)()
    cs = CombineStats()
    tm = np.zeros(5, np.float64)
    m = C.c_int64(0)
    p = _i32(pref)
    ok = lib().mosp_update_host(
        n, g.K, g.row, _nz(g.col), _nz(g.w), b.n_ins, _nz(_i32(b.ins_from)),
        _nz(_i32(b.ins_to)), _nz(_i32(b.ins_w)), b.n_del, _nz(_i32(b.del_from)),
        _nz(_i32(b.del_to)), Ku, source, _nz(p), len(p), delta, int(fused),
        np.ascontiguousarray(init_dist, np.int64).reshape(-1),
        np.ascontiguousarray(init_parents, np.int32).reshape(-1), C.byref(m), row, col, w,
        dist, par, cd, cp, st, C.byref(cs), tm)
    if not ok:
        raise RuntimeError(f"mospUpdate failed: {last_error()}")
    mm = m.value
    return {
        "updated": HostGraph(n, g.K, row, col[:mm].copy(), w[: mm * g.K].copy()),
        "dist": dist.reshape(Ku, n), "parents": par.reshape(Ku, n),
        "comb_dist": cd, "comb_parent": cp,
        "stats": [st[k].as_dict() for k in range(Ku)],
        "combined_edges": cs.n_edges, "scale": cs.scale, "comb_delta": cs.delta,
        "comb_stats": cs.search.as_dict(),
        "timings_ms": {"apply": tm[0], "upload": tm[1], "sosp": tm[2], "combined": tm[3],
                       "download": tm[4], "gpu_compute": tm[2] + tm[3]},
    }


# ------------------------------------------------------------- device side --

def device_count() -> int:
    return lib().mosp_device_count()


def set_device(device: int = 0) -> None:
    _check(lib().mosp_set_device(device), "hipSetDevice")


def synchronize() -> None:
    _check(lib().mosp_device_synchronize(), "hipDeviceSynchronize")


def _require_gpu() -> None:
    if device_count() < 1:
        raise RuntimeError("no GPU visible: the HIP engine needs an MI355X")


class DeviceBuffer:
    """Device memory owned by Python (hipMalloc through the library)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().mosp_dev_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise RuntimeError(last_error())
        _track(self)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        _check(lib().mosp_memcpy(buf.ptr, a.ctypes.data, a.nbytes, 1), "copy to device")
        return buf

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        _check(lib().mosp_memcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "copy to host")
        return out

    def free(self) -> None:
        if self.ptr:
            lib().mosp_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Workspace:
    """mosp_ws: scratch of the engine, reserved once and reused."""

    def __init__(self, capacity: int = 0, slots: int = 1, grid: int = 0):
        _require_gpu()
        self.handle = lib().mosp_ws_create()
        _track(self)
        if grid:
            lib().mosp_ws_set_grid(self.handle, grid)
        if capacity:
            self.reserve(capacity, slots)

    def reserve(self, capacity: int, slots: int = 1) -> None:
        _check(lib().mosp_ws_reserve(self.handle, capacity, slots), "workspace reserve")

    def set_grid(self, workgroups: int) -> None:
        lib().mosp_ws_set_grid(self.handle, workgroups)

    def set_barrier_timeout(self, ms: float) -> None:
        """Longest grid-barrier wait (ms) before a launch gives up with an error."""
        _check(lib().mosp_ws_set_barrier_timeout(self.handle, float(ms)), "barrier timeout")

    def set_shape(self, threads: int = 0, queue_cap: int = 0) -> None:
        """Workgroups of 1024 or 512 threads (0: automatic) and the push loop's
        LDS queue entries (256..4096, 0: all); results do not depend on them."""
        _check(lib().mosp_ws_set_shape(self.handle, int(threads), int(queue_cap)), "shape")

    def set_combined_delta_scale(self, scale: int) -> None:
        """Combined search width = the reference's rule x scale (>= 1)."""
        _check(lib().mosp_ws_set_combined_delta_scale(self.handle, int(scale)),
               "combined width scale")

    @property
    def grid(self) -> int:
        return lib().mosp_ws_grid(self.handle)

    PHASES = ("pack", "jump", "invalidate", "pull", "near_far", "unpack", "comb_count",
              "comb_scan", "comb_fill", "comb_search", "comb_unpack")

    def last_profile(self) -> dict:
        """Device-clock phase times (ms) and barrier count of the last launch."""
        ms = np.zeros(len(self.PHASES), np.float64)
        b = C.c_int32(0)
        lib().mosp_ws_last_profile(self.handle, ms, len(ms), C.byref(b))
        out = {p: round(float(x), 4) for p, x in zip(self.PHASES, ms)}
        out["barriers"] = b.value
        return out

    def async_events(self) -> dict:
        """Events of the last launch's asynchronous near-far loop."""
        ev = (C.c_int32 * 4)()
        _check(lib().mosp_ws_async_events(self.handle, ev), "mosp_ws_async_events")
        return {"lost_batches": ev[0], "short_spills": ev[1], "sweep_phases": ev[2],
                "async": bool(ev[3])}

    def close(self) -> None:
        if self.handle:
            lib().mosp_ws_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """mosp_device_graph: out-CSR + in-CSR with K weight columns in HBM."""

    def __init__(self, g: HostGraph | None = None):
        _require_gpu()
        _track(self)
        if g is None:  # empty: the target of apply()
            self.n = self.K = self.m = 0
            self.handle = lib().mosp_device_graph_create()
            return
        self.n, self.K, self.m = g.n, g.K, g.m
        self.handle = lib().mosp_device_graph_upload(g.n, g.K, g.row, _nz(g.col), _nz(g.w))
        if not self.handle:
            raise RuntimeError(f"device graph upload failed: {last_error()}")

    def apply(self, b: "Batch", out: "DeviceGraph", stream=None) -> np.ndarray:
        """Apply batch b in HBM (applyChangeBatch semantics), writing the
        updated graph into `out`; returns the weight-increase mask."""
        mask = np.zeros(max(b.n_ins, 1), np.uint32)
        s = C.c_void_p(stream) if stream is not None else None
        _check(lib().mosp_device_graph_apply(
            self.handle, b.n_ins, _nz(_i32(b.ins_from)), _nz(_i32(b.ins_to)),
            _nz(_i32(b.ins_w)), b.n_del, _nz(_i32(b.del_from)), _nz(_i32(b.del_to)),
            out.handle, mask, s), "mosp_device_graph_apply")
        out.n, out.K, out.m = self.n, self.K, int(lib().mosp_device_graph_edges(out.handle))
        return mask[: b.n_ins].copy()

    def weight_stats(self, stream=None) -> tuple[np.ndarray, np.ndarray]:
        """(max weight, weight sum) per objective, computed on the device."""
        mx = np.zeros(self.K, np.int64)
        sm = np.zeros(self.K, np.int64)
        _check(lib().mosp_device_graph_weight_stats(self.handle, mx, sm,
                                                    C.c_void_p(stream) if stream else None),
               "weight stats")
        return mx, sm

    def download(self) -> HostGraph:
        m = int(lib().mosp_device_graph_edges(self.handle))
        row = np.zeros(self.n + 1, np.int32)
        col = np.zeros(max(m, 1), np.int32)
        w = np.zeros(max(m, 1) * self.K, np.int32)
        _check(lib().mosp_device_graph_download(self.handle, row, col, w), "download")
        return HostGraph(self.n, self.K, row, col[:m].copy(), w[: m * self.K].copy())

    def view(self, k: int) -> tuple[Csr, Csr]:
        out, inn = Csr(), Csr()
        _check(lib().mosp_device_graph_view(self.handle, k, C.byref(out), C.byref(inn)), "view")
        return out, inn

    def download_reverse(self, k: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """The in-CSR built on the device: (row pointers, sources, objective k's
        weights); the order inside an in-row is unspecified."""
        _, inn = self.view(k)
        m = int(lib().mosp_device_graph_edges(self.handle))
        row = np.zeros(self.n + 1, np.int32)
        col = np.zeros(max(m, 1), np.int32)
        w = np.zeros(max(m, 1), np.int32)
        _check(lib().mosp_memcpy(row.ctypes.data, inn.row_ptr, row.nbytes, 2), "copy in-rows")
        if m > 0:
            _check(lib().mosp_memcpy(col.ctypes.data, inn.col_ind, 4 * m, 2), "copy in-columns")
            _check(lib().mosp_memcpy(w.ctypes.data, inn.weights, 4 * m, 2), "copy in-weights")
        return row, col[:m], w[:m]

    def close(self) -> None:
        if self.handle:
            lib().mosp_device_graph_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def parallel_sosp_update(csr_prefix: str, distances_in: str, tree_in: str, insert_path: str,
                         delete_path: str, objective: int, source: int, distances_out: str,
                         tree_out: str) -> bool:
    """parallelSOSPUpdate (the reference's file-based single-objective update)."""
    _require_gpu()
    enc = [str(x).encode() for x in (csr_prefix, distances_in, tree_in, insert_path,
                                     delete_path)]
    return bool(lib().mosp_parallel_sosp_update(*enc, int(objective), int(source),
                                                 str(distances_out).encode(),
                                                 str(tree_out).encode()))


def parallel_combined_graph(csr_prefix: str, tree_paths, source: int, distances_out: str,
                            tree_out: str, pref=()) -> bool:
    """parallelCombinedGraph (the reference's file-based combined graph)."""
    _require_gpu()
    paths = (C.c_char_p * max(len(tree_paths), 1))(*[str(t).encode() for t in tree_paths])
    pr = _i32(pref) if len(pref) else np.zeros(1, np.int32)
    return bool(lib().mosp_parallel_combined_graph(str(csr_prefix).encode(), paths,
                                                   len(tree_paths), int(source),
                                                   str(distances_out).encode(),
                                                   str(tree_out).encode(), pr, len(pref)))


def _dev(a, dtype) -> DeviceBuffer:
    return DeviceBuffer.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype)).reshape(-1))


def from_scratch(dg: DeviceGraph, k: int, source: int, delta: int, max_w: int, ws: Workspace):
    """sospFromScratchGpu on objective k; returns (dist, parent, stats) on host."""
    out, _ = dg.view(k)
    d = DeviceBuffer(8 * dg.n)
    p = DeviceBuffer(4 * dg.n)
    st = SospStats()
    _check(lib().mosp_sosp_from_scratch(C.byref(out), source, delta, max_w, ws.handle,
                                        d.ptr, p.ptr, C.byref(st), None), "sospFromScratchGpu")
    return d.to_numpy(np.int64, dg.n), p.to_numpy(np.int32, dg.n), st.as_dict()


def initial_trees(g: HostGraph, K: int = 0, source: int = 0, ws: Workspace | None = None):
    """Canonical trees of every objective by the GPU from-scratch search (the
    product's replacement of `mospPrep init`)."""
    K = K or g.K
    dg = DeviceGraph(g)
    ws = ws or Workspace(g.n)
    w = g.w.reshape(-1, g.K).astype(np.int64)
    dist = np.zeros((K, g.n), np.int64)
    par = np.zeros((K, g.n), np.int32)
    for k in range(K):
        maxw = max(int(w[:, k].max()) if g.m else 1, 1)
        delta = default_delta(g.m, g.n, int(w[:, k].sum()))
        dist[k], par[k], _ = from_scratch(dg, k, source, delta, maxw, ws)
    dg.close()
    return dist, par


def sosp_update(dg: DeviceGraph, k: int, changes_from, changes_to, heads, source: int,
                delta: int, max_w: int, ws: Workspace, dist, parent):
    """sospUpdateGpu on objective k with host arrays in/out."""
    out, inn = dg.view(k)
    cf, ct, hd = (_dev(_nz(_i32(a)), np.int32) for a in (changes_from, changes_to, heads))
    ch = Changes(cf.ptr, ct.ptr, len(changes_from), hd.ptr, len(heads))
    d = _dev(dist, np.int64)
    p = _dev(parent, np.int32)
    st = SospStats()
    _check(lib().mosp_sosp_update(C.byref(out), C.byref(inn), C.byref(ch), source, delta, max_w,
                                  ws.handle, d.ptr, p.ptr, C.byref(st), None), "sospUpdateGpu")
    return d.to_numpy(np.int64, dg.n), p.to_numpy(np.int32, dg.n), st.as_dict()


def sosp_update_noalloc(dg: DeviceGraph, k: int, changes_from, changes_to, heads, source: int,
                        delta: int, max_w: int, dist, parent):
    """mosp_sosp_update_noalloc on objective k (no workspace: the library's
    static scratch) with host arrays in/out."""
    out, inn = dg.view(k)
    cf, ct, hd = (_dev(_nz(_i32(a)), np.int32) for a in (changes_from, changes_to, heads))
    ch = Changes(cf.ptr, ct.ptr, len(changes_from), hd.ptr, len(heads))
    d = _dev(dist, np.int64)
    p = _dev(parent, np.int32)
    st = SospStats()
    _check(lib().mosp_sosp_update_noalloc(C.byref(out), C.byref(inn), C.byref(ch), source, delta,
                                          max_w, d.ptr, p.ptr, C.byref(st), None),
           "mosp_sosp_update_noalloc")
    return d.to_numpy(np.int64, dg.n), p.to_numpy(np.int32, dg.n), st.as_dict()


def combined_sosp(parents: np.ndarray, source: int, pref, ws: Workspace, delta: int = 0):
    """combinedGraphSospGpu on K parent arrays (objective-major)."""
    K, n = parents.shape
    dp = _dev(parents.reshape(-1), np.int32)
    d = DeviceBuffer(8 * n)
    p = DeviceBuffer(4 * n)
    cs = CombineStats()
    pa = _i32(pref)
    _check(lib().mosp_combined_sosp(dp.ptr, n, K, _nz(pa), len(pa), source, delta, ws.handle,
                                    d.ptr, p.ptr, C.byref(cs), None), "combinedGraphSospGpu")
    return d.to_numpy(np.int64, n), p.to_numpy(np.int32, n), {"n_edges": cs.n_edges, "scale": cs.scale,
                                              "delta": cs.delta, **cs.search.as_dict()}


def path_costs(g: HostGraph, parent, source: int) -> np.ndarray:
    """mospPathCosts on the GPU: n x K costs (DISTANCE_INF if unreachable)."""
    dg = DeviceGraph(g)
    p = _dev(_i32(parent), np.int32)
    c = DeviceBuffer(8 * g.n * g.K)
    _check(lib().mosp_path_costs_gpu(dg.handle, p.ptr, source, c.ptr, None), "mospPathCosts")
    out = c.to_numpy(np.int64, g.n * g.K).reshape(g.n, g.K)
    dg.close()
    return out


class Session:
    """A MOSP update resident in HBM: the updated graph, the old trees, the
    batch and every output stay on the device; `run()` is one pass of the hot
    path (K updates + combined graph + its search) in ONE persistent launch,
    reading the old trees and writing new buffers, so it can be repeated."""

    def __init__(self, original: HostGraph, batch: Batch, init_dist: np.ndarray,
                 init_parents: np.ndarray, K: int = 0, source: int = 0, pref=(),
                 delta: int = 0, grid: int = 0, delta_scale: float = 1.0,
                 comb_delta_scale: int = 1):
        self.K = min(K or original.K, original.K)
        self.n, self.source, self.pref = original.n, source, _i32(pref)
        self.updated, self.mask = apply_batch(original, batch)
        self.batch = batch
        self.maxw, wsum = objective_bounds(original, batch, self.K)
        m = max(original.m, 1)
        # The reference's default width per objective (original m and weight
        # sum, src/mospUpdate.cu:177-181); delta_scale multiplies it.
        self.delta_base = np.array([delta if delta > 0 else
                                    default_delta(m, self.n, int(wsum[k]))
                                    for k in range(self.K)], np.int64)
        self.set_delta_scale(1.0 if delta > 0 else delta_scale)
        self.dg = DeviceGraph(self.updated)
        self.ws = Workspace(self.n, self.K, grid)
        if comb_delta_scale > 1:
            self.ws.set_combined_delta_scale(comb_delta_scale)
        self.heads = _dev(_nz(_i32(batch.ins_to)), np.int32)
        self._keep = []
        changes = (Changes * self.K)()
        for k in range(self.K):
            cf, ct = changed_lists(batch, self.mask, k)
            a, b = _dev(_nz(cf), np.int32), _dev(_nz(ct), np.int32)
            self._keep += [a, b]
            changes[k] = Changes(a.ptr, b.ptr, len(cf), self.heads.ptr, batch.n_ins)
        self.changes = changes
        self.init_dist = _dev(np.asarray(init_dist)[: self.K].reshape(-1), np.int64)
        self.init_par = _dev(np.asarray(init_parents)[: self.K].reshape(-1), np.int32)
        self.dist = DeviceBuffer(8 * self.K * self.n)
        self.par = DeviceBuffer(4 * self.K * self.n)
        self.cdist = DeviceBuffer(8 * self.n)
        self.cpar = DeviceBuffer(4 * self.n)
        self.stats = (SospStats * self.K)()
        self.cstats = CombineStats()

    def set_delta_scale(self, scale: float) -> None:
        """Near-far widths = max(1, int(base width * scale)); outputs do not
        depend on them."""
        self.delta = np.maximum(1, (self.delta_base * float(scale)).astype(np.int64))

    def run(self, stream=None) -> None:
        s = C.c_void_p(stream) if stream is not None else None
        dl = (C.c_int64 * self.K)(*[int(x) for x in self.delta])
        mw = (C.c_int64 * self.K)(*[int(x) for x in self.maxw])
        _check(lib().mosp_update_fused(self.dg.handle, self.K, self.changes, self.source, dl, mw,
                                       _nz(self.pref), len(self.pref), self.ws.handle,
                                       self.init_dist.ptr, self.init_par.ptr, self.dist.ptr,
                                       self.par.ptr, self.cdist.ptr, self.cpar.ptr, self.stats,
                                       C.byref(self.cstats), s), "mosp_update_fused")

    def results(self) -> dict:
        n, K = self.n, self.K
        return {"dist": self.dist.to_numpy(np.int64, K * n).reshape(K, n),
                "parents": self.par.to_numpy(np.int32, K * n).reshape(K, n),
                "comb_dist": self.cdist.to_numpy(np.int64, n),
                "comb_parent": self.cpar.to_numpy(np.int32, n),
                "stats": [self.stats[k].as_dict() for k in range(self.K)],
                "combined_edges": self.cstats.n_edges,
                "kernel_ms": self.stats[0].kernel_ms,
                "sosp_phase_ms": self.cstats.sosp_phase_ms,
                "combined_phase_ms": self.cstats.combined_phase_ms}

    def close(self) -> None:
        self.dg.close()
        self.ws.close()
