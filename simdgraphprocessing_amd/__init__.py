"""simdgraphprocessing_amd -- MI355X (gfx950) set-intersection pattern engine.

Python binding over the C ABI in ``include/eh.h`` (``libeh.so``, built from
``csrc/`` with hipcc).  Every call goes to the HIP library; there is no CPU
fallback: importing works without a GPU, but creating a :class:`Context` fails
loudly when the library or the device is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeh.so")
REFERENCE_SRC = os.environ.get("EH_REFERENCE_SRC", "/root/reference")  # only read by build()

EH_BITSET, EH_PSHORT, EH_UINTEGER, EH_BITSET_NEW, EH_NEW_TYPE = 0, 1, 2, 7, 8
# kernel classes of a counting pass, in the order of eh_run_stats.part_ms (EH_PART_*)
PARTS = ("k_small_rows", "k_bstat", "k_bshl", "k_local4_mid", "k_local4_big", "k_count_big", "k_count")
STATIC_MAX_PAIRS = 1 << 20  # EH_STATIC_MAX_PAIRS: pairs eh_intersect_rows_static stages at most
TYPE_NAMES = {0: "bitset", 1: "pshort", 2: "uint"}
LAYOUTS = {"uint": 0, "pshort": 1, "bs": 2, "bitset": 2, "hybrid": 3, "hybrid3": 4, "learned": 5,
           "cost": 6, "bitset_new": 7, "new_type": 8, "hybrid_block": 9}
FILTERS = {"all": 0, "lower": 1, "mask": 2}
SCHEDULES = {"plain": 0, "tuned": 1}
ORDERINGS = {"degree": 0, "rev_degree": 1, "strong_run": 2, "bfs": 3, "shingles": 4, "random": 5, "the_game": 6}
MLP_PARAMS = 5508


class EHError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class RowView(C.Structure):
    _fields_ = [("data", C.c_void_p), ("cardinality", C.c_uint64), ("number_of_bytes", C.c_uint64),
                ("type", C.c_int32), ("reserved", C.c_int32)]


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("max_degree", C.c_uint32), ("entries", C.c_uint64)]


class MatrixInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("max_row", C.c_uint32), ("cardinality", C.c_uint64),
                ("rows_by_type", C.c_uint64 * 3), ("payload_bytes", C.c_uint64),
                ("ref_row_bytes", C.c_uint64), ("wedges", C.c_uint64), ("build_ms", C.c_double)]


class RunOpts(C.Structure):
    _fields_ = [("schedule", C.c_int32), ("lanes_per_row", C.c_int32), ("task_work", C.c_uint32),
                ("lds_words", C.c_uint32), ("waves_per_block", C.c_int32), ("repeat", C.c_int32),
                ("small_degree", C.c_int32), ("stream_ratio", C.c_int32), ("bstat_split", C.c_int32),
                ("local_max", C.c_int32), ("time_parts", C.c_int32), ("bsh_max", C.c_int32),
                ("short_queues", C.c_int32)]


class BfsInfo(C.Structure):
    _fields_ = [("path_length", C.c_uint32), ("rounds", C.c_uint32), ("reached", C.c_uint64),
                ("kernel_ms", C.c_double), ("bottom_up", C.c_uint32), ("reserved", C.c_uint32)]


BFS_MODES = {"top_down": 0, "symmetric": 1, "in_rows": 2}


class RunStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("sched_ms", C.c_double), ("tasks", C.c_uint64),
                ("alg_bytes", C.c_uint64), ("pairs", C.c_uint64 * 4), ("work_units", C.c_uint64),
                ("schedule", C.c_int32), ("lanes_per_row", C.c_int32), ("task_work", C.c_uint32),
                ("lds_words", C.c_uint32), ("part_ms", C.c_double * 8)]

    def as_dict(self):
        return {"kernel_ms": self.kernel_ms, "sched_ms": self.sched_ms, "tasks": self.tasks,
                "alg_bytes": self.alg_bytes, "pairs": list(self.pairs), "work_units": self.work_units,
                "schedule": self.schedule, "lanes_per_row": self.lanes_per_row,
                "task_work": self.task_work, "lds_words": self.lds_words,
                "part_ms": {name: self.part_ms[q] for q, name in enumerate(PARTS) if self.part_ms[q] > 0}}


class GenSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n", C.c_uint32), ("scale", C.c_uint32),
                ("edge_factor", C.c_uint32), ("m", C.c_uint64), ("seed", C.c_uint64),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("gamma", C.c_double),
                ("communities", C.c_uint32), ("community_size", C.c_uint32), ("relabel", C.c_int32),
                ("reserved", C.c_int32)]


# every symbol declared in include/eh.h
EXPORTS = [
    "eh_abi_version", "eh_status_str", "eh_last_error", "eh_ctx_create", "eh_ctx_destroy",
    "eh_ctx_device_name", "eh_graph_from_csr", "eh_graph_from_edges", "eh_graph_generate",
    "eh_graph_relabel_by_degree", "eh_graph_reorder", "eh_graph_info_get", "eh_graph_download", "eh_graph_download_perm",
    "eh_graph_free", "eh_matrix_build", "eh_matrix_build_host", "eh_matrix_info_get", "eh_matrix_row",
    "eh_matrix_download", "eh_matrix_row_types", "eh_matrix_free", "eh_intersect", "eh_intersect_rows",
    "eh_intersect_rows_static", "eh_intersect_row_sets", "eh_union", "eh_difference", "eh_triangle_count", "eh_clique_count", "eh_bfs_path", "eh_count_profile", "eh_batch_generate", "eh_batch_info",
    "eh_batch_graph_params", "eh_batch_graph_download", "eh_batch_count", "eh_batch_alg_bytes", "eh_batch_free",
    "eh_batch_rebuild", "eh_batch_matrix",
    "eh_selector_set_params", "eh_selector_get_params", "eh_selector_set_rule", "eh_selector_set_schedule",
    "eh_matrix_features",
    "eh_matrix_build_typed", "eh_mlp_forward", "eh_mlp_loss_grad", "eh_selector_train",
]

_lib = None
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)


def build(verbose=False):
    """Compile libeh.so for gfx950 (hipcc; cross-compiles without a GPU) and the C++ apps."""
    jobs = f"-j{min(os.cpu_count() or 4, 16)}"
    # the device library, then the host C++ apps and tests that link it
    steps = [("csrc", []), ("apps", [])]
    # where the reference's sources are, its apps are also compiled, unchanged, against the
    # C++ surface (apps/Makefile target `ref`); nothing of the reference is copied
    if os.path.isdir(os.path.join(REFERENCE_SRC, "apps")):
        steps.append(("apps", ["ref", f"REF={REFERENCE_SRC}"]))
    for sub, extra in steps:
        r = subprocess.run(["make", "-C", os.path.join(HERE, sub), jobs] + extra, capture_output=not verbose,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{sub} build failed:\n" + (r.stdout or "")[-4000:] + (r.stderr or "")[-4000:])


def lib():
    """Load libeh.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run simdgraphprocessing_amd.build() first")
    L = C.CDLL(LIB_PATH)
    L.eh_status_str.restype = C.c_char_p
    L.eh_last_error.restype = C.c_char_p
    L.eh_ctx_create.argtypes = [C.c_int, _pp]
    L.eh_ctx_destroy.argtypes = [_vp]
    L.eh_ctx_destroy.restype = None
    L.eh_ctx_device_name.argtypes = [_vp, C.c_char_p, C.c_size_t]
    L.eh_graph_from_csr.argtypes = [_vp, C.c_uint32, _vp, _vp, _pp]
    L.eh_graph_from_edges.argtypes = [_vp, C.c_uint32, C.c_uint64, _vp, _vp, C.c_int, _pp]
    L.eh_graph_generate.argtypes = [_vp, C.POINTER(GenSpec), _pp]
    L.eh_graph_relabel_by_degree.argtypes = [_vp]
    L.eh_graph_reorder.argtypes = [_vp, C.c_int, C.c_uint64]
    L.eh_graph_info_get.argtypes = [_vp, C.POINTER(GraphInfo)]
    L.eh_graph_download.argtypes = [_vp, _vp, _vp]
    L.eh_graph_download_perm.argtypes = [_vp, _vp]
    L.eh_graph_free.argtypes = [_vp]
    L.eh_graph_free.restype = None
    L.eh_matrix_build.argtypes = [_vp, C.c_int, _vp, C.c_int, _pp]
    L.eh_matrix_build_host.argtypes = [_vp, C.c_uint32, _vp, _vp, C.c_int, _vp, C.c_int, _pp]
    L.eh_matrix_info_get.argtypes = [_vp, C.POINTER(MatrixInfo)]
    L.eh_matrix_row.argtypes = [_vp, C.c_uint32, _vp, C.c_size_t, C.POINTER(RowView)]
    L.eh_matrix_download.argtypes = [_vp, _vp, _vp]
    L.eh_matrix_row_types.argtypes = [_vp, _vp]
    L.eh_matrix_free.argtypes = [_vp]
    L.eh_matrix_free.restype = None
    L.eh_intersect.argtypes = [_vp, C.POINTER(RowView), C.POINTER(RowView), C.c_int, _vp, C.c_size_t,
                               C.POINTER(RowView)]
    L.eh_intersect_rows.argtypes = [_vp, _vp, _vp, C.c_uint64, _vp]
    L.eh_intersect_rows_static.argtypes = [_vp, _vp, _vp, C.c_uint64, _vp]
    L.eh_intersect_row_sets.argtypes = [_vp, _vp, _vp, C.c_uint64, _vp, _vp, C.c_uint64, _vp]
    for name in ("eh_union", "eh_difference"):
        getattr(L, name).argtypes = [_vp, C.POINTER(RowView), C.POINTER(RowView), _vp, C.c_size_t,
                                     C.POINTER(RowView)]
    L.eh_bfs_path.argtypes = [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _vp, C.c_uint32,
                              C.POINTER(BfsInfo)]
    L.eh_triangle_count.argtypes = [_vp, C.POINTER(RunOpts), C.POINTER(C.c_uint64), C.POINTER(RunStats)]
    L.eh_clique_count.argtypes = [_vp, C.c_uint32, C.POINTER(RunOpts), C.POINTER(C.c_uint64),
                                  C.POINTER(RunStats)]
    L.eh_count_profile.argtypes = [_vp, C.c_uint32, C.POINTER(RunOpts), _vp, _vp, _vp, C.c_uint64,
                                   C.POINTER(C.c_uint64)]
    L.eh_batch_generate.argtypes = [_vp, C.c_uint64, C.c_uint32, C.c_int, _pp]
    L.eh_batch_info.argtypes = [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.eh_batch_graph_params.argtypes = [_vp, _vp, _vp]
    L.eh_batch_graph_download.argtypes = [_vp, C.c_uint32, _vp, _vp, C.c_uint64]
    L.eh_batch_count.argtypes = [_vp, C.c_int, C.POINTER(RunOpts), C.POINTER(RunOpts), _vp, _vp,
                                 C.POINTER(C.c_double)]
    L.eh_batch_alg_bytes.argtypes = [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.eh_batch_free.argtypes = [_vp]
    L.eh_batch_free.restype = None
    L.eh_batch_rebuild.argtypes = [_vp, C.c_int, _vp]
    L.eh_batch_matrix.argtypes = [_vp, _pp]
    L.eh_selector_set_params.argtypes = [_vp, _vp]
    L.eh_selector_get_params.argtypes = [_vp, _vp]
    L.eh_selector_set_rule.argtypes = [_vp, C.c_uint32, C.c_int32]
    L.eh_selector_set_schedule.argtypes = [_vp, C.c_int32, C.c_int32]
    L.eh_matrix_features.argtypes = [_vp, _vp]
    L.eh_matrix_build_typed.argtypes = [_vp, C.c_int, _vp, _pp]
    L.eh_mlp_forward.argtypes = [_vp, _vp, _vp, C.c_int, _vp]
    L.eh_mlp_loss_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(C.c_float), _vp]
    L.eh_selector_train.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint64,
                                    C.POINTER(C.c_float)]
    _lib = L
    return L


def _check(status):
    if status != 0:
        L = lib()
        raise EHError(status, f"{L.eh_status_str(status).decode()}: {L.eh_last_error().decode()}")


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _opts(schedule="tuned", lanes=0, task_work=0, lds_words=0, waves=0, repeat=1, small_degree=0,
          stream_ratio=0, bstat=0, local_max=0, time_parts=False, bsh=0, short_queues=0):
    o = RunOpts()
    o.bsh_max = bsh
    o.short_queues = short_queues
    o.time_parts = int(bool(time_parts))
    o.bstat_split = bstat
    o.local_max = local_max
    o.small_degree = small_degree
    o.stream_ratio = stream_ratio
    o.schedule = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
    o.lanes_per_row = lanes
    o.task_work = task_work
    o.lds_words = lds_words
    o.waves_per_block = waves
    o.repeat = repeat
    return o


class Context:
    """A device plus one HIP stream (eh_ctx)."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        _check(L.eh_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().eh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def device_name(self):
        buf = C.create_string_buffer(256)
        _check(lib().eh_ctx_device_name(self.h, buf, 256))
        return buf.value.decode()

    # ---- graphs
    def graph_from_csr(self, n, off, nbr):
        off = np.ascontiguousarray(off, np.uint64)
        nbr = np.ascontiguousarray(nbr, np.uint32)
        h = C.c_void_p()
        _check(lib().eh_graph_from_csr(self.h, n, _ptr(off), _ptr(nbr), C.byref(h)))
        return Graph(self, h)

    def graph_from_edges(self, n, src, dst, relabel=True):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        h = C.c_void_p()
        _check(lib().eh_graph_from_edges(self.h, n, len(src), _ptr(src), _ptr(dst), int(relabel), C.byref(h)))
        return Graph(self, h)

    def generate(self, kind, relabel=True, **kw):
        s = GenSpec()
        s.kind = {"er": 0, "rmat": 1, "chung_lu": 2, "er_exact": 3}[kind]
        s.relabel = int(relabel)
        s.a, s.b, s.c = kw.get("a", 0.57), kw.get("b", 0.19), kw.get("c", 0.19)
        s.gamma = kw.get("gamma", 2.1)
        s.n = kw.get("n", 0)
        s.scale = kw.get("scale", 0)
        s.edge_factor = kw.get("edge_factor", 16)
        s.m = kw.get("m", 0)
        s.seed = kw.get("seed", 0)
        s.communities = kw.get("communities", 0)
        s.community_size = kw.get("community_size", 256)
        h = C.c_void_p()
        _check(lib().eh_graph_generate(self.h, C.byref(s), C.byref(h)))
        return Graph(self, h)

    def generate_er(self, n, m, seed, relabel=True):
        return self.generate("er", relabel, n=n, m=m, seed=seed)

    def generate_er_exact(self, n, m, seed, relabel=True):
        """G(n, m): exactly m distinct non-loop pairs, the first m distinct pairs of generate_er's
        draw stream."""
        return self.generate("er_exact", relabel, n=n, m=m, seed=seed)

    def generate_rmat(self, scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, relabel=True):
        return self.generate("rmat", relabel, scale=scale, edge_factor=edge_factor, seed=seed, a=a, b=b, c=c)

    def generate_chung_lu(self, n, mean_degree, gamma, seed, communities=0, community_size=256, relabel=True):
        return self.generate("chung_lu", relabel, n=n, m=n * mean_degree // 2, gamma=gamma, seed=seed,
                             communities=communities, community_size=community_size)

    def matrix_from_csr(self, n, off, nbr, filter="lower", policy="hybrid", mask=None):
        off = np.ascontiguousarray(off, np.uint64)
        nbr = np.ascontiguousarray(nbr, np.uint32)
        mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        h = C.c_void_p()
        _check(lib().eh_matrix_build_host(self.h, n, _ptr(off), _ptr(nbr), FILTERS[filter], _ptr(mk),
                                          LAYOUTS[policy], C.byref(h)))
        return SparseMatrix(self, h)

    # ---- set ops
    def intersect(self, a, b, materialize=True, out_cap=None):
        """a, b: (type, cardinality, payload bytes) in the reference formats."""
        L = lib()
        keep = []

        def view(s):
            t, card, payload = s
            p = np.ascontiguousarray(payload, np.uint8)
            keep.append(p)
            return RowView(p.ctypes.data if len(p) else None, card, len(p), t, 0)

        va, vb = view(a), view(b)
        cap = out_cap if out_cap is not None else 16 * (a[1] + b[1]) + len(a[2]) + len(b[2]) + 64
        out = np.zeros(cap, np.uint8)
        ov = RowView()
        _check(L.eh_intersect(self.h, C.byref(va), C.byref(vb), int(materialize), _ptr(out), cap, C.byref(ov)))
        return int(ov.cardinality), int(ov.type), out[:ov.number_of_bytes].copy()

    def _setop(self, fn, a, b, cap):
        keep = []

        def view(s):
            t, card, payload = s
            p = np.ascontiguousarray(payload, np.uint8)
            keep.append(p)
            return RowView(p.ctypes.data if len(p) else None, card, len(p), t, 0)

        va, vb = view(a), view(b)
        out = np.zeros(max(cap, 8), np.uint8)
        ov = RowView()
        _check(getattr(lib(), fn)(self.h, C.byref(va), C.byref(vb), _ptr(out), len(out), C.byref(ov)))
        return int(ov.cardinality), int(ov.type), out[:ov.number_of_bytes].copy()

    def union(self, a, b):
        """ops::set_union(Set<bitset>* A, Set<X>* B): A's bitset with B's elements added."""
        return self._setop("eh_union", a, b, len(a[2]) + 8)

    def difference(self, a, b):
        """ops::set_difference(C, bitset A, bitset B): the elements of A not in B (uint)."""
        return self._setop("eh_difference", a, b, 256 * len(a[2]) + 8)

    # ---- learned selector
    def set_selector(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_PARAMS
        _check(lib().eh_selector_set_params(self.h, _ptr(p)))

    def set_selector_rule(self, guard_words=8, fixed_arm=-1):
        """The LEARNED policy's bitset guard (words per id) and, for the ablations, a fixed arm
        (0 bitset, 1 pshort, 2 uint, 3 uint searched) in place of the network's (-1)."""
        _check(lib().eh_selector_set_rule(self.h, int(guard_words), int(fixed_arm)))

    def set_selector_schedule(self, bstat=0, bsh=0):
        """The tuned triangle schedule the LEARNED policy selects for (the pass's bstat_split and
        bsh_max; 0: their defaults, bstat < 0: every pair A-side): only rows that some A-side or
        edge-centric pair of that schedule reads are candidates."""
        _check(lib().eh_selector_set_schedule(self.h, int(bstat), int(bsh)))

    def get_selector(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().eh_selector_get_params(self.h, _ptr(p)))
        return p

    def mlp_forward(self, params, x):
        p = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], 4), np.float32)
        _check(lib().eh_mlp_forward(self.h, _ptr(p), _ptr(x), x.shape[0], _ptr(y)))
        return y

    def mlp_loss_grad(self, params, x, target, mask):
        p = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        _check(lib().eh_mlp_loss_grad(self.h, _ptr(p), _ptr(x), _ptr(t), _ptr(m), x.shape[0], C.byref(loss), _ptr(g)))
        return float(loss.value), g

    def train_selector(self, x, target, mask, epochs=200, batch=256, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        loss = C.c_float()
        _check(lib().eh_selector_train(self.h, _ptr(x), _ptr(t), _ptr(m), x.shape[0], epochs, batch, lr, seed,
                                       C.byref(loss)))
        return float(loss.value)

    # ---- batches
    def generate_batch(self, seed, ngraphs, policy="hybrid"):
        h = C.c_void_p()
        _check(lib().eh_batch_generate(self.h, seed, ngraphs, LAYOUTS[policy], C.byref(h)))
        return Batch(self, h)


class Graph:
    """Canonical symmetric graph on the device (eh_graph)."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def __del__(self):
        if getattr(self, "h", None):
            lib().eh_graph_free(self.h)
            self.h = None

    def info(self):
        gi = GraphInfo()
        _check(lib().eh_graph_info_get(self.h, C.byref(gi)))
        return {"n": gi.n, "max_degree": gi.max_degree, "entries": gi.entries}

    def download(self):
        i = self.info()
        off = np.zeros(i["n"] + 1, np.uint64)
        nbr = np.zeros(max(i["entries"], 1), np.uint32)
        _check(lib().eh_graph_download(self.h, _ptr(off), _ptr(nbr)))
        return off, nbr[:i["entries"]]

    def perm(self):
        p = np.zeros(self.info()["n"], np.uint32)
        _check(lib().eh_graph_download_perm(self.h, _ptr(p)))
        return p

    def relabel_by_degree(self):
        _check(lib().eh_graph_relabel_by_degree(self.h))

    def reorder(self, ordering, seed=0):
        """One of ORDERINGS (MutableGraph::reorder_*) on the device."""
        _check(lib().eh_graph_reorder(self.h, ORDERINGS[ordering], seed))

    def build_matrix_typed(self, types, filter="lower"):
        """A matrix whose rows take the given arms (0 bitset, 1 pshort, 2 uint, 3 uint searched;
        a bitset only within the context's guard)."""
        t = np.ascontiguousarray(types, np.uint8)
        assert t.shape == (self.info()["n"],)
        h = C.c_void_p()
        _check(lib().eh_matrix_build_typed(self.h, FILTERS[filter], _ptr(t), C.byref(h)))
        return SparseMatrix(self.ctx, h)

    def build_matrix(self, filter="lower", policy="hybrid", mask=None):
        mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        h = C.c_void_p()
        _check(lib().eh_matrix_build(self.h, FILTERS[filter], _ptr(mk), LAYOUTS[policy], C.byref(h)))
        return SparseMatrix(self.ctx, h)


class SparseMatrix:
    """Built per-row layouts on the device (eh_matrix).  owner: the batch a borrowed arena
    matrix belongs to (kept alive here, and the matrix is not freed)."""

    def __init__(self, ctx, h, owner=None):
        self.ctx, self.h, self.owner = ctx, h, owner

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "owner", None) is None:
            lib().eh_matrix_free(self.h)
        self.h = None

    def info(self):
        mi = MatrixInfo()
        _check(lib().eh_matrix_info_get(self.h, C.byref(mi)))
        return {"n": mi.n, "max_row": mi.max_row, "cardinality": mi.cardinality,
                "rows_by_type": {"bitset": mi.rows_by_type[0], "pshort": mi.rows_by_type[1],
                                 "uint": mi.rows_by_type[2]},
                "payload_bytes": mi.payload_bytes, "ref_row_bytes": mi.ref_row_bytes,
                "wedges": mi.wedges, "build_ms": mi.build_ms}

    def row(self, i, cap=1 << 22):
        buf = np.zeros(cap, np.uint8)
        v = RowView()
        _check(lib().eh_matrix_row(self.h, i, _ptr(buf), cap, C.byref(v)))
        return int(v.type), int(v.cardinality), buf[:v.number_of_bytes].copy()

    def download(self):
        i = self.info()
        off = np.zeros(i["n"] + 1, np.uint64)
        nbr = np.zeros(max(i["cardinality"], 1), np.uint32)
        _check(lib().eh_matrix_download(self.h, _ptr(off), _ptr(nbr)))
        return off, nbr[:i["cardinality"]]

    def row_types(self):
        t = np.zeros(self.info()["n"], np.uint8)
        _check(lib().eh_matrix_row_types(self.h, _ptr(t)))
        return t

    def features(self):
        f = np.zeros((self.info()["n"], 16), np.float32)
        _check(lib().eh_matrix_features(self.h, _ptr(f)))
        return f

    def triangle_count(self, schedule="tuned", **kw):
        c = C.c_uint64()
        st = RunStats()
        o = _opts(schedule, **kw)
        _check(lib().eh_triangle_count(self.h, C.byref(o), C.byref(c), C.byref(st)))
        return int(c.value), st.as_dict()

    def clique_count(self, k=4, schedule="tuned", **kw):
        c = C.c_uint64()
        st = RunStats()
        o = _opts(schedule, **kw)
        _check(lib().eh_clique_count(self.h, k, C.byref(o), C.byref(c), C.byref(st)))
        return int(c.value), st.as_dict()

    def profile(self, k=3, schedule="tuned", **kw):
        """One counting pass with per-task shader-clock cycles: (rows, kinds, cycles)."""
        o = _opts(schedule, **kw)
        nt = C.c_uint64()
        st = lib().eh_count_profile(self.h, k, C.byref(o), None, None, None, 0, C.byref(nt))
        if st not in (0, 6):
            _check(st)
        n = int(nt.value)
        rows = np.zeros(max(n, 1), np.uint32)
        kinds = np.zeros(max(n, 1), np.uint32)
        cyc = np.zeros(max(n, 1), np.uint64)
        _check(lib().eh_count_profile(self.h, k, C.byref(o), _ptr(rows), _ptr(kinds), _ptr(cyc), n, C.byref(nt)))
        return rows[:n], kinds[:n], cyc[:n]

    def bfs_path(self, start, max_depth=0xEFFFFFF, n_vertices=0, mode="top_down", level_cap=256):
        """n_path's BFS (apps/n_path.cpp:27-87) from row `start`; mode: top_down, symmetric
        (rows are symmetric) or in_rows (rows [n, 2n) are in-neighbourhoods)."""
        sizes = np.zeros(max(level_cap, 1), np.uint64)
        bi = BfsInfo()
        _check(lib().eh_bfs_path(self.h, n_vertices, start, max_depth, BFS_MODES[mode], _ptr(sizes), level_cap,
                                 C.byref(bi)))
        return {"path_length": bi.path_length, "rounds": bi.rounds, "reached": bi.reached,
                "level_sizes": sizes[:min(bi.rounds, level_cap)].tolist(), "kernel_ms": bi.kernel_ms,
                "bottom_up": bi.bottom_up}

    def intersect_rows(self, a_rows, b_rows):
        a = np.ascontiguousarray(a_rows, np.uint32)
        b = np.ascontiguousarray(b_rows, np.uint32)
        out = np.zeros(len(a), np.uint64)
        _check(lib().eh_intersect_rows(self.h, _ptr(a), _ptr(b), len(a), _ptr(out)))
        return out

    def intersect_row_sets(self, sets, materialize=False):
        """sets: a list of row-id lists (1 to 16 rows each) -> (counts, [sorted elements per set]
        or None)."""
        row_off = np.zeros(len(sets) + 1, np.uint64)
        row_off[1:] = np.cumsum([len(r) for r in sets])
        ids = np.ascontiguousarray(np.concatenate([np.asarray(r, np.uint32) for r in sets]) if sets else
                                   np.zeros(0, np.uint32), np.uint32)
        counts = np.zeros(max(len(sets), 1), np.uint64)
        eoff = np.zeros(len(sets) + 1, np.uint64)
        L = lib()
        _check(L.eh_intersect_row_sets(self.h, _ptr(ids), _ptr(row_off), len(sets), _ptr(counts), None, 0,
                                       _ptr(eoff)))
        if not materialize:
            return counts[:len(sets)], None
        elems = np.zeros(max(int(eoff[-1]), 1), np.uint32)
        _check(L.eh_intersect_row_sets(self.h, _ptr(ids), _ptr(row_off), len(sets), _ptr(counts), _ptr(elems),
                                       len(elems), _ptr(eoff)))
        return counts[:len(sets)], [elems[int(eoff[q]):int(eoff[q + 1])] for q in range(len(sets))]

    def intersect_rows_static(self, a_rows, b_rows):
        """intersect_rows without device allocation: staged in the library's static array
        of STATIC_MAX_PAIRS pairs (more raises EHError with EH_ERR_CAPACITY)."""
        a = np.ascontiguousarray(a_rows, np.uint32)
        b = np.ascontiguousarray(b_rows, np.uint32)
        out = np.zeros(len(a), np.uint64)
        _check(lib().eh_intersect_rows_static(self.h, _ptr(a), _ptr(b), len(a), _ptr(out)))
        return out


class Batch:
    """Many independent graphs in one device arena (eh_batch)."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def __del__(self):
        if getattr(self, "h", None):
            lib().eh_batch_free(self.h)
            self.h = None

    def info(self):
        g = C.c_uint32()
        v = C.c_uint64()
        e = C.c_uint64()
        _check(lib().eh_batch_info(self.h, C.byref(g), C.byref(v), C.byref(e)))
        return {"graphs": g.value, "vertices": v.value, "edges": e.value}

    def rebuild(self, policy="hybrid", types=None):
        """Build the arena's layouts again under `policy`, or with the per-row arms `types`
        (a uint8 per arena row, as build_matrix_typed takes them).  Matrices borrowed from the
        batch before (``matrix()``) are invalid afterwards."""
        t = None if types is None else np.ascontiguousarray(types, np.uint8)
        if t is not None and len(t) != self.info()["vertices"]:
            raise ValueError("one type per arena row")
        _check(lib().eh_batch_rebuild(self.h, LAYOUTS[policy], _ptr(t)))

    def matrix(self):
        """The arena's oriented matrix (borrowed: valid until rebuild)."""
        h = C.c_void_p()
        _check(lib().eh_batch_matrix(self.h, C.byref(h)))
        return SparseMatrix(self.ctx, h, owner=self)

    def alg_bytes(self):
        """Algorithmic bytes of the triangle and 4-clique passes, summed over the graphs."""
        t, k = C.c_uint64(), C.c_uint64()
        _check(lib().eh_batch_alg_bytes(self.h, C.byref(t), C.byref(k)))
        return int(t.value), int(k.value)

    def params(self):
        ng = self.info()["graphs"]
        n = np.zeros(ng, np.uint32)
        m = np.zeros(ng, np.uint64)
        _check(lib().eh_batch_graph_params(self.h, _ptr(n), _ptr(m)))
        return n, m

    def graph(self, g, cap=1 << 24):
        n = int(self.params()[0][g])
        off = np.zeros(n + 1, np.uint64)
        nbr = np.zeros(cap, np.uint32)
        _check(lib().eh_batch_graph_download(self.h, g, _ptr(off), _ptr(nbr), cap))
        return n, off, nbr[:int(off[-1])].copy()

    def count(self, mode="persistent", schedule="tuned", triangles=True, cliques=True, tri_opts=None,
              k4_opts=None, **kw):
        """Per-graph triangle and 4-clique counts.  ``kw`` sets both passes' options;
        ``tri_opts`` / ``k4_opts`` (dicts) override them for one pass."""
        ng = self.info()["graphs"]
        tri = np.zeros(ng, np.uint64) if triangles else None
        k4 = np.zeros(ng, np.uint64) if cliques else None
        ms = C.c_double()
        o3 = _opts(schedule, **{**kw, **(tri_opts or {})})
        o4 = _opts(schedule, **{**kw, **(k4_opts or {})})
        _check(lib().eh_batch_count(self.h, {"persistent": 0, "per_graph": 1}[mode], C.byref(o3), C.byref(o4),
                                    _ptr(tri), _ptr(k4), C.byref(ms)))
        return tri, k4, ms.value
