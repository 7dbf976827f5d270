"""CPU oracle for the set-intersection engine -- TEST INFRASTRUCTURE ONLY.

This package is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import it.  It wraps ``liboracle.so``
(built from ``eh_oracle.c`` by ``oracle/Makefile``), a C restatement of the
reference algorithm; each C function cites the reference file:line it follows.

It also knows how to run the reference's own programs built into
``oracle/_ref/`` (``make -C oracle ref``) on a graph written in the reference's
binary format, for cross-checking the restatement.
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

BITSET, PSHORT, UINTEGER, BITSET_NEW, NEW_TYPE = 0, 1, 2, 7, 8
POLICY = {"uint": 0, "pshort": 1, "bs": 2, "hybrid": 3, "hybrid3": 4,
          "bitset_new": 7, "new_type": 8, "hybrid_block": 9}
_BUILDER = {BITSET: "bitset", PSHORT: "pshort", BITSET_NEW: "bitset_new", NEW_TYPE: "new_type"}
ALG = {"scalar": 0, "standard": 1, "galloping": 2, "select": 3}

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcSet(C.Structure):
    _fields_ = [("data", C.c_void_p), ("cardinality", C.c_uint64),
                ("number_of_bytes", C.c_uint64), ("type", C.c_int32), ("pad", C.c_int32)]


_lib = None


def build():
    """Compile liboracle.so (and the reference programs when the sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    L.orc_row_type.argtypes = [_u32p, C.c_size_t, C.c_int]
    L.orc_row_image.argtypes = [_u8p, _u32p, C.c_size_t, C.c_int]
    L.orc_row_image.restype = C.c_size_t
    L.orc_perm.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _u32p]
    L.orc_perm.restype = None
    L.orc_order.argtypes = [C.c_uint32, _u64p, _u32p, C.c_int, C.c_uint64, _u32p]
    L.orc_order.restype = C.c_int
    L.orc_row_image_bound.argtypes = [_u32p, C.c_size_t]
    L.orc_bfs_path.argtypes = [C.c_uint32, _u64p, _u32p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                               C.c_uint32, C.c_void_p]
    L.orc_bfs_path.restype = C.c_uint32
    L.orc_bfs_path_mt.argtypes = [C.c_uint32, _u64p, _u32p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                  C.c_uint32, C.c_void_p, C.c_int]
    L.orc_bfs_path_mt.restype = C.c_uint32
    L.orc_bitset_difference.argtypes = [_u8p, C.c_size_t, _u8p, C.c_size_t, _u32p]
    L.orc_bitset_difference.restype = C.c_size_t
    L.orc_row_image_bound.restype = C.c_size_t
    for name in ("orc_uint_build", "orc_pshort_build", "orc_bitset_build", "orc_bitset_new_build",
                 "orc_new_type_build"):
        f = getattr(L, name)
        f.argtypes = [_u8p, _u32p, C.c_size_t]
        f.restype = C.c_size_t
    L.orc_decode.argtypes = [C.c_int, _u8p, C.c_size_t, C.c_size_t, _u32p]
    L.orc_decode.restype = C.c_size_t
    L.orc_isect_uint.argtypes = [C.c_int, _u32p, C.c_size_t, _u32p, C.c_size_t, C.c_void_p]
    L.orc_isect_uint.restype = C.c_size_t
    L.orc_set_intersect.argtypes = [C.POINTER(OrcSet), C.POINTER(OrcSet), C.POINTER(OrcSet)]
    L.orc_set_intersect.restype = C.c_uint64
    L.orc_edges_to_csr.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p,
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
    L.orc_relabel_by_degree.argtypes = [C.c_uint32, _u64p, _u32p, _u32p, _u64p, _u32p]
    L.orc_orient_lower.argtypes = [C.c_uint32, _u64p, _u32p, _u64p, C.c_void_p]
    L.orc_orient_lower.restype = C.c_uint64
    L.orc_load_binary.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_void_p),
                                  C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
    L.orc_write_binary.argtypes = [C.c_char_p, C.c_uint32, _u64p, _u32p, C.c_void_p]
    L.orc_free.argtypes = [C.c_void_p]
    L.orc_matrix_build.argtypes = [C.c_uint32, _u64p, _u32p, C.c_int, C.c_int, C.c_int]
    L.orc_matrix_build.restype = C.c_void_p
    L.orc_matrix_free.argtypes = [C.c_void_p]
    L.orc_matrix_get_row.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(OrcSet)]
    L.orc_matrix_type_counts.argtypes = [C.c_void_p, _u64p]
    L.orc_matrix_bytes.argtypes = [C.c_void_p]
    L.orc_matrix_bytes.restype = C.c_uint64
    L.orc_matrix_cardinality.argtypes = [C.c_void_p]
    L.orc_matrix_cardinality.restype = C.c_uint64
    L.orc_matrix_max_row.argtypes = [C.c_void_p]
    L.orc_matrix_max_row.restype = C.c_uint32
    L.orc_triangle_count.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.orc_triangle_count.restype = C.c_uint64
    L.orc_triangle_count_sample.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    L.orc_triangle_count_sample.restype = C.c_uint64
    L.orc_clique_count.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.orc_clique_count.restype = C.c_uint64
    L.orc_clique_count_sample.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                                          C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    L.orc_clique_count_sample.restype = C.c_uint64
    L.orc_count_rows.argtypes = [C.c_void_p, C.c_int, C.c_int, _u64p]
    L.orc_count_rows.restype = C.c_uint64
    L.orc_pair_mix.argtypes = [C.c_void_p, _u64p]
    L.orc_pair_mix.restype = C.c_uint64
    L.orc_brute_triangles.argtypes = [C.c_uint32, _u64p, _u32p]
    L.orc_brute_triangles.restype = C.c_uint64
    L.orc_brute_cliques.argtypes = [C.c_uint32, _u64p, _u32p, C.c_int]
    L.orc_brute_cliques.restype = C.c_uint64
    L.orc_mix64.argtypes = [C.c_uint64]
    L.orc_mix64.restype = C.c_uint64
    L.orc_stream_key.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_stream_key.restype = C.c_uint64
    L.orc_draw.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_draw.restype = C.c_uint64
    L.orc_gen_er.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, _u32p, _u32p]
    L.orc_gen_rmat.argtypes = [C.c_uint32, C.c_uint64, _u32p, C.c_uint64, _u32p, _u32p]
    L.orc_rmat_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, _u32p]
    L.orc_chung_lu_cdf.argtypes = [C.c_uint32, C.c_double, _u64p]
    L.orc_gen_chung_lu.argtypes = [C.c_uint32, C.c_uint64, _u64p, C.c_uint64, _u32p, _u32p]
    L.orc_gen_communities.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.c_void_p, C.c_void_p]
    L.orc_gen_communities.restype = C.c_uint64
    L.orc_batch_params.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint64)]
    L.orc_mlp_forward.argtypes = [_f32p, _f32p, C.c_int, _f64p]
    L.orc_mlp_loss_grad.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_int, _f64p]
    L.orc_mlp_loss_grad.restype = C.c_double
    _lib = L
    return L


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _take(ptr, n, dtype):
    """Copy n elements from a malloc'ed pointer into numpy and free it."""
    L = lib()
    size = np.dtype(dtype).itemsize * n
    buf = (C.c_char * max(size, 1)).from_address(ptr.value) if ptr.value else None
    out = np.frombuffer(bytes(buf)[:size], dtype=dtype).copy() if buf else np.zeros(0, dtype)
    L.orc_free(ptr)
    return out


# ---------------------------------------------------------------- graphs

def edges_to_csr(n, src, dst):
    L = lib()
    src, dst = _u32(src), _u32(dst)
    po, pn = C.c_void_p(), C.c_void_p()
    L.orc_edges_to_csr(n, len(src), src, dst, C.byref(po), C.byref(pn))
    off = _take(po, n + 1, np.uint64)
    nbr = _take(pn, int(off[-1]), np.uint32)
    return off, nbr


def relabel_by_degree(n, off, nbr):
    L = lib()
    off, nbr = _u64(off), _u32(nbr)
    new2old = np.zeros(n, np.uint32)
    off2 = np.zeros(n + 1, np.uint64)
    nbr2 = np.zeros(max(len(nbr), 1), np.uint32)
    L.orc_relabel_by_degree(n, off, nbr, new2old, off2, nbr2)
    return new2old, off2, nbr2[:len(nbr)]


def orient_lower(n, off, nbr):
    L = lib()
    off, nbr = _u64(off), _u32(nbr)
    off2 = np.zeros(n + 1, np.uint64)
    nbr2 = np.zeros(max(len(nbr) // 2 + 1, 1), np.uint32)
    m = L.orc_orient_lower(n, off, nbr, off2, nbr2.ctypes.data)
    return off2, nbr2[:m].copy()


def load_binary(path):
    L = lib()
    n = C.c_uint32()
    po, pn, pe = C.c_void_p(), C.c_void_p(), C.c_void_p()
    rc = L.orc_load_binary(path.encode(), C.byref(n), C.byref(po), C.byref(pn), C.byref(pe))
    if rc != 0:
        raise IOError(f"orc_load_binary({path}) failed: {rc}")
    off = _take(po, n.value + 1, np.uint64)
    nbr = _take(pn, int(off[-1]), np.uint32)
    ext = _take(pe, n.value, np.uint64)
    return n.value, off, nbr, ext


def write_binary(path, n, off, nbr, ext=None):
    L = lib()
    ext_p = None if ext is None else _u64(ext).ctypes.data
    rc = L.orc_write_binary(path.encode(), n, _u64(off), _u32(nbr), ext_p)
    if rc != 0:
        raise IOError(path)


def brute_triangles(n, off, nbr):
    return int(lib().orc_brute_triangles(n, _u64(off), _u32(nbr)))


def brute_cliques(n, off, nbr, k):
    return int(lib().orc_brute_cliques(n, _u64(off), _u32(nbr), k))


# ---------------------------------------------------------------- layouts

def row_type(data, policy="hybrid"):
    d = _u32(data)
    return lib().orc_row_type(d, len(d), POLICY[policy])


def row_image(data, policy="hybrid"):
    L = lib()
    d = _u32(data)
    out = np.zeros(L.orc_row_image_bound(d, len(d)), np.uint8)
    n = L.orc_row_image(out, d, len(d), POLICY[policy])
    return out[:n].copy()


def layout_build(kind, data):
    L = lib()
    d = _u32(data)
    out = np.zeros(L.orc_row_image_bound(d, len(d)), np.uint8)
    n = getattr(L, f"orc_{kind}_build")(out, d, len(d))
    return out[:n].copy()


def decode(type_, payload, card):
    p = np.ascontiguousarray(payload, dtype=np.uint8)
    if len(p) == 0:
        p = np.zeros(8, np.uint8)
        nbytes = 0
    else:
        nbytes = len(p)
    out = np.zeros(max(card, 1) + 64, np.uint32)
    k = lib().orc_decode(type_, p, card, nbytes, out)
    return out[:k].copy()


def isect_uint(alg, a, b, materialize=False):
    L = lib()
    a, b = _u32(a), _u32(b)
    if materialize:
        out = np.zeros(min(len(a), len(b)) + 8, np.uint32)
        k = L.orc_isect_uint(ALG[alg], a, len(a), b, len(b), out.ctypes.data)
        return out[:k].copy()
    return int(L.orc_isect_uint(ALG[alg], a, len(a), b, len(b), None))


def make_set(values, type_):
    """A typed set over sorted values: returns (OrcSet, keepalive buffer)."""
    v = _u32(values)
    if type_ == UINTEGER:
        buf = v.view(np.uint8).copy()
        nb = buf.nbytes
    else:
        buf = layout_build(_BUILDER[type_], v)
        nb = len(buf)
    buf = np.concatenate([buf, np.zeros(64, np.uint8)])
    s = OrcSet(buf.ctypes.data, len(v), nb, type_, 0)
    return s, buf


def set_intersect(a_vals, a_type, b_vals, b_type, materialize=True):
    """Typed-set intersection through the dispatcher; returns (count, values, result_type)."""
    L = lib()
    A, ka = make_set(a_vals, a_type)
    B, kb = make_set(b_vals, b_type)
    if not materialize:
        return int(L.orc_set_intersect(None, C.byref(A), C.byref(B))), None, None
    n = max(len(a_vals), 1) + max(len(b_vals), 1)
    hi = int(max(np.max(a_vals) if len(a_vals) else 0, np.max(b_vals) if len(b_vals) else 0))
    out = np.zeros(16 * n + 8 * (hi // 64 + 4) + 64, np.uint8)
    Cs = OrcSet(out.ctypes.data, 0, 0, 0, 0)
    c = int(L.orc_set_intersect(C.byref(Cs), C.byref(A), C.byref(B)))
    vals = decode(Cs.type, out[:Cs.number_of_bytes], Cs.cardinality) if c else np.zeros(0, np.uint32)
    return c, vals, Cs.type


def set_intersect_bytes(a_vals, a_type, b_vals, b_type):
    """Materialised typed-set intersection: (count, result type, result bytes)."""
    L = lib()
    A, ka = make_set(a_vals, a_type)
    B, kb = make_set(b_vals, b_type)
    n = max(len(a_vals), 1) + max(len(b_vals), 1)
    hi = int(max(np.max(a_vals) if len(a_vals) else 0, np.max(b_vals) if len(b_vals) else 0))
    out = np.zeros(16 * n + 8 * (hi // 64 + 4) + 64, np.uint8)
    Cs = OrcSet(out.ctypes.data, 0, 0, 0, 0)
    c = int(L.orc_set_intersect(C.byref(Cs), C.byref(A), C.byref(B)))
    return c, int(Cs.type), out[:Cs.number_of_bytes].copy()


def set_payload(values, type_):
    """(type, cardinality, payload bytes) of a typed set in the reference formats."""
    v = _u32(values)
    if type_ == UINTEGER:
        return type_, len(v), v.view(np.uint8).copy()
    if len(v) == 0:
        return type_, 0, np.zeros(0, np.uint8)
    return type_, len(v), layout_build(_BUILDER[type_], v)


class Matrix:
    """SparseMatrix built from a symmetric CSR (lower=True keeps nbr < node)."""

    def __init__(self, n, off, nbr, policy="hybrid", lower=True, threads=8):
        self.n = n
        self._off, self._nbr = _u64(off), _u32(nbr)
        off64 = self._off
        if len(off64) != n + 1 or int(off64[0]) != 0 or int(off64[-1]) > len(self._nbr) or \
                np.any(off64[1:] < off64[:-1]):
            raise ValueError("Matrix: malformed CSR offsets")
        if int(off64[-1]) and int(self._nbr[:int(off64[-1])].max()) >= n:
            raise ValueError("Matrix: neighbour id out of range")
        self.h = lib().orc_matrix_build(n, self._off, self._nbr, int(lower), POLICY[policy], threads)
        if not self.h:
            raise MemoryError("orc_matrix_build could not allocate the row images")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_matrix_free(self.h)
            self.h = None

    def row(self, i):
        s = OrcSet()
        lib().orc_matrix_get_row(self.h, i, C.byref(s))
        nb = s.number_of_bytes
        payload = np.ctypeslib.as_array((C.c_uint8 * max(nb, 1)).from_address(s.data))[:nb].copy() \
            if nb else np.zeros(0, np.uint8)
        return s.type, int(s.cardinality), payload

    def type_counts(self):
        c = np.zeros(3, np.uint64)
        lib().orc_matrix_type_counts(self.h, c)
        return {"bitset": int(c[0]), "pshort": int(c[1]), "uint": int(c[2])}

    def bytes(self):
        return int(lib().orc_matrix_bytes(self.h))

    def cardinality(self):
        return int(lib().orc_matrix_cardinality(self.h))

    def max_row(self):
        return int(lib().orc_matrix_max_row(self.h))

    def triangles(self, threads=8):
        t = C.c_double()
        c = lib().orc_triangle_count(self.h, threads, C.byref(t))
        return int(c), t.value

    def triangles_sample(self, threads, stride, phase=0):
        t = C.c_double()
        e = C.c_uint64()
        c = lib().orc_triangle_count_sample(self.h, threads, stride, phase, C.byref(e), C.byref(t))
        return int(c), int(e.value), t.value

    def cliques(self, k=4, threads=8):
        t = C.c_double()
        c = lib().orc_clique_count(self.h, k, threads, C.byref(t))
        return int(c), t.value

    def cliques_sample(self, k, threads, stride, phase=0):
        t = C.c_double()
        e = C.c_uint64()
        c = lib().orc_clique_count_sample(self.h, k, threads, stride, phase, C.byref(e), C.byref(t))
        return int(c), int(e.value), t.value

    def count_rows(self, k=3, threads=8):
        out = np.zeros(self.n, np.uint64)
        total = lib().orc_count_rows(self.h, k, threads, out)
        return int(total), out

    def pair_mix(self):
        mix = np.zeros(4, np.uint64)
        tot = lib().orc_pair_mix(self.h, mix)
        return int(tot), [int(x) for x in mix]


# ---------------------------------------------------------------- generators

def stream_key(seed, stream):
    return int(lib().orc_stream_key(seed, stream))


def draw(key, ctr):
    return int(lib().orc_draw(key, ctr))


def gen_er(n, m, seed, stream=0):
    src = np.zeros(m, np.uint32)
    dst = np.zeros(m, np.uint32)
    lib().orc_gen_er(n, m, stream_key(seed, stream), src, dst)
    return src, dst


def gen_er_exact(n, m, seed):
    """G(n, m) with exactly m distinct non-loop pairs: the first m distinct pairs of gen_er's
    draw stream (first occurrences by np.unique, ordered by draw index).  Returns (src, dst)
    with src < dst, in first-draw order."""
    assert m <= n * (n - 1) // 2
    K = m + m // 4 + 4096
    while True:
        src, dst = gen_er(n, K, seed)
        lo, hi = np.minimum(src, dst).astype(np.uint64), np.maximum(src, dst).astype(np.uint64)
        keys = (lo << np.uint64(32)) | hi
        keys[src == dst] = np.uint64(0xFFFFFFFFFFFFFFFF)
        uk, first = np.unique(keys, return_index=True)
        first = np.sort(first[uk != np.uint64(0xFFFFFFFFFFFFFFFF)])
        if len(first) >= m:
            first = first[:m]
            return lo[first].astype(np.uint32), hi[first].astype(np.uint32)
        K *= 2


def rmat_thresholds(a=0.57, b=0.19, c=0.19):
    thr = np.zeros(3, np.uint32)
    lib().orc_rmat_thresholds(a, b, c, thr)
    return thr


def gen_rmat(scale, edge_factor, seed, a=0.57, b=0.19, c=0.19):
    m = edge_factor << scale
    src = np.zeros(m, np.uint32)
    dst = np.zeros(m, np.uint32)
    lib().orc_gen_rmat(scale, m, rmat_thresholds(a, b, c), stream_key(seed, 0), src, dst)
    return src, dst


def chung_lu_cdf(n, gamma):
    cdf = np.zeros(n, np.uint64)
    lib().orc_chung_lu_cdf(n, gamma, cdf)
    return cdf


def gen_chung_lu(n, mean_degree, gamma, seed, ncomm=0, csize=256):
    m = n * mean_degree // 2
    cdf = chung_lu_cdf(n, gamma)
    src = np.zeros(m, np.uint32)
    dst = np.zeros(m, np.uint32)
    lib().orc_gen_chung_lu(n, m, cdf, stream_key(seed, 0), src, dst)
    if ncomm:
        L = lib()
        k = L.orc_gen_communities(n, ncomm, csize, seed, None, None)
        cs = np.zeros(k, np.uint32)
        cd = np.zeros(k, np.uint32)
        L.orc_gen_communities(n, ncomm, csize, seed, cs.ctypes.data, cd.ctypes.data)
        src = np.concatenate([src, cs])
        dst = np.concatenate([dst, cd])
    return src, dst


def batch_params(seed, g):
    n = C.c_uint32()
    m = C.c_uint64()
    lib().orc_batch_params(seed, g, C.byref(n), C.byref(m))
    return n.value, m.value


def canonical_oriented(n, src, dst):
    """edges -> symmetric CSR -> degree relabel -> orientation nbr < node."""
    off, nbr = edges_to_csr(n, src, dst)
    new2old, roff, rnbr = relabel_by_degree(n, off, nbr)
    ooff, onbr = orient_lower(n, roff, rnbr)
    return {"sym": (off, nbr), "relabel": (new2old, roff, rnbr), "oriented": (ooff, onbr)}


# ---------------------------------------------------------------- MLP fp64 reference

def mlp_forward(params, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros((x.shape[0], 4), np.float64)
    lib().orc_mlp_forward(np.ascontiguousarray(params, np.float32), x, x.shape[0], out)
    return out


def mlp_loss_grad(params, x, target, mask):
    x = np.ascontiguousarray(x, np.float32)
    grad = np.zeros(len(params), np.float64)
    loss = lib().orc_mlp_loss_grad(np.ascontiguousarray(params, np.float32), x,
                                   np.ascontiguousarray(target, np.float32),
                                   np.ascontiguousarray(mask, np.float32), x.shape[0], grad)
    return float(loss), grad


# ---------------------------------------------------------------- reference programs

ORDERINGS = {"degree": 0, "rev_degree": 1, "strong_run": 2, "bfs": 3, "shingles": 4, "random": 5, "the_game": 6}


def perm(seed, stream, n):
    out = np.zeros(max(n, 1), np.uint32)
    lib().orc_perm(seed, stream, n, out)
    return out[:n]


def order(n, off, nbr, ordering, seed=0):
    """new2old of one of the reference's vertex orderings (MutableGraph::reorder_*)."""
    out = np.zeros(max(n, 1), np.uint32)
    rc = lib().orc_order(n, _u64(off), _u32(nbr), ORDERINGS[ordering], seed, out)
    if rc != 0:
        raise ValueError(ordering)
    return out[:n]


def apply_order(n, off, nbr, new2old):
    """Renumber a symmetric CSR by new2old; rows re-sorted."""
    off = np.asarray(off, np.int64)
    nbr = np.asarray(nbr, np.int64)
    o2n = np.empty(n, np.int64)
    o2n[np.asarray(new2old, np.int64)] = np.arange(n)
    deg = np.diff(off)[np.asarray(new2old, np.int64)]
    noff = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    rows = [np.sort(o2n[nbr[off[v]:off[v + 1]]]) for v in np.asarray(new2old, np.int64)]
    nnbr = np.concatenate(rows).astype(np.uint32) if rows and len(nbr) else np.zeros(0, np.uint32)
    return noff, nnbr


def edge_list_graph(lines):
    """The reference's edge-list loader (src/MutableGraph.hpp:517-598): ids in order of
    first appearance, both directions, rows sorted and deduplicated."""
    ext, rows = {}, []
    for a, b in lines:
        for x in (a, b):
            if x not in ext:
                ext[x] = len(ext)
                rows.append(set())
        rows[ext[a]].add(ext[b])
        rows[ext[b]].add(ext[a])
    n = len(rows)
    off = np.zeros(n + 1, np.uint64)
    nbr = []
    for i, r in enumerate(rows):
        nbr.extend(sorted(r))
        off[i + 1] = len(nbr)
    ids = np.zeros(n, np.uint64)
    for x, i in ext.items():
        ids[i] = x
    return n, off, np.array(nbr, np.uint32), ids


def directed_edge_list_graph(lines):
    """The reference's directed edge-list loader (src/MutableGraph.hpp:893-981): ids in order
    of first appearance (source before destination), out-rows and in-rows sorted and
    deduplicated.  Returns (n, out_rows, in_rows, ext_ids)."""
    ext, outs, ins = {}, [], []
    for a, b in lines:
        for x in (a, b):
            if x not in ext:
                ext[x] = len(ext)
                outs.append(set())
                ins.append(set())
        outs[ext[a]].add(ext[b])
        ins[ext[b]].add(ext[a])
    ids = np.zeros(len(outs), np.uint64)
    for x, i in ext.items():
        ids[i] = x
    rows = lambda side: [np.array(sorted(r), np.uint32) for r in side]
    return len(outs), rows(outs), rows(ins), ids


def load_directed_binary(path_or_bytes):
    """The directed binary format (src/MutableGraph.hpp:636-714): size_t n, then per node
    {u64 external id; size_t out count; out ids; size_t in count; in ids}.
    Returns (n, out_rows, in_rows, ext_ids)."""
    raw = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    pos = 0

    def take(count, dtype):
        nonlocal pos
        arr = np.frombuffer(raw, dtype, count, pos)
        pos += arr.nbytes
        return arr

    n = int(take(1, np.uint64)[0])
    ids, outs, ins = np.zeros(n, np.uint64), [], []
    for i in range(n):
        ids[i] = take(1, np.uint64)[0]
        outs.append(take(int(take(1, np.uint64)[0]), np.uint32).copy())
        ins.append(take(int(take(1, np.uint64)[0]), np.uint32).copy())
    if pos != len(raw):
        raise IOError("trailing bytes in directed binary file")
    return n, outs, ins, ids


def bfs_path(n, off, nbr, start, max_depth=0xEFFFFFF, cap=64, threads=1):
    """n_path's BFS (apps/n_path.cpp:41-85) over rows [0, n): returns
    {path_length, rounds, level_sizes, reached}.  threads > 1: each round's union over that many
    threads (apps/n_path.cpp:48-77), same results."""
    rounds, reached = C.c_uint32(), C.c_uint64()
    sizes = np.zeros(cap, np.uint64)
    if threads > 1:
        p = lib().orc_bfs_path_mt(n, _u64(off), _u32(nbr), start, max_depth, C.addressof(rounds),
                                  sizes.ctypes.data, cap, C.addressof(reached), threads)
    else:
        p = lib().orc_bfs_path(n, _u64(off), _u32(nbr), start, max_depth, C.addressof(rounds), sizes.ctypes.data,
                               cap, C.addressof(reached))
    return {"path_length": int(p), "rounds": rounds.value, "level_sizes": sizes[:min(rounds.value, cap)].tolist(),
            "reached": reached.value}


def bitset_difference(a_bytes, b_bytes):
    """set_difference of two bitset images (difference.hpp:5-84) as a sorted uint32 array."""
    a = np.ascontiguousarray(a_bytes, np.uint8)
    b = np.ascontiguousarray(b_bytes, np.uint8)
    out = np.zeros(max(8 * len(a), 8), np.uint32)
    k = lib().orc_bitset_difference(a if len(a) else np.zeros(8, np.uint8), len(a),
                                    b if len(b) else np.zeros(8, np.uint8), len(b), out)
    return out[:k].copy()


def run_ref_converter(lines, outdir):
    """Run the reference's edge-list converter built in _ref/; returns {name: (n, off, nbr, ext)}."""
    exe = os.path.join(REF_DIR, "undirectedEdgeListToBinary")
    src = os.path.join(outdir, "edges.txt")
    with open(src, "w") as f:
        for a, b in lines:
            f.write(f"{a} {b}\n")
    subprocess.run([exe, src, outdir], capture_output=True, text=True, timeout=600, check=True, cwd=outdir)
    out = {}
    for name in ORDERINGS:
        path = os.path.join(outdir, f"u_{name}.bin")
        if os.path.exists(path):
            out[name] = load_binary(path)
    return out


def ref_available():
    return os.path.exists(os.path.join(REF_DIR, "undirected_triangle_counting"))


def run_ref_stats(app, n, off, nbr, threads=4, layout="hybrid", k=None, timeout=600):
    """Run a reference program built in _ref/ on a symmetric CSR (already
    degree-relabelled); returns {count, seconds (counting phase), row_bytes}."""
    exe = os.path.join(REF_DIR, app)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "g.bin")
        write_binary(path, n, off, nbr)
        cmd = [exe, f"--graph={path}", "--input_type=binary", f"--t={threads}", f"--layout={layout}"]
        if k is not None:
            cmd.append(f"--n={k}")
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=td).stdout
    cnt = re.search(r"Number of (?:triangles|cliques): (\d+)", out) or re.search(r"path length: (\d+)", out)
    tm = re.findall(r"Time\[(?:UNDIRECTED TRIANGLE COUNTING|n_cliques)\]: ([0-9.e+-]+) s", out)
    rb = re.search(r"ROW DATA SIZE \(Bytes\): (\d+)", out)
    if not cnt:
        raise RuntimeError(f"reference run failed: {out[-2000:]}")
    return {"count": int(cnt.group(1)), "seconds": float(tm[-1]) if tm else float("nan"),
            "row_bytes": int(rb.group(1)) if rb else None}


def run_ref(app, n, off, nbr, threads=4, layout="hybrid", k=None, timeout=600):
    """(count, seconds of the counting phase) of a reference program; see run_ref_stats."""
    r = run_ref_stats(app, n, off, nbr, threads, layout, k, timeout)
    return r["count"], r["seconds"]


def run_ref_blocks(pairs, workdir):
    """Run _ref/ref_blocks (the reference's block-bitset builders and intersections) over
    (a, b) pairs of sorted distinct uint32 arrays; per pair returns a list of seven
    (card, type, bytes): bitset_new(a), bitset_new(b), new_type(a), new_type(b),
    bn(a) & bn(b), uint(a) & bn(b), nt(a) & nt(b)."""
    src, dst = os.path.join(workdir, "pairs.bin"), os.path.join(workdir, "blocks.bin")
    with open(src, "wb") as f:
        for a, b in pairs:
            for v in (a, b):
                v = _u32(v)
                f.write(np.uint32(len(v)).tobytes() + v.tobytes())
    subprocess.run([os.path.join(REF_DIR, "ref_blocks"), src, dst], check=True, capture_output=True, timeout=600)
    raw = open(dst, "rb").read()
    pos, out = 0, []
    for _ in pairs:
        rec = []
        for _ in range(7):
            card, typ, nb = np.frombuffer(raw, np.uint64, 3, pos).tolist()
            pos += 24
            rec.append((card, typ, np.frombuffer(raw, np.uint8, nb, pos).copy()))
            pos += nb
        out.append(rec)
    return out
