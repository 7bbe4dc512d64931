"""MI355X spatial-join-count-over-shells engine: Python host mirror.

Thin ctypes bindings over the C-ABI in ``include/sjcs.h`` (library
``lib/libsjcs.so``, built in-tree by ``make -C spatialjoincountovershells_amd``).
The names follow the reference's interface for this path
(src/astr_rcount/main.c, include/objects/ArraySTR/ArraySTR.h):

* :class:`Engine` -- the JoinContext of ``astr_rcount``: shell thresholds
  (``mrcc->radiuses`` squared, main.c:645-649), the halo set (main.c:635-642)
  and u64 counts that accumulate over particle chunks (main.c:246-262, 651).
* :class:`ArraySTR` -- the index lifecycle ``constructArraySTR`` /
  ``insertDataArraySTR`` / ``makeTree*ArraySTR`` /
  ``multiRangeCountWithPriodicBoundArraySTR`` / ``destructArraySTR``.
* :class:`RTree` -- ``rt_rcount``'s index: ``constructRTree`` /
  ``insertDataRTree`` / ``multiRangeCountWithPriodicBoundRTree`` /
  ``destructRTree``.
* :func:`shell_thresholds` -- the CLI's ``--radius min:max:K`` to squared
  FP32 thresholds (include/support_functions.h:18-34, main.c:506-533).

There is no CPU fallback: without the built library or without a HIP
device every computing call raises :class:`SJCSError`.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsjcs.so")
HOST_LIB_PATH = os.path.join(_HERE, "lib", "libsjcs_host.so")
CLI_PATH = os.path.join(_HERE, "bin", "sjcs_rcount")

SJCS_OK = 0
SJCS_ERR_INVALID = -1
SJCS_ERR_DEVICE = -2
SJCS_ERR_NOMEM = -3
SJCS_ERR_STATE = -4

# Every symbol include/sjcs.h declares.
EXPORTED = (
    "sjcs_engine_create", "sjcs_engine_destroy", "sjcs_last_error",
    "sjcs_set_shells", "sjcs_set_halos", "sjcs_set_halos_device",
    "sjcs_reset_counts", "sjcs_add_particles", "sjcs_add_particles_device",
    "sjcs_sync", "sjcs_get_counts", "sjcs_last_chunk_pair_tests",
    "sjcs_set_option", "sjcs_get_option", "sjcs_autotune", "sjcs_last_chunk_timings",
    "sjcs_phase_timings",
    "sjcs_index_create", "sjcs_index_insert", "sjcs_index_build",
    "sjcs_index_count", "sjcs_index_destroy", "sjcs_index_count_noalloc",
    "sjcs_device_alloc", "sjcs_device_free", "sjcs_copy_to_device", "sjcs_copy_to_host",
    "sjcs_host_alloc", "sjcs_host_free", "sjcs_get_counts_range", "sjcs_get_counts_range_u32",
    "sjcs_gadget_decode_params", "sjcs_add_compressed_particles",
    "sjcs_add_compressed_particles_device",
    "sjcs_decode_compressed_device",
    "sjcs_server_create", "sjcs_server_submit", "sjcs_server_test", "sjcs_server_wait",
    "sjcs_server_times", "sjcs_server_release", "sjcs_server_error", "sjcs_server_destroy",
    "sjcs_rtree_create", "sjcs_rtree_insert", "sjcs_rtree_shape", "sjcs_rtree_leaves",
    "sjcs_rtree_count", "sjcs_rtree_count_device", "sjcs_rtree_error", "sjcs_rtree_destroy",
)


class SJCSError(RuntimeError):
    """A C-ABI call returned an error code."""

    def __init__(self, code, msg):
        super().__init__(f"sjcs error {code}: {msg}")
        self.code = code


class GadgetDecode(ctypes.Structure):
    """sjcs_gadget_decode: what decoding one compressed particle file needs
    (GadgetDecodeParams + GadId, include/Gadget/GadgetHeader.h:30-37,
    include/Gadget/GadId.h:5-14)."""
    _fields_ = [("id_start", ctypes.c_int64), ("ng1d", ctypes.c_int64),
                ("nt1d", ctypes.c_int32), ("fused", ctypes.c_int32),
                ("cell_interval", ctypes.c_float), ("disp_min", ctypes.c_float),
                ("unitsep", ctypes.c_double)]


GADGET_HEADER_BYTES = 256


def read_compressed(path: str):
    """(header bytes, uint16 triplets [n][3]) of a compressed particle file:
    4-byte marker, 256-byte header, 8 bytes of markers, npart[1] triplets
    (src/Gadget/BufVector3D.c:111-136)."""
    with open(path, "rb") as f:
        f.seek(4)
        hdr = f.read(GADGET_HEADER_BYTES)
        if len(hdr) != GADGET_HEADER_BYTES:
            raise ValueError(f"{path}: shorter than a Gadget header")
        n = int(np.frombuffer(hdr, "<i4", 1, 4)[0])
        if n < 0:
            raise ValueError(f"{path}: npart[1] = {n}")
        f.seek(8, os.SEEK_CUR)
        uv = np.frombuffer(f.read(6 * n), "<u2")
        if uv.size != 3 * n:
            raise ValueError(f"{path}: shorter than its header says")
    return hdr, uv.reshape(n, 3)


def gadget_header(npart: int, npart_total: int, box: float = 1000.0, disp_min: float = -1.0,
                  disp_max: float = 1.0, pos_bits: int = 16, id_start: int = 1) -> bytes:
    """A 256-byte Gadget header (include/Gadget/GadgetHeader.h:4-27) with the
    fields the decode reads: npart[1], npartTotal[1] (low word, an int) and
    npartTotal[2] (high word), BoxSize, disp_min, disp_max, pos_bits,
    id_start, id_end."""
    import struct
    h = bytearray(GADGET_HEADER_BYTES)
    struct.pack_into("<6i", h, 0, 0, npart, 0, 0, 0, 0)
    lo = npart_total & 0xFFFFFFFF
    lo = lo - (1 << 32) if lo >= (1 << 31) else lo
    struct.pack_into("<6i", h, 96, 0, lo, npart_total >> 32, 0, 0, 0)
    struct.pack_into("<i", h, 124, 1)
    struct.pack_into("<d", h, 128, box)
    struct.pack_into("<ffi", h, 172, disp_min, disp_max, pos_bits)
    struct.pack_into("<qq", h, 184, id_start, id_start + npart - 1)
    return bytes(h)


def write_compressed(path: str, header: bytes, uv) -> None:
    """A compressed particle file: record markers around the header and
    around the uint16 triplets (read back by read_compressed and by the
    reference's readCompressedParticles)."""
    import struct
    uv = np.ascontiguousarray(uv, dtype="<u2").reshape(-1, 3)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", GADGET_HEADER_BYTES) + bytes(header) +
                struct.pack("<i", GADGET_HEADER_BYTES) + struct.pack("<i", uv.nbytes & 0x7FFFFFFF))
        uv.tofile(f)
        f.write(struct.pack("<i", uv.nbytes & 0x7FFFFFFF))


def gadget_decode_params(header: bytes, ntile: int = 1) -> GadgetDecode:
    """sjcs_gadget_decode_params of a 256-byte header (host only)."""
    if len(header) != GADGET_HEADER_BYTES:
        raise ValueError("a Gadget header is 256 bytes")
    lib = load_library()
    buf = ctypes.create_string_buffer(bytes(header), GADGET_HEADER_BYTES)
    out = GadgetDecode()
    rc = lib.sjcs_gadget_decode_params(buf, ntile, ctypes.byref(out))
    if rc != SJCS_OK:
        raise SJCSError(rc, "bad Gadget header (negative particle total)")
    return out


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads lib/libsjcs.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise SJCSError(SJCS_ERR_STATE,
                        f"{path} is not built; run `make -C {_HERE}` first")
    lib = ctypes.CDLL(path)
    vp, u64, i32, sz = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "sjcs_engine_create": (i32, [i32, pp]),
        "sjcs_engine_destroy": (None, [vp]),
        "sjcs_last_error": (ctypes.c_char_p, [vp]),
        "sjcs_set_shells": (i32, [vp, vp, i32]),
        "sjcs_set_halos": (i32, [vp, vp, sz, u64]),
        "sjcs_set_halos_device": (i32, [vp, vp, u64]),
        "sjcs_reset_counts": (i32, [vp]),
        "sjcs_add_particles": (i32, [vp, vp, u64]),
        "sjcs_add_particles_device": (i32, [vp, vp, u64]),
        "sjcs_sync": (i32, [vp]),
        "sjcs_get_counts": (i32, [vp, vp]),
        "sjcs_last_chunk_pair_tests": (i32, [vp, ctypes.POINTER(u64)]),
        "sjcs_set_option": (i32, [vp, ctypes.c_char_p, ctypes.c_int64]),
        "sjcs_last_chunk_timings": (i32, [vp, ctypes.POINTER(ctypes.c_float)]),
        "sjcs_phase_timings": (i32, [vp, ctypes.POINTER(ctypes.c_float)]),
        "sjcs_get_option": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
        "sjcs_server_create": (i32, [vp, i32, u64, pp]),
        "sjcs_server_submit": (i32, [vp, vp, sz, u64, vp, i32, vp, ctypes.POINTER(u64)]),
        "sjcs_server_test": (i32, [vp, u64]),
        "sjcs_server_wait": (i32, [vp, u64]),
        "sjcs_server_times": (i32, [vp, u64, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
        "sjcs_server_release": (i32, [vp, u64]),
        "sjcs_server_error": (ctypes.c_char_p, [vp]),
        "sjcs_server_destroy": (None, [vp]),
        "sjcs_autotune": (i32, [vp, vp, u64, ctypes.POINTER(ctypes.c_float)]),
        "sjcs_index_create": (i32, [vp, u64, pp]),
        "sjcs_index_insert": (i32, [vp, vp, u64]),
        "sjcs_index_build": (i32, [vp]),
        "sjcs_index_count": (i32, [vp, vp, sz, u64, vp, i32, vp]),
        "sjcs_index_destroy": (None, [vp]),
        "sjcs_index_count_noalloc": (i32, [vp, vp, u64, vp, i32, vp]),
        "sjcs_device_alloc": (i32, [vp, sz, pp]),
        "sjcs_device_free": (i32, [vp, vp]),
        "sjcs_copy_to_device": (i32, [vp, vp, vp, sz]),
        "sjcs_copy_to_host": (i32, [vp, vp, vp, sz]),
        "sjcs_host_alloc": (i32, [sz, pp]),
        "sjcs_host_free": (None, [vp]),
        "sjcs_get_counts_range": (i32, [vp, u64, u64, vp]),
        "sjcs_get_counts_range_u32": (i32, [vp, u64, u64, vp]),
        "sjcs_gadget_decode_params": (i32, [vp, i32, ctypes.POINTER(GadgetDecode)]),
        "sjcs_add_compressed_particles": (i32, [vp, vp, u64, ctypes.POINTER(GadgetDecode)]),
        "sjcs_add_compressed_particles_device": (i32, [vp, vp, u64,
                                                       ctypes.POINTER(GadgetDecode)]),
        "sjcs_decode_compressed_device": (i32, [vp, vp, u64, ctypes.POINTER(GadgetDecode), vp]),
        "sjcs_rtree_create": (i32, [vp, pp]),
        "sjcs_rtree_insert": (i32, [vp, vp, u64]),
        "sjcs_rtree_shape": (i32, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                   ctypes.POINTER(u64), ctypes.POINTER(i32)]),
        "sjcs_rtree_leaves": (i32, [vp, vp, vp]),
        "sjcs_rtree_count": (i32, [vp, vp, sz, u64, vp, i32, vp]),
        "sjcs_rtree_count_device": (i32, [vp, vp, u64, vp, i32, vp]),
        "sjcs_rtree_error": (ctypes.c_char_p, [vp]),
        "sjcs_rtree_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _points(a, what: str) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"{what} must have shape (n, 3)")
    return a


def _shells(r2) -> np.ndarray:
    r2 = np.ascontiguousarray(r2, dtype=np.float32).reshape(-1)
    if r2.size == 0:
        raise ValueError("at least one shell")
    return r2


class Engine:
    """Device-resident halos, shells and u64 counts; one particle chunk at a
    time (the reference's per-file loop, src/astr_rcount/main.c:246-262)."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        rc = self._lib.sjcs_engine_create(device, ctypes.byref(h))
        if rc != SJCS_OK:
            raise SJCSError(rc, self._lib.sjcs_last_error(None).decode())
        self._h = h
        self.nrads = 0
        self.nhalos = 0

    def _check(self, rc):
        if rc != SJCS_OK:
            raise SJCSError(rc, self._lib.sjcs_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sjcs_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_shells(self, radiuses_sq):
        r2 = _shells(radiuses_sq)
        self._check(self._lib.sjcs_set_shells(self._h, _ptr(r2), r2.size))
        self.nrads = r2.size

    def set_halos(self, points):
        p = _points(points, "halo points")
        self._check(self._lib.sjcs_set_halos(self._h, _ptr(p), 12, p.shape[0]))
        self.nhalos = p.shape[0]

    def set_halos_device(self, ptr: int, n: int):
        self._check(self._lib.sjcs_set_halos_device(self._h, ctypes.c_void_p(ptr), n))
        self.nhalos = n

    def reset_counts(self):
        self._check(self._lib.sjcs_reset_counts(self._h))

    def add_particles(self, xyz):
        p = _points(xyz, "particles")
        self._check(self._lib.sjcs_add_particles(self._h, _ptr(p), p.shape[0]))

    def add_particles_device(self, ptr: int, n: int):
        self._check(self._lib.sjcs_add_particles_device(self._h, ctypes.c_void_p(ptr), n))

    def add_compressed_particles(self, uv, params: GadgetDecode):
        """One compressed particle file's triplets as a chunk, decoded on
        the device (sjcs_add_compressed_particles)."""
        uv = np.ascontiguousarray(uv, dtype=np.uint16).reshape(-1, 3)
        self._check(self._lib.sjcs_add_compressed_particles(self._h, _ptr(uv), uv.shape[0],
                                                            ctypes.byref(params)))

    def add_compressed_particles_device(self, ptr: int, n: int, params: GadgetDecode):
        self._check(self._lib.sjcs_add_compressed_particles_device(
            self._h, ctypes.c_void_p(ptr), n, ctypes.byref(params)))

    def decode_compressed(self, uv, params: GadgetDecode) -> np.ndarray:
        """The device decode alone: triplets in, packed xyz out."""
        uv = np.ascontiguousarray(uv, dtype=np.uint16).reshape(-1, 3)
        n = uv.shape[0]
        out = np.zeros((n, 3), np.float32)
        d_uv = self.device_array(uv if n else np.zeros(3, np.uint16))
        d_xyz = self.device_array(out if n else np.zeros(3, np.float32))
        try:
            self._check(self._lib.sjcs_decode_compressed_device(
                self._h, ctypes.c_void_p(d_uv), n, ctypes.byref(params), ctypes.c_void_p(d_xyz)))
            self._check(self._lib.sjcs_copy_to_host(self._h, _ptr(out), ctypes.c_void_p(d_xyz),
                                                    out.nbytes))
        finally:
            self.device_free(d_uv)
            self.device_free(d_xyz)
        return out

    def sync(self):
        self._check(self._lib.sjcs_sync(self._h))

    def counts(self) -> np.ndarray:
        out = np.zeros((self.nhalos, self.nrads), dtype=np.uint64)
        self._check(self._lib.sjcs_get_counts(self._h, _ptr(out)))
        return out

    def counts_range(self, first: int, n: int) -> np.ndarray:
        """Rows [first, first + n) of the counts."""
        out = np.zeros((n, self.nrads), dtype=np.uint64)
        self._check(self._lib.sjcs_get_counts_range(self._h, first, n, _ptr(out)))
        return out

    def counts_range_u32(self, first: int, n: int) -> np.ndarray:
        """Rows [first, first + n) of the counts as uint32; SJCSError
        (SJCS_ERR_STATE) once the totals are held in 64 bits."""
        out = np.zeros((n, self.nrads), dtype=np.uint32)
        self._check(self._lib.sjcs_get_counts_range_u32(self._h, first, n, _ptr(out)))
        return out

    def counts_fit_u32(self) -> bool:
        """Whether counts_range_u32 can serve the current totals."""
        return self._lib.sjcs_get_counts_range_u32(self._h, 0, 0, None) == SJCS_OK

    def last_chunk_pair_tests(self) -> int:
        v = ctypes.c_uint64()
        self._check(self._lib.sjcs_last_chunk_pair_tests(self._h, ctypes.byref(v)))
        return v.value

    def last_chunk_timings(self):
        ms = (ctypes.c_float * 3)()
        self._check(self._lib.sjcs_last_chunk_timings(self._h, ms))
        return tuple(ms)

    def phase_timings(self) -> dict:
        """Device milliseconds: the last chunk's upload, index build and
        count, the last halo preparation and the last read-out."""
        ms = (ctypes.c_float * 5)()
        self._check(self._lib.sjcs_phase_timings(self._h, ms))
        return dict(zip(("upload", "index_build", "count", "halo_sort", "readout"), ms))

    def device_array(self, host: np.ndarray) -> int:
        """Copies a host array into new device memory; returns the device
        address (free it with :meth:`device_free`)."""
        host = np.ascontiguousarray(host)
        ptr = ctypes.c_void_p()
        self._check(self._lib.sjcs_device_alloc(self._h, host.nbytes, ctypes.byref(ptr)))
        self._check(self._lib.sjcs_copy_to_device(self._h, ptr, _ptr(host), host.nbytes))
        return ptr.value

    def device_free(self, ptr: int):
        self._check(self._lib.sjcs_device_free(self._h, ctypes.c_void_p(ptr)))

    def set_option(self, key: str, value: int):
        self._check(self._lib.sjcs_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key: str) -> int:
        v = ctypes.c_int64()
        self._check(self._lib.sjcs_get_option(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, d_xyz: int, n: int) -> float:
        """Times candidate settings on a device-resident chunk and keeps the
        fastest (see sjcs_autotune); returns its sort + count milliseconds."""
        ms = ctypes.c_float()
        self._check(self._lib.sjcs_autotune(self._h, ctypes.c_void_p(d_xyz), n, ctypes.byref(ms)))
        return ms.value


class ArraySTR:
    """The reference index lifecycle on the device (include/objects/ArraySTR/
    ArraySTR.h:52-58): construct(n) -> insert(points) -> make_tree() ->
    multi_range_count(centrals, radiuses_sq) -> close()."""

    def __init__(self, engine: Engine, ndata: int):
        self._e = engine
        self._lib = engine._lib
        h = ctypes.c_void_p()
        engine._check(self._lib.sjcs_index_create(engine._h, ndata, ctypes.byref(h)))
        self._h = h

    def insert(self, xyz):
        p = _points(xyz, "particles")
        self._e._check(self._lib.sjcs_index_insert(self._h, _ptr(p), p.shape[0]))

    def make_tree(self):
        self._e._check(self._lib.sjcs_index_build(self._h))

    def multi_range_count(self, centrals, radiuses_sq, counts=None) -> np.ndarray:
        c = _points(centrals, "centrals")
        r2 = _shells(radiuses_sq)
        if counts is None:
            counts = np.zeros((c.shape[0], r2.size), dtype=np.uint64)
        if counts.dtype != np.uint64 or not counts.flags.c_contiguous or \
                counts.shape != (c.shape[0], r2.size):
            raise ValueError("counts must be C-contiguous uint64 (n, nrads)")
        self._e._check(self._lib.sjcs_index_count(self._h, _ptr(c), 12, c.shape[0], _ptr(r2),
                                                  r2.size, _ptr(counts)))
        return counts

    def multi_range_count_noalloc(self, d_centrals: int, n: int, radiuses_sq, d_counts: int):
        """sjcs_index_count_noalloc: centres and u64 counts [n][K] already on
        the device (counts +=); the library allocates nothing for it."""
        r2 = _shells(radiuses_sq)
        self._e._check(self._lib.sjcs_index_count_noalloc(
            self._h, ctypes.c_void_p(d_centrals), n, _ptr(r2), r2.size,
            ctypes.c_void_p(d_counts)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sjcs_index_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class RTree:
    """rt_rcount's index (src/objects/RTree/RTree.c, Build.c): a Guttman
    R-tree built on the host by inserting points one at a time in order --
    the order decides which particles share a leaf, and rt_rcount's count
    depends on that -- and counted on the GPU with rt_rcount's leaf walk
    (src/objects/RTree/MultiRangeCount.c:13-114).  ``engine`` None gives a
    host-only tree: insert and inspect, no count."""

    def __init__(self, engine: "Engine | None" = None):
        self._e = engine
        self._lib = engine._lib if engine is not None else load_library()
        h = ctypes.c_void_p()
        rc = self._lib.sjcs_rtree_create(engine._h if engine is not None else None,
                                         ctypes.byref(h))
        if rc != SJCS_OK:
            raise SJCSError(rc, "sjcs_rtree_create")
        self._h = h

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise SJCSError(rc, self._lib.sjcs_rtree_error(self._h).decode())
        return rc

    def insert(self, xyz):
        """insertDataRTree for each point, in order (ids continue)."""
        p = _points(xyz, "particles")
        self._check(self._lib.sjcs_rtree_insert(self._h, _ptr(p), p.shape[0]))

    def shape(self) -> dict:
        n, nl, nn, ht = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        self._check(self._lib.sjcs_rtree_shape(self._h, ctypes.byref(n), ctypes.byref(nl),
                                               ctypes.byref(nn), ctypes.byref(ht)))
        return {"points": n.value, "leaves": nl.value, "nodes": nn.value, "height": ht.value}

    def leaves(self):
        """(ids leaf by leaf, leaf sizes), leaves depth first."""
        sh = self.shape()
        ids = np.zeros(max(sh["points"], 1), np.uint64)
        lens = np.zeros(max(sh["leaves"], 1), np.uint32)
        self._check(self._lib.sjcs_rtree_leaves(self._h, _ptr(ids), _ptr(lens)))
        return ids[:sh["points"]], lens[:sh["leaves"]]

    def multi_range_count(self, centrals, radiuses_sq, counts=None) -> np.ndarray:
        """threadJoin + multiRangeCountWithPriodicBoundRTree over every
        centre: counts [n][nrads] (uint64, added to when given)."""
        c = _points(centrals, "centrals")
        r2 = _shells(radiuses_sq)
        if counts is None:
            counts = np.zeros((c.shape[0], r2.size), dtype=np.uint64)
        if counts.dtype != np.uint64 or not counts.flags.c_contiguous or \
                counts.shape != (c.shape[0], r2.size):
            raise ValueError("counts must be C-contiguous uint64 (n, nrads)")
        self._check(self._lib.sjcs_rtree_count(self._h, _ptr(c), 12, c.shape[0], _ptr(r2),
                                               r2.size, _ptr(counts)))
        return counts

    def multi_range_count_device(self, d_centrals: int, n: int, radiuses_sq, d_counts: int):
        """sjcs_rtree_count_device: centres (packed xyz) and u64 counts
        [n][K] already on the device (counts +=)."""
        r2 = _shells(radiuses_sq)
        self._check(self._lib.sjcs_rtree_count_device(self._h, ctypes.c_void_p(d_centrals), n,
                                                      _ptr(r2), r2.size,
                                                      ctypes.c_void_p(d_counts)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sjcs_rtree_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


class Server:
    """Serving loop over a built index (sjcs_server_*): submit halo batches
    with their own shells, wait for their counts."""

    def __init__(self, index: ArraySTR, slots: int = 0, mem_budget: int = 0):
        self._ix = index
        self._lib = index._lib
        h = ctypes.c_void_p()
        index._e._check(self._lib.sjcs_server_create(index._h, slots, mem_budget, ctypes.byref(h)))
        self._h = h
        self._live = {}  # ticket -> (centres, thresholds, counts) kept alive

    def _check(self, rc):
        if rc < 0:
            raise SJCSError(rc, self._lib.sjcs_server_error(self._h).decode())
        return rc

    def submit(self, centrals, radiuses_sq) -> int:
        c = _points(centrals, "centrals")
        r2 = _shells(radiuses_sq)
        counts = np.zeros((c.shape[0], r2.size), dtype=np.uint64)
        t = ctypes.c_uint64()
        self._check(self._lib.sjcs_server_submit(self._h, _ptr(c), 12, c.shape[0], _ptr(r2),
                                                 r2.size, _ptr(counts), ctypes.byref(t)))
        self._live[t.value] = (c, r2, counts)
        return t.value

    def done(self, ticket: int) -> bool:
        return self._check(self._lib.sjcs_server_test(self._h, ticket)) == 1

    def wait(self, ticket: int) -> np.ndarray:
        self._check(self._lib.sjcs_server_wait(self._h, ticket))
        return self._live[ticket][2]

    def times(self, ticket: int):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.sjcs_server_times(self._h, ticket, ctypes.byref(a), ctypes.byref(b),
                                                ctypes.byref(c)))
        return a.value, b.value, c.value

    def release(self, ticket: int):
        self._check(self._lib.sjcs_server_release(self._h, ticket))
        self._live.pop(ticket, None)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sjcs_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        self.close()


def shell_thresholds(minrad: float, maxrad: float, nrads: int):
    """(radiuses, radiuses_sq) as the reference computes them for
    ``--radius minrad:maxrad:nrads`` after its min/max swap
    (include/support_functions.h:18-34 with the exponent a + i*c evaluated
    as fma(i, c, a), as gcc -O2 compiles it on an FMA host; squares in FP32,
    src/astr_rcount/main.c:648-649).  Computed by the C host library."""
    host = load_host_library()
    r = np.zeros(nrads, dtype=np.float32)
    r2 = np.zeros(nrads, dtype=np.float32)
    host.sjcs_host_shell_thresholds(ctypes.c_double(minrad), ctypes.c_double(maxrad), nrads,
                                    _ptr(r), _ptr(r2))
    return r, r2


_host = None


def load_host_library(path: str = HOST_LIB_PATH) -> ctypes.CDLL:
    """The C host library (CLI parsing, file formats, thresholds, output)."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(path):
        raise SJCSError(SJCS_ERR_STATE,
                        f"{path} is not built; run `make -C {_HERE}` first")
    lib = ctypes.CDLL(path)
    lib.sjcs_host_shell_thresholds.restype = None
    lib.sjcs_host_shell_thresholds.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               ctypes.c_void_p, ctypes.c_void_p]
    _host = lib
    return lib
