"""Python bindings for the mdmc C ABI (include/mdmc.h).

ctypes over ``framework_amd/lib/libmdmc.so``; torch is used only for device
memory and streams.  Every call goes to the HIP library: if it is not built,
``lib()`` raises instead of falling back to anything else.

Reference interface mirrored (DMC_CUDA/DMC_CUDA/include/dmc.hpp:40-57):
  setup_device_luts()   -> setup()
  cleanup_device_luts() -> cleanup()
  run_dmc(vertices, triangles, grid, xyz_min, xyz_max, iso, num_smooth)
                        -> run_dmc(grid, xyz_min, xyz_max, iso, num_smooth)
                           returning (vertices (V,3) float32, triangles (2Q,3) uint32)
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmdmc.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "mdmc.h")

MDMC_OK = 0
STATUS_NAMES = {1: "vertex overflow", 2: "quad overflow", 4: "ring lookup failed", 8: "scan timeout",
                16: "internal inconsistency", 32: "more than 2^32-1 cells, vertices or quads"}

_lib = None


class MDMCError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = ""
        if _lib is not None:
            msg = _lib.mdmc_last_error().decode(errors="replace")
            name = _lib.mdmc_status_string(code).decode()
        else:
            name = str(code)
        super().__init__(f"{where}: {name} ({code}) {msg}".strip())
        self.code = code


class Counts(ctypes.Structure):
    _fields_ = [("num_active", ctypes.c_uint32), ("num_vertices", ctypes.c_uint32),
                ("num_quads", ctypes.c_uint32), ("status", ctypes.c_uint32)]


class Mesh(ctypes.Structure):
    _fields_ = [("vertices", ctypes.POINTER(ctypes.c_float)), ("quads", ctypes.POINTER(ctypes.c_int32)),
                ("num_vertices", ctypes.c_uint32), ("num_quads", ctypes.c_uint32),
                ("num_active", ctypes.c_uint32)]


class DeviceMesh(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("quads", ctypes.c_void_p),
                ("num_vertices", ctypes.c_uint32), ("num_quads", ctypes.c_uint32),
                ("num_active", ctypes.c_uint32)]


_U3 = ctypes.c_uint32 * 3
_F3 = ctypes.c_float * 3
_VP = ctypes.c_void_p


class Volume(ctypes.Structure):
    """mdmc_volume: one volume of a batch (device pointers)."""
    _fields_ = [("field", _VP), ("dims", ctypes.c_uint32 * 3), ("iso", ctypes.c_float),
                ("xyz_min", ctypes.c_float * 3), ("xyz_max", ctypes.c_float * 3),
                ("vertices", _VP), ("vertex_capacity", ctypes.c_uint64),
                ("quads", _VP), ("quad_capacity", ctypes.c_uint64)]


# the same layout as a numpy record, so that a batch table is filled with a
# few array assignments instead of per-field ctypes writes
_VOLUME_DTYPE = np.dtype({
    "names": ["field", "dims", "iso", "xyz_min", "xyz_max", "vertices", "vertex_capacity", "quads", "quad_capacity"],
    "formats": ["<u8", ("<u4", (3,)), "<f4", ("<f4", (3,)), ("<f4", (3,)), "<u8", "<u8", "<u8", "<u8"],
    "offsets": [getattr(Volume, n).offset for n, _ in Volume._fields_],
    "itemsize": ctypes.sizeof(Volume)})


class GenSpec(ctypes.Structure):
    """mdmc_gen_spec: one synthetic field of a batched generation."""
    _fields_ = [("out", _VP), ("dims", ctypes.c_uint32 * 3), ("kind", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("nblobs", ctypes.c_uint32), ("sigma_lo", ctypes.c_double), ("sigma_hi", ctypes.c_double),
                ("amp_lo", ctypes.c_double), ("amp_hi", ctypes.c_double), ("period", ctypes.c_double),
                ("noise_amp", ctypes.c_double)]


MAX_BATCH = 1024
OPT_BATCH_PLANES, OPT_CLASSIFY_BLOCKS_PER_CU, OPT_EMIT_BLOCKS_PER_CU, OPT_PT_PLANES, OPT_PT_BAND, OPT_PT_ROWS = \
    0, 1, 2, 3, 4, 5


def declared_functions() -> list[str]:
    """Names of the functions include/mdmc.h declares."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdmc_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libmdmc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with framework_amd.build.build()")
    # torch first: it brings its own HIP runtime, which the library then
    # shares; the other order leaves two runtimes in the process
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    L.mdmc_status_string.restype = ctypes.c_char_p
    L.mdmc_status_string.argtypes = [ctypes.c_int]
    L.mdmc_last_error.restype = ctypes.c_char_p
    L.mdmc_setup.restype = ctypes.c_int
    L.mdmc_cleanup.restype = ctypes.c_int
    L.mdmc_run.argtypes = [_VP, _U3, _F3, _F3, ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(Mesh)]
    L.mdmc_run.restype = ctypes.c_int
    L.mdmc_mesh_free.argtypes = [ctypes.POINTER(Mesh)]
    L.mdmc_run_device.argtypes = [_VP, _U3, _F3, _F3, ctypes.c_float, ctypes.c_uint32, _VP,
                                  ctypes.POINTER(DeviceMesh)]
    L.mdmc_run_device.restype = ctypes.c_int
    L.mdmc_device_mesh_free.argtypes = [ctypes.POINTER(DeviceMesh)]
    L.mdmc_device_mesh_free.restype = None
    L.mdmc_ctx_create.argtypes = [ctypes.POINTER(_VP)]
    L.mdmc_ctx_create.restype = ctypes.c_int
    L.mdmc_ctx_destroy.argtypes = [_VP]
    L.mdmc_ctx_destroy.restype = None
    L.mdmc_count_device.argtypes = [_VP, _VP, _U3, ctypes.c_float, _VP, _VP]
    L.mdmc_count_device.restype = ctypes.c_int
    L.mdmc_emit_device.argtypes = [_VP, _VP, _F3, _F3, _VP, ctypes.c_uint64, _VP, ctypes.c_uint64, _VP, _VP]
    L.mdmc_emit_device.restype = ctypes.c_int
    L.mdmc_extract_device.argtypes = [_VP, _VP, _U3, _F3, _F3, ctypes.c_float, _VP, ctypes.c_uint64, _VP,
                                      ctypes.c_uint64, _VP, _VP]
    L.mdmc_extract_device.restype = ctypes.c_int
    L.mdmc_run_out_of_core.argtypes = [_VP, _U3, _F3, _F3, ctypes.c_float, ctypes.c_uint64, ctypes.POINTER(Mesh)]
    L.mdmc_run_out_of_core.restype = ctypes.c_int
    L.mdmc_out_of_core_slab_planes.argtypes = [_U3, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32)]
    L.mdmc_out_of_core_slab_planes.restype = ctypes.c_int
    L.mdmc_write_obj.argtypes = [ctypes.c_char_p, _VP, ctypes.c_uint64, _VP, ctypes.c_uint64]
    L.mdmc_write_obj.restype = ctypes.c_int
    L.mdmc_smooth_device.argtypes = [_VP, _VP, ctypes.c_uint64, _VP, ctypes.c_uint64, ctypes.c_uint32, _VP, _VP]
    L.mdmc_smooth_device.restype = ctypes.c_int
    L.mdmc_cells_device.argtypes = [_VP, _VP, _VP, _VP, _VP, ctypes.c_uint64, _VP]
    L.mdmc_cells_device.restype = ctypes.c_int
    L.mdmc_field_blobs_device.argtypes = [_VP, _VP, _U3, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, _VP]
    L.mdmc_field_blobs_device.restype = ctypes.c_int
    L.mdmc_field_gyroid_device.argtypes = [_VP, _VP, _U3, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, _VP]
    L.mdmc_field_gyroid_device.restype = ctypes.c_int
    for name in ("mdmc_batch_count_device", "mdmc_batch_emit_device", "mdmc_batch_extract_device"):
        f = getattr(L, name)
        f.argtypes = [_VP, ctypes.POINTER(Volume), ctypes.c_uint32, _VP, _VP]
        f.restype = ctypes.c_int
    L.mdmc_batch_generate_device.argtypes = [_VP, ctypes.POINTER(GenSpec), ctypes.c_uint32, _VP]
    L.mdmc_batch_generate_device.restype = ctypes.c_int
    L.mdmc_ctx_set_option.argtypes = [_VP, ctypes.c_int, ctypes.c_int]
    L.mdmc_ctx_set_option.restype = ctypes.c_int
    L.mdmc_ctx_profile.argtypes = [_VP, ctypes.c_int]
    L.mdmc_ctx_profile.restype = ctypes.c_int
    L.mdmc_ctx_profile_read.argtypes = [_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32)]
    L.mdmc_ctx_profile_read.restype = ctypes.c_int
    _lib = L
    return L


def _check(rc: int, where: str) -> None:
    if rc != MDMC_OK:
        raise MDMCError(rc, where)


def setup() -> None:
    """dmc::setup_device_luts (include/dmc.hpp:40)."""
    _check(lib().mdmc_setup(), "mdmc_setup")


def cleanup() -> None:
    """dmc::cleanup_device_luts (include/dmc.hpp:48)."""
    _check(lib().mdmc_cleanup(), "mdmc_cleanup")


def _default_box(dims_xyz):
    X, Y, Z = dims_xyz
    return (0.0, 0.0, 0.0), (float(max(X - 1, 1)), float(max(Y - 1, 1)), float(max(Z - 1, 1)))


def triangles_from_quads(quads: np.ndarray) -> np.ndarray:
    """The reference's triangle list: (q0,q1,q2),(q2,q3,q0) per quad (dmc.cu:2039-2045)."""
    q = np.asarray(quads).reshape(-1, 4)
    t = np.empty((q.shape[0], 2, 3), dtype=np.uint32)
    t[:, 0, 0], t[:, 0, 1], t[:, 0, 2] = q[:, 0], q[:, 1], q[:, 2]
    t[:, 1, 0], t[:, 1, 1], t[:, 1, 2] = q[:, 2], q[:, 3], q[:, 0]
    return t.reshape(-1, 3)


def _take_mesh(L, m):
    try:
        V, Q = m.num_vertices, m.num_quads
        v = np.ctypeslib.as_array(m.vertices, shape=(V * 3,)).copy().reshape(V, 3) if V else np.zeros((0, 3), np.float32)
        q = np.ctypeslib.as_array(m.quads, shape=(Q * 4,)).copy().reshape(Q, 4) if Q else np.zeros((0, 4), np.int32)
    finally:
        L.mdmc_mesh_free(ctypes.byref(m))
    return v, q


def run_host(field: np.ndarray, iso: float, xyz_min=None, xyz_max=None, num_smooth: int = 0):
    """mdmc_run: host field in, host (vertices (V,3) float32, quads (Q,4) int32) out."""
    field = np.ascontiguousarray(field, dtype=np.float32)
    Z, Y, X = field.shape
    dmin, dmax = _default_box((X, Y, Z))
    mn = _F3(*(xyz_min if xyz_min is not None else dmin))
    mx = _F3(*(xyz_max if xyz_max is not None else dmax))
    m = Mesh()
    L = lib()
    _check(L.mdmc_run(field.ctypes.data_as(_VP), _U3(X, Y, Z), mn, mx, ctypes.c_float(iso),
                      ctypes.c_uint32(num_smooth), ctypes.byref(m)), "mdmc_run")
    return _take_mesh(L, m)


class _DeviceArray:
    """A library device buffer seen by torch through __cuda_array_interface__."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (int(ptr), False),
                                         "version": 2, "strides": None}


def run_device(field, iso: float, xyz_min=None, xyz_max=None, num_smooth: int = 0, stream=None):
    """mdmc_run_device: a (Z, Y, X) float32 CUDA tensor in, the library's
    device mesh out, returned as (vertices (V,3), quads (Q,4), counts) torch
    tensors on the device (copies; the library mesh is freed here)."""
    import torch

    if not field.is_cuda or field.dtype != torch.float32 or not field.is_contiguous():
        raise ValueError("field must be a contiguous float32 CUDA tensor")
    Z, Y, X = field.shape
    dmin, dmax = _default_box((X, Y, Z))
    mn = _F3(*(xyz_min if xyz_min is not None else dmin))
    mx = _F3(*(xyz_max if xyz_max is not None else dmax))
    st = stream if stream is not None else torch.cuda.current_stream()
    m = DeviceMesh()
    L = lib()
    _check(L.mdmc_run_device(_VP(field.data_ptr()), _U3(X, Y, Z), mn, mx, ctypes.c_float(iso),
                             ctypes.c_uint32(num_smooth), _VP(st.cuda_stream), ctypes.byref(m)), "mdmc_run_device")
    V, Q, A = m.num_vertices, m.num_quads, m.num_active
    try:
        dev = field.device
        v = (torch.as_tensor(_DeviceArray(m.vertices, (V, 3), "<f4"), device=dev).clone() if V
             else torch.zeros((0, 3), dtype=torch.float32, device=dev))
        q = (torch.as_tensor(_DeviceArray(m.quads, (Q, 4), "<i4"), device=dev).clone() if Q
             else torch.zeros((0, 4), dtype=torch.int32, device=dev))
        torch.cuda.synchronize(dev)
    finally:
        L.mdmc_device_mesh_free(ctypes.byref(m))
    return v, q, {"num_active": A, "num_vertices": V, "num_quads": Q}


def run_out_of_core(field: np.ndarray, iso: float, device_budget: int, xyz_min=None, xyz_max=None):
    """mdmc_run_out_of_core: the host field goes through the device in z-slabs
    within device_budget bytes; same mesh as run_host."""
    field = np.ascontiguousarray(field, dtype=np.float32)
    Z, Y, X = field.shape
    dmin, dmax = _default_box((X, Y, Z))
    mn = _F3(*(xyz_min if xyz_min is not None else dmin))
    mx = _F3(*(xyz_max if xyz_max is not None else dmax))
    m = Mesh()
    L = lib()
    _check(L.mdmc_run_out_of_core(field.ctypes.data_as(_VP), _U3(X, Y, Z), mn, mx, ctypes.c_float(iso),
                                  int(device_budget), ctypes.byref(m)), "mdmc_run_out_of_core")
    return _take_mesh(L, m)


def out_of_core_slab_planes(dims_xyz, device_budget: int) -> int:
    """Cell planes per slab of run_out_of_core for these dims and budget."""
    S = ctypes.c_uint32()
    _check(lib().mdmc_out_of_core_slab_planes(_U3(*dims_xyz), int(device_budget), ctypes.byref(S)),
           "mdmc_out_of_core_slab_planes")
    return S.value


def write_obj(path: str, vertices: np.ndarray, quads: np.ndarray) -> None:
    """mdmc_write_obj: the mesh as Wavefront OBJ (two triangles per quad)."""
    v = np.ascontiguousarray(vertices, dtype=np.float32)
    q = np.ascontiguousarray(quads, dtype=np.int32)
    _check(lib().mdmc_write_obj(os.fsencode(path), v.ctypes.data_as(_VP) if len(v) else None, len(v),
                                q.ctypes.data_as(_VP) if len(q) else None, len(q)), "mdmc_write_obj")


def run_dmc(grid: np.ndarray, xyz_min, xyz_max, iso: float, num_smooth: int = 0):
    """dmc::run_dmc (include/dmc.hpp:56-57) with the reference's triangle output."""
    v, q = run_host(grid, iso, xyz_min, xyz_max, num_smooth)
    return v, triangles_from_quads(q)


class Context:
    """A device-resident extraction context (mdmc_ctx) on the current device."""

    def __init__(self):
        import torch  # device memory and streams
        self._torch = torch
        L = lib()
        h = _VP()
        _check(L.mdmc_ctx_create(ctypes.byref(h)), "mdmc_ctx_create")
        self._h = h
        self._counts = torch.zeros(4, dtype=torch.int32, device="cuda")
        self.last_dims = None

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().mdmc_ctx_destroy(self._h)
            self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream):
        torch = self._torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return _VP(s.cuda_stream)

    @staticmethod
    def _dims(field):
        if field.dim() != 3 or field.dtype.is_floating_point is False:
            raise ValueError("field must be a 3-D float32 tensor (Z, Y, X)")
        Z, Y, X = field.shape
        return X, Y, Z

    def counts_tensor(self):
        return self._counts

    def read_counts(self) -> dict:
        c = self._counts.cpu().numpy().view(np.uint32)
        return {"num_active": int(c[0]), "num_vertices": int(c[1]), "num_quads": int(c[2]), "status": int(c[3])}

    def count(self, field, iso: float, counts=None, stream=None) -> None:
        if field.dtype != self._torch.float32 or not field.is_contiguous() or not field.is_cuda:
            raise ValueError("field must be a contiguous float32 CUDA tensor")
        X, Y, Z = self._dims(field)
        c = self._counts if counts is None else counts
        _check(lib().mdmc_count_device(self._h, _VP(field.data_ptr()), _U3(X, Y, Z), ctypes.c_float(iso),
                                       _VP(c.data_ptr()), self._stream(stream)), "mdmc_count_device")
        self.last_dims = (X, Y, Z)

    def emit(self, field, vertices, quads, xyz_min=None, xyz_max=None, counts=None, stream=None) -> None:
        dmin, dmax = _default_box(self.last_dims)
        mn = _F3(*(xyz_min if xyz_min is not None else dmin))
        mx = _F3(*(xyz_max if xyz_max is not None else dmax))
        c = self._counts if counts is None else counts
        vptr = vertices.data_ptr() if vertices is not None and vertices.numel() else None
        qptr = quads.data_ptr() if quads is not None and quads.numel() else None
        vcap = vertices.numel() // 3 if vptr else 0
        qcap = quads.numel() // 4 if qptr else 0
        _check(lib().mdmc_emit_device(self._h, _VP(field.data_ptr()), mn, mx, _VP(vptr), vcap, _VP(qptr), qcap,
                                      _VP(c.data_ptr()), self._stream(stream)), "mdmc_emit_device")

    def smooth(self, vertices, quads, num_smooth: int, counts=None, stream=None) -> None:
        """num_smooth smoothing iterations over the mesh the last emit of this
        context wrote into (vertices, quads)."""
        c = self._counts if counts is None else counts
        vptr = vertices.data_ptr() if vertices is not None and vertices.numel() else None
        qptr = quads.data_ptr() if quads is not None and quads.numel() else None
        _check(lib().mdmc_smooth_device(self._h, _VP(vptr), vertices.numel() // 3 if vptr else 0, _VP(qptr),
                                        quads.numel() // 4 if qptr else 0, int(num_smooth), _VP(c.data_ptr()),
                                        self._stream(stream)), "mdmc_smooth_device")

    def extract(self, field, iso: float, xyz_min=None, xyz_max=None, stream=None, num_smooth: int = 0):
        """Count, size the outputs exactly, emit (then num_smooth smoothing
        iterations).  Returns (vertices (V,3), quads (Q,4), counts)."""
        torch = self._torch
        self.count(field, iso, stream=stream)
        (stream or torch.cuda.current_stream()).synchronize()
        cnt = self.read_counts()
        if cnt["status"]:
            raise MDMCError(-7, f"count status {cnt['status']}")
        V, Q = cnt["num_vertices"], cnt["num_quads"]
        verts = torch.empty((V, 3), dtype=torch.float32, device=field.device)
        quads = torch.empty((Q, 4), dtype=torch.int32, device=field.device)
        self.emit(field, verts, quads, xyz_min, xyz_max, stream=stream)
        if num_smooth:
            self.smooth(verts, quads, num_smooth, stream=stream)
        (stream or torch.cuda.current_stream()).synchronize()
        cnt = self.read_counts()
        if cnt["status"]:
            raise MDMCError(-7, f"emit status {cnt['status']}: " +
                            ", ".join(n for b, n in STATUS_NAMES.items() if cnt["status"] & b))
        return verts, quads, cnt

    def extract_into(self, field, iso, vertices, quads, xyz_min=None, xyz_max=None, counts=None, stream=None):
        """One asynchronous count+emit into preallocated buffers (no host sync)."""
        X, Y, Z = self._dims(field)
        dmin, dmax = _default_box((X, Y, Z))
        mn = _F3(*(xyz_min if xyz_min is not None else dmin))
        mx = _F3(*(xyz_max if xyz_max is not None else dmax))
        c = self._counts if counts is None else counts
        _check(lib().mdmc_extract_device(self._h, _VP(field.data_ptr()), _U3(X, Y, Z), mn, mx,
                                         ctypes.c_float(iso), _VP(vertices.data_ptr()), vertices.numel() // 3,
                                         _VP(quads.data_ptr()), quads.numel() // 4, _VP(c.data_ptr()),
                                         self._stream(stream)), "mdmc_extract_device")
        self.last_dims = (X, Y, Z)

    KERNELS = ("classify", "scan", "emit", "other")

    def profile(self, enable: bool) -> None:
        """Bracket every kernel launch of this context with HIP events."""
        _check(lib().mdmc_ctx_profile(self._h, int(enable)), "mdmc_ctx_profile")

    def profile_read(self) -> dict:
        """{kernel: (total ms, launches)} since the last read."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_uint32 * 4)()
        _check(lib().mdmc_ctx_profile_read(self._h, ms, n), "mdmc_ctx_profile_read")
        return {k: (ms[i], n[i]) for i, k in enumerate(self.KERNELS)}

    def cells(self, num_active: int, stream=None, host: bool = True) -> dict:
        """Per active cell of the last count: cell index, configuration, table
        variant and first vertex (numpy arrays, or CUDA tensors with host=False)."""
        torch = self._torch
        n = max(int(num_active), 1)
        ci = torch.empty(n, dtype=torch.int32, device="cuda")
        cfg = torch.empty(n, dtype=torch.uint8, device="cuda")
        fl = torch.empty(n, dtype=torch.uint8, device="cuda")
        vb = torch.empty(n, dtype=torch.int32, device="cuda")
        _check(lib().mdmc_cells_device(self._h, _VP(ci.data_ptr()), _VP(cfg.data_ptr()), _VP(fl.data_ptr()),
                                       _VP(vb.data_ptr()), num_active, self._stream(stream)), "mdmc_cells_device")
        (stream or torch.cuda.current_stream()).synchronize()
        k = int(num_active)
        if not host:
            return {"cell_index": ci[:k], "cfg": cfg[:k], "lut2": fl[:k], "vertex_begin": vb[:k]}
        return {"cell_index": ci[:k].cpu().numpy().view(np.uint32), "cfg": cfg[:k].cpu().numpy(),
                "lut2": fl[:k].cpu().numpy(), "vertex_begin": vb[:k].cpu().numpy().view(np.uint32)}

    def field_blobs(self, dims_xyz, seed, nblobs, sigma=(8.0, 32.0), amp=(-1.0, 1.0), out=None, stream=None):
        torch = self._torch
        X, Y, Z = dims_xyz
        if out is None:
            out = torch.empty((Z, Y, X), dtype=torch.float32, device="cuda")
        _check(lib().mdmc_field_blobs_device(self._h, _VP(out.data_ptr()), _U3(X, Y, Z), int(seed), int(nblobs),
                                             float(sigma[0]), float(sigma[1]), float(amp[0]), float(amp[1]),
                                             self._stream(stream)), "mdmc_field_blobs_device")
        return out

    def field_gyroid(self, dims_xyz, period=24.0, seed=0, noise_amp=0.05, out=None, stream=None):
        torch = self._torch
        X, Y, Z = dims_xyz
        if out is None:
            out = torch.empty((Z, Y, X), dtype=torch.float32, device="cuda")
        _check(lib().mdmc_field_gyroid_device(self._h, _VP(out.data_ptr()), _U3(X, Y, Z), float(period), int(seed),
                                              float(noise_amp), self._stream(stream)), "mdmc_field_gyroid_device")
        return out

    # ---- batches (many volumes per launch sequence) and launch options

    def set_option(self, option: int, value: int) -> None:
        _check(lib().mdmc_ctx_set_option(self._h, int(option), int(value)), "mdmc_ctx_set_option")

    def batch_counts_tensor(self, n: int):
        return self._torch.zeros((n, 4), dtype=self._torch.int32, device="cuda")

    def batch_count(self, fields, isos, counts, boxes=None, stream=None, _table=None) -> None:
        """Classify + scan many volumes in one launch sequence (counts: (n, 4) int32 CUDA tensor)."""
        arr = _table if _table is not None else _volume_table(fields, isos, boxes)
        _check(lib().mdmc_batch_count_device(self._h, _table_ptr(arr), len(fields), _VP(counts.data_ptr()),
                                             self._stream(stream)), "mdmc_batch_count_device")

    def batch_emit(self, fields, isos, outputs, counts, boxes=None, stream=None, _table=None) -> None:
        arr = _table if _table is not None else _volume_table(fields, isos, boxes, outputs)
        _check(lib().mdmc_batch_emit_device(self._h, _table_ptr(arr), len(fields), _VP(counts.data_ptr()),
                                            self._stream(stream)), "mdmc_batch_emit_device")

    def batch_extract_into(self, fields, isos, outputs, counts, boxes=None, stream=None) -> None:
        """Count + emit into preallocated outputs [(verts, quads), ...] without a host sync."""
        arr = _volume_table(fields, isos, boxes, outputs)
        _check(lib().mdmc_batch_extract_device(self._h, _table_ptr(arr), len(fields), _VP(counts.data_ptr()),
                                               self._stream(stream)), "mdmc_batch_extract_device")

    def batch_extract(self, fields, isos, boxes=None, stream=None):
        """Count, size every output exactly, emit.  Returns [(verts, quads, counts dict), ...].
        The vertex (quad) arrays of a batch are consecutive views of one allocation."""
        torch = self._torch
        n = len(fields)
        table = _volume_table(fields, isos, boxes)
        if n == 0:
            return []
        counts = self.batch_counts_tensor(n)
        self.batch_count(fields, isos, counts, stream=stream, _table=table)
        (stream or torch.cuda.current_stream()).synchronize()
        c = counts.cpu().numpy().view(np.uint32)
        nv = c[:, 1].astype(np.int64)
        nq = c[:, 2].astype(np.int64)
        all_v = torch.empty((int(nv.sum()), 3), dtype=torch.float32, device="cuda")
        all_q = torch.empty((int(nq.sum()), 4), dtype=torch.int32, device="cuda")
        voff = np.concatenate(([0], np.cumsum(nv)[:-1])).astype(np.uint64)
        qoff = np.concatenate(([0], np.cumsum(nq)[:-1])).astype(np.uint64)
        table["vertices"] = np.where(nv > 0, np.uint64(all_v.data_ptr()) + np.uint64(12) * voff, 0)
        table["vertex_capacity"] = nv
        table["quads"] = np.where(nq > 0, np.uint64(all_q.data_ptr()) + np.uint64(16) * qoff, 0)
        table["quad_capacity"] = nq
        self.batch_emit(fields, isos, None, counts, stream=stream, _table=table)
        (stream or torch.cuda.current_stream()).synchronize()
        c = counts.cpu().numpy().view(np.uint32)
        bad = np.nonzero(c[:, 3])[0]
        if len(bad):
            raise MDMCError(-7, f"volume {int(bad[0])}: status {int(c[bad[0], 3])}")
        vs = torch.split(all_v, nv.tolist())
        qs = torch.split(all_q, nq.tolist())
        cl = c[:, :3].tolist()
        return [(vs[i], qs[i], {"num_active": cl[i][0], "num_vertices": cl[i][1], "num_quads": cl[i][2],
                                "status": 0}) for i in range(n)]

    def batch_generate(self, specs, stream=None):
        """Fields for many volumes in one launch.  specs: dicts with dims (X, Y, Z), kind ("blobs" /
        "gyroid") and the generator parameters; returns the (Z, Y, X) tensors."""
        torch = self._torch
        arr = (GenSpec * len(specs))()
        outs = []
        for i, sp in enumerate(specs):
            X, Y, Z = sp["dims"]
            o = torch.empty((Z, Y, X), dtype=torch.float32, device="cuda")
            outs.append(o)
            g = arr[i]
            g.out = o.data_ptr()
            g.dims[0], g.dims[1], g.dims[2] = X, Y, Z
            g.kind = 1 if sp["kind"] == "blobs" else 2
            g.seed = int(sp.get("seed", 0))
            g.nblobs = int(sp.get("nblobs", 16))
            g.sigma_lo, g.sigma_hi = map(float, sp.get("sigma", (8.0, 32.0)))
            g.amp_lo, g.amp_hi = map(float, sp.get("amp", (-1.0, 1.0)))
            g.period = float(sp.get("period", 24.0))
            g.noise_amp = float(sp.get("noise_amp", 0.05))
        _check(lib().mdmc_batch_generate_device(self._h, arr, len(specs), self._stream(stream)),
               "mdmc_batch_generate_device")
        return outs


def _volume_table(fields, isos, boxes=None, outputs=None):
    """Batch table (mdmc_volume records) for torch fields shaped (Z, Y, X)."""
    n = len(fields)
    if len(isos) != n or (boxes is not None and len(boxes) != n) or (outputs is not None and len(outputs) != n):
        raise ValueError("isos, boxes and outputs need one entry per field")
    tab = np.zeros(n, dtype=_VOLUME_DTYPE)
    if n == 0:
        return tab
    import torch

    ptrs, shapes = [], []
    for f in fields:
        if f.dtype != torch.float32 or f.dim() != 3 or not f.is_contiguous():
            raise ValueError("batch fields must be contiguous 3-D float32 CUDA tensors")
        ptrs.append(f.data_ptr())
        shapes.append(f.shape)
    tab["field"] = ptrs
    zyx = np.array(shapes, dtype=np.uint32)
    tab["dims"] = zyx[:, ::-1]
    tab["iso"] = np.asarray(isos, dtype=np.float64)
    if boxes is None:
        tab["xyz_max"] = np.maximum(zyx[:, ::-1].astype(np.int64) - 1, 1)
    else:
        tab["xyz_min"] = [b[0] for b in boxes]
        tab["xyz_max"] = [b[1] for b in boxes]
    if outputs is not None:
        for i, o in enumerate(outputs):
            if o is None:
                continue
            verts, quads = o
            if verts is not None and verts.numel():
                tab["vertices"][i] = verts.data_ptr()
                tab["vertex_capacity"][i] = verts.numel() // 3
            if quads is not None and quads.numel():
                tab["quads"][i] = quads.data_ptr()
                tab["quad_capacity"][i] = quads.numel() // 4
    return tab


def _table_ptr(tab):
    return tab.ctypes.data_as(ctypes.POINTER(Volume))


