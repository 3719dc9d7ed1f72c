"""CPU oracle for the Dual Marching Cubes extraction path.

TEST INFRASTRUCTURE ONLY.  A ctypes wrapper over ``oracle/build/liboracle.so``,
the plain-C restatement of the reference ``dmc::run_dmc``
(DMC_CUDA/DMC_CUDA/src/dmc.cu:2069-2208).  Only the tests, the smoke check and
the benchmark's CPU-baseline leg use it, always as the checker or the timed
baseline, never as the product path.

Parity status of each oracle function (how it is pinned is in the tests):
  LUT builder           pinned: SHA-256 of the reference arrays
  classification/scans  pinned: brute force on tiny grids, closed forms
  2B/3B correction      pinned: hand-built two-cell cases for all 36 configs
  edge points/vertices  pinned: closed forms (axis-aligned planes, spheres),
                        FP64 twin bound, symmetry
  quads                 pinned: mesh invariants (closed, oriented 2-manifold,
                        Euler characteristic), hand-built rings
  smoothing             pinned: hand-built rings (convex, concave, crossing,
                        outside), closed form on planes, no-op when no
                        triangle holds the edge
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


class _Result(ctypes.Structure):
    _fields_ = [
        ("num_active", ctypes.c_uint32),
        ("num_vertices", ctypes.c_uint32),
        ("num_quads", ctypes.c_uint32),
        ("num_tie_edges", ctypes.c_uint32),
        ("num_unset_vertices", ctypes.c_uint32),
        ("num_anomalies", ctypes.c_uint32),
        ("vertices", ctypes.POINTER(ctypes.c_float)),
        ("vertices64", ctypes.POINTER(ctypes.c_double)),
        ("quads", ctypes.POINTER(ctypes.c_int32)),
        ("cell_index", ctypes.POINTER(ctypes.c_uint32)),
        ("cfg", ctypes.POINTER(ctypes.c_uint8)),
        ("info", ctypes.POINTER(ctypes.c_uint8)),
        ("vertex_begin", ctypes.POINTER(ctypes.c_uint32)),
        ("quad_begin", ctypes.POINTER(ctypes.c_uint32)),
    ]


def build() -> None:
    """Compile the oracle (gcc) into oracle/build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f3 = ctypes.POINTER(ctypes.c_float)
        u3 = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_dmc_run.argtypes = [f3, u3, f3, f3, ctypes.c_float, ctypes.c_int, ctypes.POINTER(_Result)]
        L.oracle_dmc_run.restype = ctypes.c_int
        L.oracle_dmc_run_smooth.argtypes = [f3, u3, f3, f3, ctypes.c_float, ctypes.c_int, ctypes.c_uint32,
                                             ctypes.POINTER(_Result)]
        L.oracle_dmc_run_smooth.restype = ctypes.c_int
        L.oracle_dmc_run_window.argtypes = [f3, u3, u3, u3, f3, f3, ctypes.c_float, ctypes.c_int, ctypes.c_uint32,
                                            ctypes.POINTER(_Result)]
        L.oracle_dmc_run_window.restype = ctypes.c_int
        L.oracle_result_free.argtypes = [ctypes.POINTER(_Result)]
        L.oracle_luts_bytes.argtypes = [ctypes.POINTER(ctypes.c_uint8)]
        L.oracle_field_sphere.argtypes = [f3, u3, ctypes.POINTER(ctypes.c_double), ctypes.c_double]
        L.oracle_field_blobs.argtypes = [f3, u3, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.oracle_field_gyroid.argtypes = [f3, u3, ctypes.c_double, ctypes.c_uint64, ctypes.c_double]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = None
        _lib = L
    return _lib


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dims(dims):
    return (ctypes.c_uint32 * 3)(*[int(d) for d in dims])


def set_threads(n: int) -> None:
    """OpenMP threads for the oracle's independent per-cell loops (results do
    not depend on the count)."""
    lib().oracle_set_threads(int(n))


def luts_bytes() -> bytes:
    buf = (ctypes.c_uint8 * 7168)()
    lib().oracle_luts_bytes(buf)
    return bytes(buf)


def run(field: np.ndarray, iso: float, xyz_min=None, xyz_max=None, fp64: bool = False, num_smooth: int = 0,
        grid_cells=None, offset=None) -> dict:
    """Extract the mesh of ``field`` (shape (Z, Y, X), float32, x fastest),
    with ``num_smooth`` smoothing iterations.

    Defaults put vertices in voxel units (min 0, max = cells per axis).  With
    ``grid_cells`` (x, y, z) the field is a window of a grid of that many
    cells, starting at point ``offset`` (x, y, z) of it: coordinates are the
    whole grid's (default box: that grid's voxel units)."""
    field = np.ascontiguousarray(field, dtype=np.float32)
    Z, Y, X = field.shape
    if xyz_min is None:
        xyz_min = (0.0, 0.0, 0.0)
    if xyz_max is None:
        xyz_max = ((float(c) for c in grid_cells) if grid_cells is not None else
                   (float(max(X - 1, 1)), float(max(Y - 1, 1)), float(max(Z - 1, 1))))
    mn = np.asarray(xyz_min, dtype=np.float32)
    mx = np.asarray(tuple(xyz_max), dtype=np.float32)
    r = _Result()
    if grid_cells is None and offset is None:
        rc = lib().oracle_dmc_run_smooth(_fptr(field), _dims((X, Y, Z)), _fptr(mn), _fptr(mx),
                                         ctypes.c_float(iso), int(fp64), int(num_smooth), ctypes.byref(r))
    else:
        gc = grid_cells if grid_cells is not None else (X - 1, Y - 1, Z - 1)
        of = offset if offset is not None else (0, 0, 0)
        rc = lib().oracle_dmc_run_window(_fptr(field), _dims((X, Y, Z)), _dims(gc), _dims(of), _fptr(mn), _fptr(mx),
                                         ctypes.c_float(iso), int(fp64), int(num_smooth), ctypes.byref(r))
    if rc != 0:
        raise MemoryError("oracle_dmc_run failed")
    try:
        A, V, Q = r.num_active, r.num_vertices, r.num_quads

        def arr(p, n, dt):
            if n == 0 or not p:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(p, shape=(n,)).copy()

        out = {
            "num_active": A, "num_vertices": V, "num_quads": Q,
            "num_tie_edges": r.num_tie_edges, "num_unset_vertices": r.num_unset_vertices,
            "num_anomalies": r.num_anomalies,
            "vertices": arr(r.vertices, 3 * V, np.float32).reshape(V, 3),
            "quads": arr(r.quads, 4 * Q, np.int32).reshape(Q, 4),
            "cell_index": arr(r.cell_index, A, np.uint32),
            "cfg": arr(r.cfg, A, np.uint8),
            "info": arr(r.info, A, np.uint8),
            "vertex_begin": arr(r.vertex_begin, A, np.uint32),
            "quad_begin": arr(r.quad_begin, A, np.uint32),
        }
        if fp64:
            out["vertices64"] = arr(r.vertices64, 3 * V, np.float64).reshape(V, 3)
        return out
    finally:
        lib().oracle_result_free(ctypes.byref(r))


def field_sphere(dims, centre, radius) -> np.ndarray:
    X, Y, Z = dims
    out = np.empty((Z, Y, X), dtype=np.float32)
    c = (ctypes.c_double * 3)(*centre)
    lib().oracle_field_sphere(_fptr(out), _dims(dims), c, float(radius))
    return out


def field_blobs(dims, seed, nblobs, sigma=(8.0, 32.0), amp=(-1.0, 1.0)) -> np.ndarray:
    X, Y, Z = dims
    out = np.empty((Z, Y, X), dtype=np.float32)
    lib().oracle_field_blobs(_fptr(out), _dims(dims), int(seed), int(nblobs), float(sigma[0]),
                             float(sigma[1]), float(amp[0]), float(amp[1]))
    return out


def field_gyroid(dims, period=24.0, seed=0, noise_amp=0.05) -> np.ndarray:
    X, Y, Z = dims
    out = np.empty((Z, Y, X), dtype=np.float32)
    lib().oracle_field_gyroid(_fptr(out), _dims(dims), float(period), int(seed), float(noise_amp))
    return out
