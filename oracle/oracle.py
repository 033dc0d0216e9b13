"""ctypes front end of the CPU checkers (test infrastructure only).

Loads ``oracle/libdd_oracle.so`` (the C++ restatement in ``dd_oracle.cpp``)
and, when it has been built, ``oracle/_ref/libdd_ref2d.so`` (the reference's
own CPU projector sources compiled in place by ``oracle/Makefile``).  Only the
tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module; the product package never does.

Geometry parameters follow the reference's 2D ``Options`` layout
``[W, H, D, dx, dy, det, sid, sdd, shift_x, shift_y, bin_shift]``
(deeprecon/deeprecon/torch/projector/projector.py:113), with ``det`` and
``bin_shift`` already in radians for the arc detector (projector.py:110-112).
"""
from __future__ import annotations

import contextlib
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "libdd_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libdd_ref2d.so")

GEOMS = {"para": 0, "flat": 1, "arc": 2}
OPS = {
    "proj": 0,
    "proj_t": 1,
    "bproj": 2,
    "bproj_t": 3,
    "bproj_w": 4,
    "bproj_tw": 5,
}
# ops whose output is a sinogram
TO_SINO = {"proj", "bproj_t", "bproj_tw"}


@dataclass(frozen=True)
class Geom2D:
    W: int
    H: int
    D: int
    dx: float
    dy: float
    det: float
    sid: float
    sdd: float
    shx: float = 0.0
    shy: float = 0.0
    bin: float = 0.0

    def options(self):
        return [self.W, self.H, self.D, self.dx, self.dy, self.det, self.sid,
                self.sdd, self.shx, self.shy, self.bin]

    def f32(self):
        """The float parameters as the reference decodes them (``.item<float>()``)."""
        return [float(np.float32(v)) for v in (self.dx, self.dy, self.det, self.sid,
                                               self.sdd, self.shx, self.shy, self.bin)]


_lib = None
_ref = None

_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)


def _load_oracle():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        lib = ctypes.CDLL(ORACLE_SO)
        i, f, d = ctypes.c_int, ctypes.c_float, ctypes.c_double
        lib.ddo_op2d_f32.argtypes = [i, i, _F32P, _F32P, _F32P] + [i] * 5 + [f] * 8
        lib.ddo_op2d_f32.restype = i
        lib.ddo_op2d_f64.argtypes = [i, i, _F64P, _F64P, _F64P] + [i] * 5 + [d] * 8
        lib.ddo_op2d_f64.restype = i
        lib.ddo_sysmatrix2d.argtypes = [i, _F32P] + [i] * 4 + [f] * 8
        lib.ddo_sysmatrix2d.restype = ctypes.c_int64
        lib.ddo_sysmatrix_copy.argtypes = [_I64P, _I64P, _F32P]
        lib.ddo_sysmatrix_copy.restype = None
        if hasattr(lib, "ddo_op3d_f64"):
            _bind3d(lib)
        _lib = lib
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference CPU library missing: {REF_SO} (run make -C oracle ref)")
        lib = ctypes.CDLL(REF_SO)
        i, f = ctypes.c_int, ctypes.c_float
        lib.ddref_op2d.argtypes = [i, i, _F32P, _F32P, _F32P] + [i] * 5 + [f] * 8
        lib.ddref_op2d.restype = i
        lib.ddref_sysmatrix2d.argtypes = [i, _F32P] + [i] * 4 + [f] * 8
        lib.ddref_sysmatrix2d.restype = ctypes.c_int64
        lib.ddref_sysmatrix_copy.argtypes = [_I64P, _I64P, _F32P]
        lib.ddref_sysmatrix_copy.restype = None
        _ref = lib
    return _ref


def _shapes(g: Geom2D, batch: int, views: int):
    return (batch, g.H, g.W), (batch, views, g.D)


def op2d(geom: str, op: str, x: np.ndarray, angles: np.ndarray, g: Geom2D,
         mode: str = "f32") -> np.ndarray:
    """Runs one 2D op on the CPU.

    ``x`` is the op's input: an image batch ``[B, H, W]`` for ``proj``,
    ``bproj_t`` and ``bproj_tw``, otherwise a sinogram batch ``[B, V, D]``.
    ``mode`` is ``"f32"`` (mirror of the reference's FP32 evaluation),
    ``"f64"`` (FP64 oracle) or ``"ref"`` (the reference's own code).
    """
    views = int(np.asarray(angles).shape[0])
    batch = int(x.shape[0])
    img_shape, sino_shape = _shapes(g, batch, views)
    to_sino = op in TO_SINO
    exp_in = img_shape if to_sino else sino_shape
    if tuple(x.shape) != exp_in:
        raise ValueError(f"input shape {x.shape} != expected {exp_in}")
    out_shape = sino_shape if to_sino else img_shape
    gi, oi = GEOMS[geom], OPS[op]
    if mode == "f64":
        lib = _load_oracle()
        xin = np.ascontiguousarray(x, dtype=np.float64)
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        out = np.zeros(out_shape, dtype=np.float64)
        img, sino = (xin, out) if to_sino else (out, xin)
        params = g.f32()  # the op receives float32 Options
        rc = lib.ddo_op2d_f64(gi, oi, img.ctypes.data_as(_F64P), sino.ctypes.data_as(_F64P),
                              ang.ctypes.data_as(_F64P), batch, g.W, g.H, views, g.D, *params)
    else:
        xin = np.ascontiguousarray(x, dtype=np.float32).copy()
        ang = np.ascontiguousarray(angles, dtype=np.float32).copy()
        out = np.zeros(out_shape, dtype=np.float32)
        img, sino = (xin, out) if to_sino else (out, xin)
        fn = _load_ref().ddref_op2d if mode == "ref" else _load_oracle().ddo_op2d_f32
        rc = fn(gi, oi, img.ctypes.data_as(_F32P), sino.ctypes.data_as(_F32P),
                ang.ctypes.data_as(_F32P), batch, g.W, g.H, views, g.D, *g.f32())
    if rc != 0:
        raise ValueError(f"unsupported geometry/op pair: {geom}/{op}")
    return out


def sysmatrix2d(geom: str, angles: np.ndarray, g: Geom2D, mode: str = "f32"):
    """COO system matrix ``(rows, cols, vals)`` with the FP weights."""
    ang = np.ascontiguousarray(angles, dtype=np.float32).copy()
    views = int(ang.shape[0])
    if mode == "ref":
        lib = _load_ref()
        nnz = lib.ddref_sysmatrix2d(GEOMS[geom], ang.ctypes.data_as(_F32P), g.W, g.H, views,
                                    g.D, *g.f32())
        copy = lib.ddref_sysmatrix_copy
    else:
        lib = _load_oracle()
        nnz = lib.ddo_sysmatrix2d(GEOMS[geom], ang.ctypes.data_as(_F32P), g.W, g.H, views,
                                  g.D, *g.f32())
        copy = lib.ddo_sysmatrix_copy
    if nnz < 0:
        raise ValueError(f"unknown geometry {geom}")
    rows = np.empty(nnz, dtype=np.int64)
    cols = np.empty(nnz, dtype=np.int64)
    vals = np.empty(nnz, dtype=np.float32)
    copy(rows.ctypes.data_as(_I64P), cols.ctypes.data_as(_I64P), vals.ctypes.data_as(_F32P))
    return rows, cols, vals


def _bind3d(lib):
    """Binds the 3D entry points once dd_oracle3d.cpp is part of the library."""
    i, d = ctypes.c_int, ctypes.c_double
    lib.ddo_op3d_f64.argtypes = [i, i, _F64P, _F64P, _F64P, _F64P] + [i] * 7 + [d] * 12
    lib.ddo_op3d_f64.restype = i
    lib.ddo_plant3d.argtypes = [i]
    lib.ddo_plant3d.restype = None


# planted-mistake bits of dd_oracle3d.cpp (tests only)
PLANT3D = {"pix_z_shift": 1, "bin_row_shift": 2, "radius": 4, "angle_source": 8,
           "angle_columns": 16, "axial": 32, "col_pitch": 64}


@contextlib.contextmanager
def planted3d(name: str):
    """Runs the 3D oracle with one planted mistake (a flipped sign or a
    dropped factor), to show that a pinning identity fails on it."""
    lib = _load_oracle()
    lib.ddo_plant3d(PLANT3D[name])
    try:
        yield
    finally:
        lib.ddo_plant3d(0)


@dataclass(frozen=True)
class Geom3D:
    """The reference's 3D Options: [W, H, Dz, cols, rows, dx, dy, dz, col, row,
    sid, sdd, sx, sy, sz, col_shift, row_shift] (projector.py:286); for the
    spiral the 16-entry form without sz (projector.py:460).  Arc / spiral
    column size and shift are in radians."""
    W: int
    H: int
    Dz: int
    C: int
    R: int
    dx: float
    dy: float
    dz: float
    col: float
    row: float
    sid: float
    sdd: float
    shx: float = 0.0
    shy: float = 0.0
    shz: float = 0.0
    csh: float = 0.0
    rsh: float = 0.0

    def options(self, spiral: bool = False):
        o = [self.W, self.H, self.Dz, self.C, self.R, self.dx, self.dy, self.dz, self.col,
             self.row, self.sid, self.sdd, self.shx, self.shy]
        return o + ([self.csh, self.rsh] if spiral else [self.shz, self.csh, self.rsh])


SCANS3D = {"cone_flat": 3, "cone_arc": 4, "spiral": 5}


def op3d(scan: str, op: str, x: np.ndarray, angles: np.ndarray, g: Geom3D,
         spos: np.ndarray | None = None) -> np.ndarray:
    """3D op in FP64 (restatement of the reference's CUDA kernels).

    ``x`` is a volume ``[B, Dz, H, W]`` for image->sinogram ops, else a
    sinogram ``[B, V, rows, cols]``; ``spos`` is ``[4, V]`` for the spiral.
    Parameters are taken at float32 precision, as the op receives them.
    """
    lib = _load_oracle()
    views = int(np.asarray(angles).shape[0])
    batch = int(x.shape[0])
    img_shape, sino_shape = (batch, g.Dz, g.H, g.W), (batch, views, g.R, g.C)
    to_sino = op in TO_SINO
    if tuple(x.shape) != (img_shape if to_sino else sino_shape):
        raise ValueError(f"input shape {x.shape} does not match the geometry")
    xin = np.ascontiguousarray(x, dtype=np.float64)
    ang = np.ascontiguousarray(np.asarray(angles, dtype=np.float32), dtype=np.float64)
    out = np.zeros(sino_shape if to_sino else img_shape, dtype=np.float64)
    img, sino = (xin, out) if to_sino else (out, xin)
    sp = None
    if scan == "spiral":
        if spos is None:
            raise ValueError("spiral needs source_pos")
        sp = np.ascontiguousarray(np.asarray(spos, dtype=np.float32), dtype=np.float64)
    params = [float(np.float32(v)) for v in (g.dx, g.dy, g.dz, g.col, g.row, g.sid, g.sdd,
                                             g.shx, g.shy, g.shz, g.csh, g.rsh)]
    rc = lib.ddo_op3d_f64(SCANS3D[scan], OPS[op], img.ctypes.data_as(_F64P), sino.ctypes.data_as(_F64P),
                          ang.ctypes.data_as(_F64P),
                          sp.ctypes.data_as(_F64P) if sp is not None else None,
                          batch, g.W, g.H, g.Dz, views, g.C, g.R, *params)
    if rc != 0:
        raise ValueError(f"unsupported 3D call {scan}/{op}")
    return out


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def max_abs(a, b) -> float:
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))
