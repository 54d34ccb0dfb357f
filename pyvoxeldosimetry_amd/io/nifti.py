"""NIfTI-1 single-file images (``.nii``, ``.nii.gz``) without nibabel.

The reference writes and reads dose maps through nibabel
(``pyvoxeldosimetry/core/utils.py:53-152``, ``pyvoxeldosimetry/io/nifti.py:6-16``),
which is not a dependency here.  This module restates the published NIfTI-1
format (the 348-byte ``nifti_1_header`` of ``nifti1.h``, the 4-byte extender,
extensions of 16-byte multiples, the voxel data from ``vox_offset``, first index
fastest) for what the dose path needs: 3-D scalar volumes, an affine (sform,
else qform, else the voxel sizes), and extensions such as the reference's
code-44 JSON metadata.

Writing follows nibabel's choices for ``Nifti1Image(data, affine)``: sform =
affine with code 2 ("aligned"), qform code 0, pixdim from the affine's column
norms, no intensity scaling (``scl_slope`` NaN), ``vox_offset`` 352 plus the
extensions.
"""
from __future__ import annotations

import gzip
import os
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

HEADER_SIZE = 348
MAGIC_SINGLE = b"n+1\x00"

# NIfTI-1 datatype codes (nifti1.h) -> NumPy dtypes
DATATYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8,
             512: np.uint16, 768: np.uint32, 1024: np.int64, 1280: np.uint64}
CODES = {np.dtype(v).str[1:]: k for k, v in DATATYPES.items()}

# (name, struct format, offset) of the fields this module uses, nifti1.h order
_FIELDS = [
    ("sizeof_hdr", "i", 0), ("regular", "c", 38), ("dim_info", "B", 39), ("dim", "8h", 40),
    ("intent_p", "3f", 56), ("intent_code", "h", 68), ("datatype", "h", 70), ("bitpix", "h", 72),
    ("slice_start", "h", 74), ("pixdim", "8f", 76), ("vox_offset", "f", 108), ("scl_slope", "f", 112),
    ("scl_inter", "f", 116), ("slice_end", "h", 120), ("slice_code", "B", 122), ("xyzt_units", "B", 123),
    ("cal_max", "f", 124), ("cal_min", "f", 128), ("slice_duration", "f", 132), ("toffset", "f", 136),
    ("descrip", "80s", 148), ("aux_file", "24s", 228), ("qform_code", "h", 252), ("sform_code", "h", 254),
    ("quatern", "3f", 256), ("qoffset", "3f", 268), ("srow_x", "4f", 280), ("srow_y", "4f", 296),
    ("srow_z", "4f", 312), ("intent_name", "16s", 328), ("magic", "4s", 344),
]


@dataclass
class Extension:
    code: int
    content: bytes


@dataclass
class NiftiImage:
    data: np.ndarray
    affine: np.ndarray
    pixdim: Tuple[float, ...]
    descrip: str = ""
    extensions: List[Extension] = field(default_factory=list)
    datatype: int = 16
    qform_code: int = 0
    sform_code: int = 2


def _open(path: str, mode: str):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def _quaternion_affine(quatern, qoffset, pixdim) -> np.ndarray:
    """qform affine: rotation from (b, c, d) with a = sqrt(1 - b^2 - c^2 - d^2),
    columns scaled by pixdim[1:4] (the third by qfac = pixdim[0] = -1 or 1)."""
    b, c, d = (float(x) for x in quatern)
    a = np.sqrt(max(0.0, 1.0 - (b * b + c * c + d * d)))
    r = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                  [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
                  [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b]])
    qfac = -1.0 if pixdim[0] < 0 else 1.0
    aff = np.eye(4)
    aff[:3, :3] = r * np.array([pixdim[1], pixdim[2], qfac * pixdim[3]])
    aff[:3, 3] = qoffset
    return aff


def parse_header(raw: bytes) -> dict:
    """Fields of a NIfTI-1 header (either byte order); raises ValueError if it is not one."""
    if len(raw) < HEADER_SIZE:
        raise ValueError("file shorter than a NIfTI-1 header")
    for endian in "<>":
        if struct.unpack(endian + "i", raw[:4])[0] == HEADER_SIZE:
            break
    else:
        raise ValueError("sizeof_hdr is not 348: not a NIfTI-1 file")
    h = {"endian": endian}
    for name, fmt, off in _FIELDS:
        vals = struct.unpack_from(endian + fmt, raw, off)
        h[name] = vals[0] if len(vals) == 1 else vals
    if h["magic"] not in (MAGIC_SINGLE, b"ni1\x00"):
        raise ValueError(f"bad NIfTI-1 magic {h['magic']!r}")
    return h


def header_affine(h: dict) -> np.ndarray:
    """Best affine: sform if its code is set, else qform, else the voxel sizes."""
    if h["sform_code"] > 0:
        aff = np.eye(4)
        aff[0], aff[1], aff[2] = h["srow_x"], h["srow_y"], h["srow_z"]
        return aff
    if h["qform_code"] > 0:
        return _quaternion_affine(h["quatern"], h["qoffset"], h["pixdim"])
    return np.diag([h["pixdim"][1], h["pixdim"][2], h["pixdim"][3], 1.0])


def load(path: str) -> NiftiImage:
    with _open(path, "rb") as f:
        raw = f.read()
    h = parse_header(raw)
    if h["magic"] != MAGIC_SINGLE:
        raise ValueError("header/image pairs (.hdr/.img) are not supported")
    e = h["endian"]
    ndim = h["dim"][0]
    if not 1 <= ndim <= 7:
        raise ValueError(f"dim[0] = {ndim} outside 1..7")
    shape = tuple(int(x) for x in h["dim"][1:1 + ndim])
    if h["datatype"] not in DATATYPES:
        raise ValueError(f"unsupported NIfTI datatype {h['datatype']}")
    dtype = np.dtype(DATATYPES[h["datatype"]]).newbyteorder(e)
    off = int(h["vox_offset"])
    exts = []
    pos = HEADER_SIZE
    if len(raw) >= pos + 4 and raw[pos] != 0:
        pos += 4
        while pos + 8 <= off:
            esize, ecode = struct.unpack_from(e + "2i", raw, pos)
            if esize < 8 or pos + esize > off:
                break
            exts.append(Extension(ecode, raw[pos + 8:pos + esize]))
            pos += esize
    count = int(np.prod(shape))
    if off + count * dtype.itemsize > len(raw):
        raise ValueError("voxel data shorter than the header's dimensions")
    data = np.frombuffer(raw, dtype=dtype, count=count, offset=off).reshape(shape, order="F")
    slope, inter = h["scl_slope"], h["scl_inter"]
    if np.isfinite(slope) and slope != 0 and (slope != 1 or (np.isfinite(inter) and inter != 0)):
        data = data * slope + (inter if np.isfinite(inter) else 0.0)
    else:
        data = data.astype(dtype.newbyteorder("="))
    descrip = h["descrip"].split(b"\x00", 1)[0].decode("latin-1")
    return NiftiImage(data=np.ascontiguousarray(data), affine=header_affine(h), pixdim=tuple(h["pixdim"]),
                      descrip=descrip, extensions=exts, datatype=h["datatype"], qform_code=h["qform_code"],
                      sform_code=h["sform_code"])


def save(path: str, data: np.ndarray, affine: np.ndarray, pixdim_xyz: Optional[Tuple[float, float, float]] = None,
         descrip: str = "", extensions: Optional[List[Extension]] = None) -> None:
    """Single-file NIfTI-1 (gzip when the name ends in .gz), little-endian, the
    array's first index fastest in the file."""
    data = np.asarray(data)
    if data.ndim < 1 or data.ndim > 7:
        raise ValueError("NIfTI-1 holds 1 to 7 dimensions")
    code = CODES.get(data.dtype.str[1:])
    if code is None:
        raise ValueError(f"dtype {data.dtype} has no NIfTI-1 code")
    affine = np.asarray(affine, dtype=np.float64)
    if affine.shape != (4, 4):
        raise ValueError("affine must be 4x4")
    if pixdim_xyz is None:
        pixdim_xyz = tuple(float(np.linalg.norm(affine[:3, i])) for i in range(3))
    ext_blob = b""
    for ext in extensions or []:
        body = bytes(ext.content)
        esize = (8 + len(body) + 15) // 16 * 16
        ext_blob += struct.pack("<2i", esize, int(ext.code)) + body + b"\x00" * (esize - 8 - len(body))
    vox_offset = HEADER_SIZE + 4 + len(ext_blob)
    hdr = bytearray(HEADER_SIZE)
    dims = [data.ndim] + list(data.shape) + [1] * (7 - data.ndim)
    pixdim = [1.0] + list(pixdim_xyz) + [1.0] * 4
    values = {
        "sizeof_hdr": HEADER_SIZE, "regular": b"r", "dim_info": 0, "dim": dims, "intent_p": (0.0, 0.0, 0.0),
        "intent_code": 0, "datatype": code, "bitpix": data.dtype.itemsize * 8, "slice_start": 0,
        "pixdim": pixdim, "vox_offset": float(vox_offset), "scl_slope": float("nan"),
        "scl_inter": float("nan"), "slice_end": 0, "slice_code": 0, "xyzt_units": 0, "cal_max": 0.0,
        "cal_min": 0.0, "slice_duration": 0.0, "toffset": 0.0,
        "descrip": descrip.encode("latin-1", "replace")[:80], "aux_file": b"", "qform_code": 0,
        "sform_code": 2, "quatern": (0.0, 0.0, 0.0), "qoffset": (0.0, 0.0, 0.0),
        "srow_x": tuple(affine[0]), "srow_y": tuple(affine[1]), "srow_z": tuple(affine[2]),
        "intent_name": b"", "magic": MAGIC_SINGLE,
    }
    for name, fmt, off in _FIELDS:
        v = values[name]
        struct.pack_into("<" + fmt, hdr, off, *(v if isinstance(v, (list, tuple)) else (v,)))
    extender = b"\x01\x00\x00\x00" if ext_blob else b"\x00\x00\x00\x00"
    body = np.asarray(data, dtype=data.dtype.newbyteorder("<")).tobytes(order="F")
    with _open(path, "wb") as f:
        f.write(bytes(hdr) + extender + ext_blob + body)


def is_nifti_file(file_path) -> bool:
    """Same contract as ``pyvoxeldosimetry/io/nifti.py:6-16``: an existing .nii /
    .nii.gz file whose header parses."""
    file_path = str(file_path)
    if not os.path.isfile(file_path) or not file_path.endswith((".nii", ".nii.gz")):
        return False
    try:
        with _open(file_path, "rb") as f:
            parse_header(f.read(HEADER_SIZE))
        return True
    except Exception:
        return False
