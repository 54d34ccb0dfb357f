"""Kernel and dose-map files: the on-disk formats on either side of the dose path.

Mirrors ``pyvoxeldosimetry/core/utils.py``:
- ``load_kernel`` (``:17-51``): the binary ``.dat`` dose-point kernel;
- ``save_dose_map`` / ``load_dose_map`` (``:53-152``): NIfTI-1 dose maps with the
  metadata as a code-44 JSON extension (``pyvoxeldosimetry_amd.io.nifti``
  replaces nibabel);
- ``verify_file_integrity`` (``:264-301``) and ``get_file_metadata`` (``:303-330``);
- ``resample_dose_map`` (``:193-231``): order-3 spline resampling of a dose map,
  on the GPU (``pyvoxeldosimetry_amd.resample``).

The ``.dat`` layout is the one the reference's reader reads (little-endian):
``int32 dims[3]``, ``float32 voxel_size, total_energy, scaling``, ``int32
timestamp[6]`` (Y M D h m s), ``int8 user[32]`` (80 bytes), then ``float32``
data in C order; the kernel is ``data.reshape(dims) * scaling``.  The kernel
README (``data/dose_kernels/F18/README.md:13-17``) documents a 24-byte header
(dims, voxel_size, total_energy, scaling) instead; ``load_kernel`` reads both,
telling them apart by the file size.  ``verify_file_integrity`` checks the file
size against the layout the reader would use; the reference's check,
``(prod(dims) + 11) * 4`` bytes (``:290``), matches neither layout and so
rejects every file its own reader accepts.
"""
from __future__ import annotations

import json
import os
from datetime import datetime
from pathlib import Path
from typing import Any, Dict, Optional, Tuple, Union

import numpy as np

from ..io import nifti

DAT_HEADER = 80            # the reference reader's header
DAT_HEADER_SHORT = 24      # the kernel README's header
META_EXTENSION_CODE = 44   # user-defined extension holding JSON metadata (utils.py:98-102)


def _dat_layout(path: Union[str, Path]) -> Tuple[np.ndarray, int]:
    """(dims, header bytes) of a .dat kernel file; ValueError if the size fits neither header."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        dims = np.fromfile(f, dtype="<i4", count=3)
    if len(dims) != 3 or (dims <= 0).any():
        raise ValueError(f"{path}: no valid kernel dimensions in the header")
    n = int(np.prod(dims.astype(np.int64)))
    for hdr in (DAT_HEADER, DAT_HEADER_SHORT):
        if size == hdr + 4 * n:
            return dims, hdr
    raise ValueError(f"{path}: {size} bytes is neither {DAT_HEADER} nor {DAT_HEADER_SHORT} header bytes "
                     f"plus {n} float32 voxels")


def load_kernel(filename: Union[str, Path]) -> Tuple[np.ndarray, Dict[str, Any]]:
    """Kernel array (float32, dims from the header, times the scaling factor) and
    its metadata, as ``core/utils.py:17-51`` returns them."""
    dims, hdr = _dat_layout(filename)
    with open(filename, "rb") as f:
        f.seek(12)
        voxel_size, total_energy, scaling = (np.float32(x) for x in np.fromfile(f, dtype="<f4", count=3))
        created, user = None, ""
        if hdr == DAT_HEADER:
            timestamp = np.fromfile(f, dtype="<i4", count=6)
            user = bytes(np.fromfile(f, dtype=np.int8, count=32).astype(np.uint8)).decode(errors="replace")
            user = user.strip("\x00")
            created = datetime(*(int(t) for t in timestamp)).strftime("%Y-%m-%d %H:%M:%S")
        kernel = np.fromfile(f, dtype="<f4").astype(np.float32)
    kernel = kernel.reshape(tuple(int(d) for d in dims)) * scaling
    metadata = {
        "dimensions": dims.astype(np.int32),
        "voxel_size": voxel_size,
        "total_energy": total_energy,
        "scaling_factor": scaling,
        "creation_date": created,
        "created_by": user,
    }
    return kernel, metadata


def save_kernel(filename: Union[str, Path], kernel: np.ndarray, voxel_size: float, total_energy: float = 0.0,
                scaling: float = 1.0, created: Optional[datetime] = None, created_by: str = "") -> None:
    """Write a kernel in the layout ``load_kernel`` reads (80-byte header; the
    stored values are ``kernel / scaling`` so that reading multiplies them back).
    The reference has no writer; this one lets the GPU-generated kernels be kept on disk."""
    kernel = np.asarray(kernel)
    if kernel.ndim != 3:
        raise ValueError("kernel must be 3-D")
    if not scaling:
        raise ValueError("scaling must be non-zero")
    created = created or datetime.now()
    user = created_by.encode()[:32]
    with open(filename, "wb") as f:
        np.asarray(kernel.shape, dtype="<i4").tofile(f)
        np.asarray([voxel_size, total_energy, scaling], dtype="<f4").tofile(f)
        np.asarray([created.year, created.month, created.day, created.hour, created.minute, created.second],
                   dtype="<i4").tofile(f)
        f.write(user + b"\x00" * (32 - len(user)))
        (kernel.astype(np.float64) / scaling).astype("<f4").tofile(f)


def save_dose_map(filename: Union[str, Path], dose_map: np.ndarray, voxel_size: Tuple[float, float, float],
                  metadata: Dict[str, Any], affine: Optional[np.ndarray] = None) -> str:
    """NIfTI-1 dose map as ``core/utils.py:53-111`` writes it: ``.nii.gz`` appended
    to other names, a default affine of the voxel sizes centred on the volume, float32
    data, ``pixdim`` = voxel sizes, the creation line in ``descrip`` and the metadata
    (updated with creation date, voxel size, dimensions, affine) as a code-44 JSON
    extension.  The creation date is the current time and ``created_by`` is kept from
    ``metadata`` (the reference writes fixed strings).  Returns the file name."""
    if hasattr(dose_map, "detach"):
        dose_map = dose_map.detach().cpu().numpy()
    dose_map = np.asarray(dose_map)
    filename = str(filename)
    if not filename.endswith((".nii", ".nii.gz")):
        filename += ".nii.gz"
    voxel_size = tuple(float(v) for v in voxel_size)
    if affine is None:
        affine = np.diag(list(voxel_size) + [1.0])
        affine[:3, 3] = np.array(dose_map.shape) * np.array(voxel_size) / -2.0
    affine = np.asarray(affine, dtype=np.float64)
    metadata.update({
        "creation_date": datetime.now().strftime("%Y-%m-%d %H:%M:%S"),
        "created_by": metadata.get("created_by", "pyvoxeldosimetry"),
        "voxel_size": voxel_size,
        "dimensions": tuple(int(s) for s in dose_map.shape),
        "affine_matrix": affine.tolist(),
    })
    descrip = f"Created by {metadata['created_by']} at {metadata['creation_date']}"
    ext = nifti.Extension(META_EXTENSION_CODE, json.dumps(metadata, default=_json_default).encode("utf-8"))
    nifti.save(filename, dose_map.astype(np.float32), affine, pixdim_xyz=voxel_size, descrip=descrip,
               extensions=[ext])
    return filename


def _json_default(o):
    if isinstance(o, np.ndarray):
        return o.tolist()
    if isinstance(o, np.generic):
        return o.item()
    raise TypeError(f"{type(o).__name__} is not JSON serialisable")


def load_dose_map(filename: Union[str, Path]) -> Tuple[np.ndarray, Dict[str, Any]]:
    """Dose array and metadata as ``core/utils.py:113-152`` returns them: voxel size
    from ``pixdim``, dimensions, affine, data type, description, then the code-44
    JSON extension merged over them (unreadable extensions are skipped)."""
    img = nifti.load(str(filename))
    metadata: Dict[str, Any] = {
        "voxel_size": tuple(float(v) for v in img.pixdim[1:4]),
        "dimensions": img.data.shape,
        "affine_matrix": img.affine.tolist(),
        "data_type": str(img.data.dtype),
        "description": img.descrip,
    }
    for ext in img.extensions:
        if ext.code == META_EXTENSION_CODE:
            try:
                metadata.update(json.loads(ext.content.rstrip(b"\x00").decode("utf-8")))
            except (ValueError, UnicodeDecodeError):
                pass
    return img.data, metadata


def verify_file_integrity(filename: Union[str, Path]) -> bool:
    """True for a NIfTI-1 file whose header parses (``.nii``, ``.gz``) or a ``.dat``
    kernel whose size matches its header and dimensions (``core/utils.py:264-301``)."""
    try:
        ext = Path(filename).suffix.lower()
        if ext in (".nii", ".gz"):
            with nifti._open(str(filename), "rb") as f:
                nifti.parse_header(f.read(nifti.HEADER_SIZE))
            return True
        if ext == ".dat":
            _dat_layout(filename)
            return True
        return False
    except Exception:
        return False


def get_file_metadata(filename: Union[str, Path]) -> Dict[str, Any]:
    """Metadata of a dose map (``.nii``, ``.gz``) or kernel (``.dat``) file
    (``core/utils.py:303-330``); ValueError for other suffixes."""
    ext = Path(filename).suffix.lower()
    if ext in (".nii", ".gz"):
        return load_dose_map(filename)[1]
    if ext == ".dat":
        return load_kernel(filename)[1]
    raise ValueError(f"Unsupported file format: {ext}")


def resample_dose_map(dose_map, input_spacing: Tuple[float, float, float],
                      output_spacing: Tuple[float, float, float]):
    """``resample_dose_map`` (``core/utils.py:193-231``): (resampled dose, affine) with an
    order-3 spline on the GPU; see ``pyvoxeldosimetry_amd.resample.resample_dose_map``."""
    from ..resample import resample_dose_map as _rdm
    return _rdm(dose_map, input_spacing, output_spacing)
