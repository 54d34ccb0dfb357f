"""Kernel (.dat) and dose-map (NIfTI-1) files: layouts pinned by hand-packed bytes.

The ``.dat`` tests pack the header in the order the reference's reader reads it
(``pyvoxeldosimetry/core/utils.py:29-38``).  The NIfTI tests check the written
header at the byte offsets of the published ``nifti1.h`` layout and read a file
assembled by hand (qform quaternion of a 90-degree rotation about z, intensity
scaling, big-endian, int16), so neither side is checked only against itself.
nibabel is not installed, so no file written by nibabel is available.
"""
import gzip
import json
import math
import struct

import numpy as np
import pytest

from pyvoxeldosimetry_amd.core import utils
from pyvoxeldosimetry_amd.io import nifti


def _dat_bytes(arr, voxel, energy, scaling, ts=(2025, 2, 8, 9, 50, 56), user=b"devhliu", short=False):
    hdr = struct.pack("<3i3f", *arr.shape, voxel, energy, scaling)
    if not short:
        hdr += struct.pack("<6i", *ts) + user.ljust(32, b"\x00")
    return hdr + np.asarray(arr, dtype="<f4").tobytes(order="C")


def test_dat_reader_follows_the_reference_header(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.random((4, 5, 6)).astype(np.float32)
    p = tmp_path / "kernel_2mm.dat"
    p.write_bytes(_dat_bytes(a, 2.0, 0.1337, 0.5))
    k, m = utils.load_kernel(p)
    assert k.dtype == np.float32 and k.shape == (4, 5, 6)
    np.testing.assert_array_equal(k, a * np.float32(0.5))
    assert list(m["dimensions"]) == [4, 5, 6]
    assert m["voxel_size"] == np.float32(2.0) and m["total_energy"] == np.float32(0.1337)
    assert m["scaling_factor"] == np.float32(0.5)
    assert m["creation_date"] == "2025-02-08 09:50:56" and m["created_by"] == "devhliu"
    assert utils.verify_file_integrity(p)
    assert utils.get_file_metadata(p)["created_by"] == "devhliu"


def test_dat_reader_accepts_the_readme_short_header(tmp_path):
    a = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    p = tmp_path / "k.dat"
    p.write_bytes(_dat_bytes(a, 1.0, 0.5, 2.0, short=True))
    k, m = utils.load_kernel(p)
    np.testing.assert_array_equal(k, 2.0 * a)
    assert m["creation_date"] is None and m["created_by"] == ""
    assert utils.verify_file_integrity(p)


def test_dat_writer_round_trip_and_integrity_failures(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.random((7, 7, 7))
    p = tmp_path / "w.dat"
    utils.save_kernel(p, a, 3.0, total_energy=0.2, scaling=1e-3, created_by="gpu")
    k, m = utils.load_kernel(p)
    np.testing.assert_allclose(k, a, rtol=2e-7)
    assert m["created_by"] == "gpu" and m["voxel_size"] == np.float32(3.0)
    raw = p.read_bytes()
    (tmp_path / "trunc.dat").write_bytes(raw[:-4])
    assert not utils.verify_file_integrity(tmp_path / "trunc.dat")
    with pytest.raises(ValueError):
        utils.load_kernel(tmp_path / "trunc.dat")
    # the size the reference's check expects, (N + 11) * 4 bytes, fits no reader layout
    (tmp_path / "ref44.dat").write_bytes(raw[:44] + raw[80:])
    assert not utils.verify_file_integrity(tmp_path / "ref44.dat")
    assert not utils.verify_file_integrity(tmp_path / "missing.dat")
    (tmp_path / "x.txt").write_text("x")
    assert not utils.verify_file_integrity(tmp_path / "x.txt")


def test_nifti_header_at_spec_offsets(tmp_path):
    dm = np.arange(3 * 4 * 5, dtype=np.float32).reshape(3, 4, 5)
    aff = np.array([[2.0, 0, 0, -3], [0, 2.5, 0, -5], [0, 0, 3.0, -7.5], [0, 0, 0, 1]])
    meta = {"nuclide": "Y90", "created_by": "test"}
    f = utils.save_dose_map(tmp_path / "d.nii", dm, (2.0, 2.5, 3.0), meta, affine=aff)
    raw = open(f, "rb").read()
    u = lambda fmt, off: struct.unpack_from("<" + fmt, raw, off)  # noqa: E731
    assert u("i", 0) == (348,) and raw[344:348] == b"n+1\x00"
    assert u("8h", 40) == (3, 3, 4, 5, 1, 1, 1, 1)
    assert u("h", 70) == (16,) and u("h", 72) == (32,)                 # float32
    assert u("3f", 80) == (2.0, 2.5, 3.0)                               # pixdim[1:4]
    assert u("h", 254) == (2,)                                          # sform_code aligned
    np.testing.assert_allclose(u("4f", 280), aff[0]); np.testing.assert_allclose(u("4f", 312), aff[2])
    assert raw[348] == 1                                                # extension follows
    esize, ecode = u("2i", 352)
    assert ecode == 44 and esize % 16 == 0
    vox = int(u("f", 108)[0])
    assert vox == 352 + esize
    ext = json.loads(raw[360:352 + esize].rstrip(b"\x00"))
    assert ext["nuclide"] == "Y90" and ext["dimensions"] == [3, 4, 5]
    first = np.frombuffer(raw, dtype="<f4", count=3, offset=vox)
    assert list(first) == [dm[0, 0, 0], dm[1, 0, 0], dm[2, 0, 0]]       # first index fastest
    assert raw[148:148 + 15] == b"Created by test"
    d, m = utils.load_dose_map(f)
    np.testing.assert_array_equal(d, dm)
    np.testing.assert_allclose(np.array(m["affine_matrix"]), aff)
    assert m["nuclide"] == "Y90" and list(m["voxel_size"]) == [2.0, 2.5, 3.0]


def test_nifti_default_affine_and_gzip(tmp_path):
    dm = np.random.default_rng(2).random((6, 4, 2)).astype(np.float32)
    f = utils.save_dose_map(tmp_path / "dose", dm, (1.0, 2.0, 4.0), {})
    assert f.endswith(".nii.gz")
    with gzip.open(f, "rb") as g:
        assert struct.unpack("<i", g.read(4))[0] == 348
    d, m = utils.load_dose_map(f)
    np.testing.assert_array_equal(d, dm)
    np.testing.assert_allclose(np.array(m["affine_matrix"])[:3, 3], [-3.0, -4.0, -4.0])   # centred
    assert nifti.is_nifti_file(f) and utils.verify_file_integrity(f)
    assert not nifti.is_nifti_file(str(tmp_path / "none.nii"))
    assert utils.get_file_metadata(f)["dimensions"] == [6, 4, 2]
    with pytest.raises(ValueError):
        utils.get_file_metadata(tmp_path / "a.txt")


def _hand_nifti(data, endian, datatype, quat, qoff, pixdim, slope, inter):
    hdr = bytearray(348)
    p = lambda fmt, off, *v: struct.pack_into(endian + fmt, hdr, off, *v)  # noqa: E731
    p("i", 0, 348)
    p("8h", 40, 3, *data.shape, 1, 1, 1, 1)
    p("h", 70, datatype)
    p("h", 72, data.dtype.itemsize * 8)
    p("8f", 76, *pixdim)
    p("f", 108, 352.0)
    p("2f", 112, slope, inter)
    p("h", 252, 1)                 # qform_code scanner, sform_code 0
    p("3f", 256, *quat)
    p("3f", 268, *qoff)
    hdr[344:348] = b"n+1\x00"
    body = np.asarray(data, dtype=data.dtype.newbyteorder(endian)).tobytes(order="F")
    return bytes(hdr) + b"\x00" * 4 + body


@pytest.mark.parametrize("endian", ["<", ">"])
def test_nifti_reader_on_hand_assembled_qform_file(tmp_path, endian):
    data = np.arange(2 * 3 * 4, dtype=np.int16).reshape(2, 3, 4)
    s = math.sqrt(0.5)                                        # 90 degrees about z: (b, c, d) = (0, 0, sin 45)
    raw = _hand_nifti(data, endian, 4, (0.0, 0.0, s), (10.0, 20.0, 30.0), (1.0, 2.0, 3.0, 4.0, 1, 1, 1, 1),
                      2.0, 1.0)
    f = tmp_path / "h.nii"
    f.write_bytes(raw)
    img = nifti.load(str(f))
    np.testing.assert_array_equal(img.data, data * 2.0 + 1.0)           # scl_slope / scl_inter
    rot = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    np.testing.assert_allclose(img.affine[:3, :3], rot * np.array([2.0, 3.0, 4.0]), atol=1e-6)
    np.testing.assert_allclose(img.affine[:3, 3], [10.0, 20.0, 30.0])
    # qfac = -1 flips the third column
    raw2 = _hand_nifti(data, endian, 4, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (-1.0, 1.0, 1.0, 1.0, 1, 1, 1, 1),
                       0.0, 0.0)
    f.write_bytes(raw2)
    img2 = nifti.load(str(f))
    np.testing.assert_allclose(np.diag(img2.affine), [1.0, 1.0, -1.0, 1.0])
    np.testing.assert_array_equal(img2.data, data)                     # slope 0 = no scaling


def test_nifti_rejects_non_nifti(tmp_path):
    f = tmp_path / "bad.nii"
    f.write_bytes(b"\x00" * 400)
    with pytest.raises(ValueError):
        nifti.load(str(f))
    assert not nifti.is_nifti_file(str(f))
    assert not utils.verify_file_integrity(f)
