"""On-disk outputs (IO/writeData.m, IO/writeNII.m, IO/readNII.m), CPU only.

The header is checked by reading each NIfTI-1 field at its byte offset in
nifti1.h (a reader independent of the writer's sequential packing); the
data by its column-major float32 bytes at offset 352; complex volumes by
the magnitude/phase pair multiplying back; the motion file by
scipy.io.loadmat."""
import os
import struct

import numpy as np
import pytest

from disorder_amd import nifti as NI

FIELDS = {  # name: (offset, format) from nifti1.h
    "sizeof_hdr": (0, "<i"), "dim": (40, "<8h"), "datatype": (70, "<h"), "bitpix": (72, "<h"),
    "pixdim": (76, "<8f"), "vox_offset": (108, "<f"), "scl_slope": (112, "<f"), "qform_code": (252, "<h"),
    "sform_code": (254, "<h"), "srow_x": (280, "<4f"), "srow_y": (296, "<4f"), "srow_z": (312, "<4f"),
    "magic": (344, "4s"),
}


def _fields(path):
    raw = open(path, "rb").read()
    return {k: struct.unpack_from(f, raw, o) for k, (o, f) in FIELDS.items()}, raw


def test_header_and_data_layout(tmp_path):
    rng = np.random.default_rng(0)
    v = rng.normal(size=(6, 5, 4)).astype(np.float32)
    MT = np.array([[-2.0, 0, 0, 10], [0, -3, 0, 20], [0, 0, 4, -5], [0, 0, 0, 1]])
    p = str(tmp_path / "a.nii")
    NI.write_nii(p, v, [2.0, 3.0, 4.0], MT)
    f, raw = _fields(p)
    assert f["sizeof_hdr"] == (348,) and f["magic"] == (b"n+1\x00",)
    assert f["dim"] == (3, 6, 5, 4, 1, 1, 1, 1)
    assert f["datatype"] == (16,) and f["bitpix"] == (32,) and f["vox_offset"] == (352.0,)
    assert f["pixdim"][1:4] == (2.0, 3.0, 4.0) and f["scl_slope"] == (0.0,)
    assert f["qform_code"] == (0,) and f["sform_code"] == (1,)
    assert f["srow_x"] == (-2.0, 0.0, 0.0, 10.0) and f["srow_z"] == (0.0, 0.0, 4.0, -5.0)
    assert len(raw) == 352 + v.size * 4
    assert np.array_equal(np.frombuffer(raw, np.float32, offset=352), v.reshape(-1, order="F"))
    img, MS, MT2 = NI.read_nii(p)
    assert np.array_equal(img, v) and np.allclose(MS, [2, 3, 4]) and np.allclose(MT2, MT)


def test_default_orientation_is_spacing(tmp_path):
    p = str(tmp_path / "b.nii")
    NI.write_nii(p, np.zeros((2, 2, 2)), [1.5, 2.0, 0.5])
    f, _ = _fields(p)
    assert f["srow_x"] == (1.5, 0, 0, 0) and f["srow_y"] == (0, 2.0, 0, 0) and f["srow_z"] == (0, 0, 0.5, 0)


def test_complex_pair_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    x = (rng.normal(size=(5, 4, 3)) + 1j * rng.normal(size=(5, 4, 3))).astype(np.complex64)
    stem = str(tmp_path / "rec")
    paths = NI.write_pair(stem, "Di", x)
    assert [os.path.basename(q) for q in paths] == ["rec_DiPh.nii", "rec_Di.nii"]
    mag = NI.read_nii(stem + "_Di.nii")[0]
    ph = NI.read_nii(stem + "_DiPh.nii")[0]
    assert np.array_equal(mag, np.abs(x)) and np.array_equal(ph, np.angle(x))
    back, _, _ = NI.read_pair(stem, "Di")
    assert np.abs(back - x).max() <= 4e-7 * np.abs(x).max()
    # a real volume (or a complex one with zero imaginary part) writes one file
    assert len(NI.write_pair(stem, "Aq", np.abs(x).astype(np.complex64))) == 1


def test_trailing_dims_are_stacked(tmp_path):
    x = np.arange(4 * 3 * 2 * 2 * 3, dtype=np.float32).reshape((4, 3, 2, 2, 3), order="F")
    stem = str(tmp_path / "s")
    NI.write_pair(stem, "Di", x)
    f, _ = _fields(stem + "_Di.nii")
    assert f["dim"][:5] == (4, 4, 3, 2, 6)
    img = NI.read_nii(stem + "_Di.nii")[0]
    assert np.array_equal(img.reshape(-1, order="F"), x.reshape(-1, order="F"))
    NI.write_pair(stem, "Aq", x[..., :1, :1])
    assert _fields(stem + "_Aq.nii")[0]["dim"][0] == 3


def test_write_data_layout(tmp_path):
    from scipy.io import loadmat
    rng = np.random.default_rng(2)
    x = (rng.normal(size=(4, 4, 3)) + 1j * rng.normal(size=(4, 4, 3))).astype(np.complex64)
    T = rng.normal(size=(7, 6))
    A = np.array([[0.0, 1.2, 0, -3], [1.1, 0, 0, 4], [0, 0, 2.0, 5], [0, 0, 0, 1]])
    paths = NI.write_data(str(tmp_path), "scan01", {"x": x, "d": np.abs(x)}, T=T, A=A)
    names = sorted(os.path.basename(p) for p in paths)
    assert names == ["scan01_Aq.nii", "scan01_AqPh.nii", "scan01_Di.nii", "scan01_Tr.mat"]
    assert all(os.path.dirname(p) == str(tmp_path / "An-Ve") for p in paths)
    MS, MT = NI.generate_nii_information(A)
    assert np.allclose(MT[0], [0, -1.2, 0, 3]) and np.allclose(MT[1], [-1.1, 0, 0, -4]) and np.allclose(MT[2], A[2])
    assert np.allclose(MS, [1.1, 1.2, 2.0, 1.0])
    img, ms, mt = NI.read_nii(str(tmp_path / "An-Ve" / "scan01_Di.nii"))
    assert np.allclose(ms, MS[:3]) and np.allclose(mt[:3], MT[:3], atol=1e-6)
    m = loadmat(str(tmp_path / "An-Ve" / "scan01_Tr.mat"))
    assert m["T"].dtype == np.float32 and m["T"].shape == (1, 1, 1, 1, 7, 6)
    assert np.array_equal(m["T"][0, 0, 0, 0], T.astype(np.float32))
    with pytest.raises(ValueError):
        NI.write_data(str(tmp_path), "bad", {"q": x})
