"""On-disk outputs in the reference's layout (IO/writeData.m, IO/writeNII.m,
IO/readNII.m, IO/generateNIIFileName.m).

* Volumes go to single-file NIfTI-1 (.nii): float32 data (datatype 16),
  column-major voxel order, 352-byte offset, the voxel spacing in pixdim,
  and the orientation as an sform (sform_code 1, srow_x/y/z = the first
  three rows of the 4x4 matrix), as make_nii/save_nii write them with the
  fields writeNII.m:35-45 sets.  A complex volume becomes a pair: the
  magnitude in <name>_<suffix>.nii and the phase (angle) in
  <name>_<suffix>Ph.nii; readNII.m multiplies them back.
* Motion goes to <name>_Tr.mat (writeData.m:52-55): variables T
  (single, [1 1 1 1 NStates 6]) and MotionInfo, in MATLAB's v5 MAT format.
* File names: <pathOu>/An-Ve/<Name>_<TypeName>, with the names of the
  reconstruction outputs Aq (uncorrected), Di (corrected) and Re (CS) and
  the geometry of generateNIIInformation (writeData.m:62-68): MT =
  diag(-1, -1, 1, 1) A, MS = the column norms of MT(1:3, 1:3).
"""
from __future__ import annotations

import gzip
import os
import struct

import numpy as np

TYPE_NAMES = {"x": "Aq", "d": "Di", "r": "Re", "T": "Tr"}
_HDR = 348
_OFFSET = 352


def generate_nii_information(A):
    """(MS [4], MT [4 x 4]) from the scanner's voxel-to-world matrix A."""
    MT = np.diag([-1.0, -1.0, 1.0, 1.0]) @ np.asarray(A, np.float64)
    MS = np.sqrt(np.sum(MT[:3, :3] ** 2, axis=0))
    return np.concatenate([MS, [1.0]]), MT


def _header(shape, MS, MT, glmax, glmin):
    dims = list(shape)
    dim = [len(dims)] + dims + [1] * (7 - len(dims))
    pixdim = [1.0] + [float(v) for v in MS[:3]] + [float(MS[3]) if len(MS) > 3 else 1.0] + [1.0] * 3
    h = bytearray(_HDR)
    struct.pack_into("<i", h, 0, _HDR)                   # sizeof_hdr
    h[38] = ord("r")                                      # regular
    struct.pack_into("<8h", h, 40, *dim)                  # dim
    struct.pack_into("<hhh", h, 70, 16, 32, 0)            # datatype float32, bitpix, slice_start
    struct.pack_into("<8f", h, 76, *pixdim)               # pixdim (qfac = 1)
    struct.pack_into("<f", h, 108, float(_OFFSET))        # vox_offset
    struct.pack_into("<ii", h, 140, int(glmax), int(glmin))
    struct.pack_into("<hh", h, 252, 0, 1)                 # qform_code, sform_code
    for r in range(3):
        struct.pack_into("<4f", h, 280 + 16 * r, *[float(v) for v in MT[r]])
    h[344:348] = b"n+1\x00"
    return bytes(h)


def write_nii(path, vol, MS=(1.0, 1.0, 1.0), MT=None):
    """One real volume (up to 7 dims, column-major) to a .nii file."""
    v = np.asarray(vol)
    if np.iscomplexobj(v):
        raise TypeError("write_nii takes real volumes; write_pair splits complex ones")
    v = v.astype(np.float32)
    MS = list(MS)
    if MT is None:
        MT = np.eye(4)
        for m in range(3):
            MT[m, m] = MS[m]
    fin = v[np.isfinite(v)]
    hi = int(np.round(fin.max())) if fin.size else 0
    lo = int(np.round(fin.min())) if fin.size else 0
    with open(path, "wb") as fh:
        fh.write(_header(v.shape, MS, np.asarray(MT, np.float64), hi, lo))
        fh.write(b"\x00" * (_OFFSET - _HDR))
        fh.write(v.reshape(-1, order="F").tobytes())


def read_nii(path):
    """(img, MS [3], MT [4 x 4]) of a .nii or .nii.gz float32/float64 file."""
    op = gzip.open if path.endswith(".gz") else open
    with op(path, "rb") as fh:
        raw = fh.read()
    if struct.unpack_from("<i", raw, 0)[0] != _HDR:
        raise ValueError(f"{path}: not a little-endian NIfTI-1 file")
    dim = struct.unpack_from("<8h", raw, 40)
    dtype = {16: np.float32, 64: np.float64}.get(struct.unpack_from("<h", raw, 70)[0])
    if dtype is None:
        raise ValueError(f"{path}: only float32/float64 data are read")
    pixdim = struct.unpack_from("<8f", raw, 76)
    off = int(struct.unpack_from("<f", raw, 108)[0])
    shape = dim[1:1 + dim[0]]
    n = int(np.prod(shape))
    img = np.frombuffer(raw, dtype=dtype, count=n, offset=off).reshape(shape, order="F")
    MT = np.eye(4)
    for r in range(3):
        MT[r] = struct.unpack_from("<4f", raw, 280 + 16 * r)
    return img, np.array(pixdim[1:4]), MT


def write_pair(stem, suffix, x, MS=(1.0, 1.0, 1.0), MT=None):
    """writeNII.m:29-45 for one volume: <stem>_<suffix>.nii (magnitude, or
    the real volume) and, when x has an imaginary part, <stem>_<suffix>Ph.nii
    (phase).  Returns the paths written."""
    x = np.asarray(x)
    M = list(x.shape) + [1] * (3 - x.ndim)
    rest = int(np.prod(M[3:])) if len(M) > 3 else 1
    x = x.reshape(M[:3] + ([rest] if rest > 1 else []), order="F")  # writeNII.m:31
    x32 = x.astype(np.complex64) if np.iscomplexobj(x) else x.astype(np.float32)
    out = []
    if np.iscomplexobj(x32) and np.any(x32.imag != 0):
        p = f"{stem}_{suffix}Ph.nii"
        write_nii(p, np.angle(x32), MS, MT)
        out.append(p)
        x32 = np.abs(x32)
    p = f"{stem}_{suffix}.nii"
    write_nii(p, np.real(x32), MS, MT)
    out.append(p)
    return out


def read_pair(stem, suffix):
    """readNII.m:17-38: the volume, times exp(i phase) when a phase file exists."""
    for ext in (".nii", ".nii.gz"):
        p = f"{stem}_{suffix}{ext}"
        if os.path.exists(p):
            break
    else:
        raise FileNotFoundError(f"{stem}_{suffix}.nii")
    x, MS, MT = read_nii(p)
    ph = f"{stem}_{suffix}Ph{ext}"
    if os.path.exists(ph):
        x = x * np.exp(1j * read_nii(ph)[0])
    return x, MS, MT


def output_stem(path_out, name):
    """generateNIIFileName.m: <path_out>/An-Ve/<name, last 225 characters>."""
    d = os.path.join(path_out, "An-Ve")
    os.makedirs(d, exist_ok=True)
    return os.path.join(d, name[-225:])


def write_data(path_out, name, volumes, T=None, motion_info=None, A=None):
    """writeData.m: every entry of volumes ({'x': uncorrected, 'd':
    corrected, 'r': regularised}; arrays, complex allowed, trailing dims
    stacked) as NIfTI pairs named by TYPE_NAMES, and the motion T [S x 6] as
    <name>_Tr.mat.  A is the voxel-to-world matrix (identity by default).
    Returns the paths written."""
    from scipy.io import savemat
    stem = output_stem(path_out, name)
    MS, MT = generate_nii_information(np.eye(4) if A is None else A)
    paths = []
    for key, vol in volumes.items():
        if key not in TYPE_NAMES or key == "T":
            raise ValueError(f"unknown output type {key!r}")
        paths += write_pair(stem, TYPE_NAMES[key], vol, MS, MT)
    if T is not None:
        T = np.asarray(T, np.float32)
        p = f"{stem}_{TYPE_NAMES['T']}.mat"
        savemat(p, {"T": T.reshape((1, 1, 1, 1) + T.shape),
                    "MotionInfo": np.zeros((0, 0)) if motion_info is None else motion_info},
                do_compression=True)
        paths.append(p)
    return paths
