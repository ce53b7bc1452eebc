"""PSF MAT-file loader.

The reference keeps the light-field PSF in a MAT-file whose variables ``H``
([M, M, Nnum, Nnum, Nz] single or double), ``Nnum`` and ``CAindex`` ([Nz, 2]
inclusive 1-based crop rows per plane) it reads through ``matfile``
(sid_main.m:30,33,137; reconstruct_S.m:118-121 for the CPU path with CAindex,
reconstruct_S.m:155 for the GPU path, which takes H only).  The file is
written with ``save -v7.3`` / ``h5create`` + ``h5write``
(external/save_psf_fast.m:39-74), i.e. HDF5 behind a 512-byte user block;
older MAT formats (v4 / v5 / v7) are read through scipy.io.loadmat.

HDF5 stores MATLAB's column-major arrays with the dimensions reversed, so a
variable read from HDF5 in C order is transposed here: the result has
MATLAB's shape and column-major (Fortran) order, as the C ABI expects for H.
"""
from __future__ import annotations

import numpy as np

from . import h5read

_NUMERIC = {"double", "single", "int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64",
            "logical"}


def is_hdf5(path: str) -> bool:
    """True when the file carries an HDF5 signature at 0, 512, 1024, ... bytes."""
    with open(path, "rb") as fh:
        fh.seek(0, 2)
        size = fh.tell()
        pos = 0
        while pos + 8 <= size:
            fh.seek(pos)
            if fh.read(8) == h5read.SIGNATURE:
                return True
            pos = 512 if pos == 0 else pos * 2
    return False


def read_mat(path: str, names=None) -> dict:
    """Numeric variables of a MAT-file as numpy arrays in MATLAB's shape and
    column-major order.  ``names`` selects variables (KeyError when one is
    missing); by default every numeric variable is returned.  Cell arrays,
    structs and char variables are skipped (or rejected when named)."""
    if not is_hdf5(path):
        import scipy.io
        kw = {} if names is None else {"variable_names": list(names)}
        raw = scipy.io.loadmat(path, squeeze_me=False, **kw)
        out = {k: v for k, v in raw.items() if not k.startswith("__") and isinstance(v, np.ndarray)
               and v.dtype.kind in "fiub"}
        for n in names or ():
            if n not in out:
                raise KeyError(f"{path}: no numeric variable {n!r}")
        return out
    f = h5read.File(path)
    keys = [k for k in f.keys() if not k.startswith("#")]
    out = {}
    for name in (names if names is not None else keys):
        if name not in keys:
            raise KeyError(f"{path}: no variable {name!r} (has {', '.join(keys)})")
        if not f.is_dataset(name):
            if names is not None:
                raise ValueError(f"{path}: variable {name!r} is a struct / object, not a numeric array")
            continue
        ds = f.dataset(name)
        cls = ds.attrs.get("MATLAB_class")
        if cls is not None and cls not in _NUMERIC:
            if names is not None:
                raise ValueError(f"{path}: variable {name!r} has MATLAB class {cls!r}, not a numeric array")
            continue
        if int(np.asarray(ds.attrs.get("MATLAB_empty", 0)).ravel()[0]):
            # an empty array is stored as its dimension vector
            dims = tuple(int(d) for d in np.asarray(ds.read()).ravel())
            out[name] = np.zeros(dims, dtype=np.float64 if cls in (None, "double") else np.float32, order="F")
            continue
        arr = ds.read()
        if ds.dtype.kind not in "fiu":
            if names is not None:
                raise ValueError(f"{path}: variable {name!r} is not numeric ({ds.dtype})")
            continue
        out[name] = arr.T   # reversed axes: MATLAB shape, column-major
    return out


def load_psf(path: str):
    """(H, Nnum, CAindex) of a PSF MAT-file.

    H: float32 [M, M, Nnum, Nnum, Nz], column-major (a single plane stored as
    4-D gets its trailing singleton back).  Nnum: int (H's lenslet size when the
    file has no Nnum).  CAindex: int32 [Nz, 2] 1-based inclusive crop rows, or
    None when absent.  Checks that H has square kernels and square lenslets,
    that Nnum agrees, and that every crop lies inside 1..M."""
    present = set(read_mat_names(path))
    if "H" not in present:
        raise KeyError(f"{path}: no variable 'H' (the PSF)")
    v = read_mat(path, [n for n in ("H", "Nnum", "CAindex") if n in present])
    H = v["H"]
    if H.ndim == 4:
        H = H[..., None]
    if H.ndim != 5 or H.shape[0] != H.shape[1] or H.shape[2] != H.shape[3]:
        raise ValueError(f"{path}: H must be [M, M, Nnum, Nnum, Nz] with square kernels and lenslets, got {H.shape}")
    H = np.asarray(H, dtype=np.float32, order="F")
    nnum = H.shape[2]
    if "Nnum" in v:
        val = float(np.asarray(v["Nnum"]).ravel()[0])
        if val != round(val) or int(round(val)) != nnum:
            raise ValueError(f"{path}: Nnum = {val} but H has {nnum} x {nnum} lenslet offsets")
    ca = None
    if "CAindex" in v:
        c = np.asarray(v["CAindex"], dtype=np.float64)
        nz, M = H.shape[4], H.shape[0]
        if c.shape != (nz, 2) or np.any(c != np.round(c)):
            raise ValueError(f"{path}: CAindex must be an integer [Nz, 2] = [{nz}, 2] array, got {c.shape}")
        if np.any(c[:, 0] < 1) or np.any(c[:, 1] > M) or np.any(c[:, 1] < c[:, 0]):
            raise ValueError(f"{path}: CAindex rows must satisfy 1 <= lo <= hi <= M = {M}")
        ca = np.asarray(c, dtype=np.int32)
    return H, nnum, ca


def read_mat_names(path: str):
    """Variable names of a MAT-file (HDF5 links of the root group, or the
    v4 / v5 directory through scipy.io.whosmat)."""
    if is_hdf5(path):
        return [k for k in h5read.File(path).keys() if not k.startswith("#")]
    import scipy.io
    return [w[0] for w in scipy.io.whosmat(path)]
