"""PSF MAT-file loading: the pure-Python HDF5 reader and the MATLAB layer.

Pinned against the HDF5 C library (through h5py, in tests/golden/make_h5_fixtures.py):
fixture files in MATLAB's -v7.3 conventions with their values, and checksums of
every numeric dataset h5py reads in HDF5 files written by other programs in this
image (a MATLAB-written file shipped with scipy, PyTables' test files).  The
MATLAB-written file's variable is also checked against its closed form."""
import hashlib
import json
import os

import numpy as np
import pytest

from sid_amd import h5read, psf_io

GOLD = os.path.join(os.path.dirname(__file__), "golden", "h5")
FIXTURES = ["psf_v73_contiguous.mat", "psf_v73_chunked.mat", "psf_latest.mat", "structures.h5"]
MATLAB_FILE = "/usr/local/lib/python3.10/dist-packages/scipy/io/matlab/tests/data/testhdf5_7.4_GLNX86.mat"


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_values(name):
    path = os.path.join(GOLD, name)
    exp = np.load(path + ".npz")
    f = h5read.File(path)
    for key in exp.files:
        got = f.read(key.replace("__", "/"))
        want = exp[key]
        assert got.dtype == want.dtype and got.shape == want.shape, key
        assert np.array_equal(got, want), key


def test_fixture_structure():
    f = h5read.File(os.path.join(GOLD, "structures.h5"))
    keys = f.keys()
    assert len([k for k in keys if k.startswith("v")]) == 40 and "grp" in keys
    assert f.visit("grp") == sorted("grp/sub/" + k for k in f.keys("grp/sub"))
    assert f.read("scalar").shape == () and float(f.read("scalar")) == 2.5
    assert h5read.File(os.path.join(GOLD, "psf_latest.mat")).sb_version == 3
    assert h5read.File(os.path.join(GOLD, "psf_v73_chunked.mat")).sb == 512


def test_external_files_match_the_hdf5_library():
    ext = json.load(open(os.path.join(GOLD, "external_checksums.json")))
    present = {p: e for p, e in ext.items() if os.path.exists(p)}
    if not present:
        pytest.skip("none of the external HDF5 files is installed here")
    n = 0
    for path, entries in present.items():
        f = h5read.File(path)
        for name, d in entries.items():
            a = np.ascontiguousarray(f.read(name))
            h = hashlib.sha256(a.dtype.str.encode() + repr(a.shape).encode() + a.tobytes()).hexdigest()
            assert (a.dtype.str, list(a.shape), h) == (d["dtype"], d["shape"], d["sha256"]), (path, name)
            n += 1
    assert n >= 50


def test_matlab_written_file():
    if not os.path.exists(MATLAB_FILE):
        pytest.skip("scipy's MATLAB -v7.3 test file is not installed here")
    v = psf_io.read_mat(MATLAB_FILE)
    assert list(v) == ["testdouble"]
    assert v["testdouble"].shape == (1, 9)                       # MATLAB 0:pi/4:2*pi, a row
    assert np.array_equal(v["testdouble"].ravel(), np.arange(9) * np.pi / 4)


@pytest.mark.parametrize("name,dtype", [("psf_v73_contiguous.mat", np.float32), ("psf_v73_chunked.mat", np.float64),
                                        ("psf_latest.mat", np.float32)])
def test_load_psf(name, dtype):
    path = os.path.join(GOLD, name)
    exp = np.load(path + ".npz")
    H, nnum, ca = psf_io.load_psf(path)
    want = exp["H"].T                                            # back to MATLAB's [M, M, n, n, nz]
    assert want.dtype == dtype
    assert H.dtype == np.float32 and H.flags.f_contiguous and H.shape == want.shape
    assert np.array_equal(H, want.astype(np.float32))
    assert nnum == want.shape[2]
    assert ca.dtype == np.int32 and np.array_equal(ca, exp["CAindex"].T)
    assert {"H", "Nnum", "CAindex"} <= set(psf_io.read_mat_names(path))


def test_load_psf_older_mat_formats(tmp_path):
    import scipy.io
    rng = np.random.default_rng(1)
    H = rng.random((5, 5, 3, 3, 2)).astype(np.float32)
    ca = np.array([[1, 5], [2, 4]], dtype=np.float64)
    p = str(tmp_path / "psf_v5.mat")
    scipy.io.savemat(p, {"H": H, "Nnum": 3.0, "CAindex": ca})
    assert not psf_io.is_hdf5(p)
    H2, nnum, ca2 = psf_io.load_psf(p)
    assert np.array_equal(H2, H) and nnum == 3 and np.array_equal(ca2, ca.astype(np.int32))
    # a single plane is stored as 4-D by MATLAB
    p1 = str(tmp_path / "one_plane.mat")
    scipy.io.savemat(p1, {"H": H[..., 0]})
    H1, nnum1, ca1 = psf_io.load_psf(p1)
    assert H1.shape == (5, 5, 3, 3, 1) and nnum1 == 3 and ca1 is None


def test_load_psf_errors(tmp_path):
    import scipy.io
    H = np.zeros((5, 5, 3, 3, 2), dtype=np.float32)
    cases = {
        "no_h.mat": {"Nnum": 3.0},
        "bad_nnum.mat": {"H": H, "Nnum": 5.0},
        "bad_shape.mat": {"H": np.zeros((5, 4, 3, 3, 2), np.float32)},
        "bad_ca.mat": {"H": H, "CAindex": np.array([[1, 6], [1, 5]], np.float64)},
        "bad_ca_shape.mat": {"H": H, "CAindex": np.array([[1, 5]], np.float64)},
    }
    for fn, d in cases.items():
        p = str(tmp_path / fn)
        scipy.io.savemat(p, d)
        with pytest.raises((KeyError, ValueError)):
            psf_io.load_psf(p)
    with pytest.raises(h5read.H5Error):
        h5read.File(os.path.join(GOLD, "extensible.h5")).read("grow")   # extensible-array index
    # a truncated HDF5 file fails with H5Error, never with a silent short read
    raw = open(os.path.join(GOLD, "psf_v73_chunked.mat"), "rb").read()
    t = tmp_path / "truncated.mat"
    t.write_bytes(raw[:3000])
    with pytest.raises(h5read.H5Error):
        psf_io.read_mat(str(t), ["H"])
    g = tmp_path / "garbage.bin"
    g.write_bytes(b"not a mat file" * 10)
    with pytest.raises(Exception):
        psf_io.load_psf(str(g))
