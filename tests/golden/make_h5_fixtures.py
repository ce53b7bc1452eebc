"""Writes the HDF5 / MATLAB -v7.3 fixtures of tests/test_psf_io.py.

Run with an interpreter that has h5py (the HDF5 C library does the writing and
the reading-back here, independently of sid_amd.h5read):

    /opt/conda/bin/python3.9 tests/golden/make_h5_fixtures.py

1. Small files in MATLAB's -v7.3 conventions (512-byte user block with a
   MATLAB header, dimensions reversed, MATLAB_class attributes) covering both
   HDF5 format generations h5py can write (libver 'earliest': superblock 0,
   version-1 object headers, symbol-table groups, version-3 layouts with
   version-1 B-tree chunk indexes; libver 'latest': superblock 3, version-2
   object headers, link messages, version-4 layouts with single-chunk,
   implicit and fixed-array chunk indexes), compact / contiguous / chunked
   storage, deflate + shuffle + fletcher32, edge chunks, a chunk B-tree of
   more than one level and a group of more than one symbol-table node.  The
   values are written to <file>.npz (C order, as HDF5 stores them).
2. Checksums (sha256 of dtype, shape and bytes as h5py reads them) of every
   numeric dataset h5py can read in HDF5 files written by other programs that
   ship in this image: scipy's MATLAB-written test file and PyTables' test
   files, to external_checksums.json.
"""
import glob
import hashlib
import json
import os
import time

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "h5")


def matlab_userblock(path):
    text = ("MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: %s HDF5 schema 1.00 ." %
            time.strftime("%a %b %d %H:%M:%S %Y", time.gmtime(0))).ljust(116).encode()
    head = text + b"\0" * 8 + b"\x00\x02" + b"IM"
    with open(path, "r+b") as fh:
        fh.write(head.ljust(512, b"\0"))


def put(f, name, mat_value, cls, expect, **kw):
    """Write a MATLAB-shaped array (reversed dims, C order) with MATLAB_class."""
    data = np.ascontiguousarray(np.asarray(mat_value).T)
    ds = f.create_dataset(name, data=data, **kw)
    ds.attrs.create("MATLAB_class", np.bytes_(cls))
    expect[name] = data
    return ds


def put_compact(f, name, mat_value, cls, expect):
    data = np.ascontiguousarray(np.asarray(mat_value).T)
    space = h5py.h5s.create_simple(data.shape)
    dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
    dcpl.set_layout(h5py.h5d.COMPACT)
    tid = h5py.h5t.py_create(data.dtype)
    dsid = h5py.h5d.create(f.id, name.encode(), tid, space, dcpl=dcpl)
    dsid.write(h5py.h5s.ALL, h5py.h5s.ALL, data)
    f[name].attrs.create("MATLAB_class", np.bytes_(cls))
    expect[name] = data


def put_implicit(f, name, data, chunks, expect):
    space = h5py.h5s.create_simple(data.shape)
    dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
    dcpl.set_chunk(chunks)
    dcpl.set_alloc_time(h5py.h5d.ALLOC_TIME_EARLY)
    tid = h5py.h5t.py_create(data.dtype)
    dsid = h5py.h5d.create(f.id, name.encode(), tid, space, dcpl=dcpl)
    dsid.write(h5py.h5s.ALL, h5py.h5s.ALL, np.ascontiguousarray(data))
    expect[name] = data


def psf_values(rng, M, n, nz, dtype):
    H = rng.random((M, M, n, n, nz)).astype(dtype)
    H[rng.random(H.shape) < 0.2] = 0
    ca = np.array([[1 + c, M - c] for c in range(nz)], dtype=np.float64)
    return H, ca


def save(path, expect, userblock):
    if userblock:
        matlab_userblock(path)
    np.savez(path + ".npz", **{k.replace("/", "__"): v for k, v in expect.items()})
    with h5py.File(path, "r") as f:   # read back through the library
        for k, v in expect.items():
            got = f[k][()]
            assert got.dtype == v.dtype and got.shape == v.shape and np.array_equal(got, v), k


def main():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(20260101)

    # 1. earliest format, contiguous single-precision PSF (what save_psf_fast's h5create/h5write makes)
    p = os.path.join(OUT, "psf_v73_contiguous.mat")
    e = {}
    with h5py.File(p, "w", userblock_size=512, libver="earliest") as f:
        H, ca = psf_values(rng, 9, 3, 3, np.float32)
        put(f, "H", H, "single", e)
        put(f, "Nnum", np.array([[3.0]]), "double", e)
        put(f, "CAindex", ca, "double", e)
    save(p, e, True)

    # 2. earliest format, chunked double PSF with shuffle + deflate (save -v7.3 compresses), edge
    #    chunks, compact CAindex, a fletcher32 dataset
    p = os.path.join(OUT, "psf_v73_chunked.mat")
    e = {}
    with h5py.File(p, "w", userblock_size=512, libver="earliest") as f:
        H, ca = psf_values(rng, 11, 3, 4, np.float64)
        put(f, "H", H, "double", e, chunks=(2, 2, 3, 4, 11), compression="gzip", compression_opts=4, shuffle=True)
        put(f, "Nnum", np.array([[3.0]]), "double", e)
        put_compact(f, "CAindex", ca, "double", e)
        put(f, "checked", rng.random((5, 7)).astype(np.float32), "single", e, chunks=(3, 2), fletcher32=True)
    save(p, e, True)

    # 3. latest format: version-2 headers, link messages, layout 4 with a fixed-array index
    #    (deflate), a single-chunk index, an implicit index, a big-endian dataset
    p = os.path.join(OUT, "psf_latest.mat")
    e = {}
    with h5py.File(p, "w", userblock_size=512, libver="latest") as f:
        H, ca = psf_values(rng, 7, 3, 2, np.float32)
        put(f, "H", H, "single", e, chunks=(1, 3, 3, 7, 7), compression="gzip")
        put(f, "Nnum", np.array([[3.0]]), "double", e)
        put(f, "CAindex", ca, "double", e, chunks=(2, 2))
        put_implicit(f, "implicit", rng.integers(0, 1000, (6, 10)).astype(np.int32), (4, 3), e)
        put(f, "bigendian", rng.random((4, 5)).astype(">f4"), "single", e)
    save(p, e, True)

    # 4. plain HDF5 structures: 40 names in the root symbol table (several nodes), a chunk
    #    B-tree of more than one level (1000 chunks), nested groups, every integer width
    p = os.path.join(OUT, "structures.h5")
    e = {}
    with h5py.File(p, "w", libver="earliest") as f:
        for i in range(40):
            name = "v%02d" % i
            f.create_dataset(name, data=np.arange(i + 1, dtype=np.int16) * (i - 20))
            e[name] = f[name][()]
        big = rng.standard_normal((100, 50))
        f.create_dataset("many_chunks", data=big, chunks=(2, 5))
        e["many_chunks"] = big
        g = f.create_group("grp").create_group("sub")
        for dt in ("u1", "i1", "<u2", ">i2", "<u4", ">i4", "<u8", ">i8", "<f2", ">f8"):
            nm = "grp/sub/t_" + dt.replace("<", "le").replace(">", "be")
            a = (rng.integers(0, 100, (3, 4)) - (50 if "i" in dt or "f" in dt else 0)).astype(dt)
            f.create_dataset(nm, data=a)
            e[nm] = a
        f.create_dataset("scalar", data=np.float64(2.5))
        e["scalar"] = np.asarray(np.float64(2.5))
    save(p, e, False)

    # 5. an extensible-array chunk index (unlimited dimension, latest format): must be refused
    p = os.path.join(OUT, "extensible.h5")
    with h5py.File(p, "w", libver="latest") as f:
        f.create_dataset("grow", data=np.arange(12.0).reshape(3, 4), maxshape=(None, 4), chunks=(1, 4))

    # external files written by other programs
    ext = {}
    sources = sorted(glob.glob("/usr/local/lib/python3.10/dist-packages/scipy/io/matlab/tests/data/testhdf5*.mat") +
                     glob.glob("/opt/conda/lib/python3.9/site-packages/tables/tests/*.h5"))
    for path in sources:
        try:
            f = h5py.File(path, "r")
        except Exception:
            continue
        entries = {}

        def visit(name, obj):
            if not isinstance(obj, h5py.Dataset):
                return
            try:
                dt = obj.dtype
                if dt.fields is not None or dt.kind not in "fiu" or dt.itemsize > 8:
                    return
                a = np.ascontiguousarray(obj[()])
            except Exception:
                return
            hsh = hashlib.sha256(a.dtype.str.encode() + repr(a.shape).encode() + a.tobytes()).hexdigest()
            entries[name] = {"dtype": a.dtype.str, "shape": list(a.shape), "sha256": hsh}

        f.visititems(visit)
        f.close()
        if entries:
            ext[path] = entries
    with open(os.path.join(OUT, "external_checksums.json"), "w") as fh:
        json.dump(ext, fh, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
