"""Minimal HDF5 reader (numpy + zlib), enough for MATLAB -v7.3 MAT-files.

The PSF of the reference lives in a MATLAB -v7.3 MAT-file (sid_main.m:30
``matfile(Input.psffile)``; written by external/save_psf_fast.m:39-74 through
``save -v7.3`` and ``h5create``/``h5write``).  Such a file is an HDF5 file
behind a 512-byte user block.  No HDF5 library is importable here, so this
module reads the on-disk format directly, following the published HDF5 File
Format Specification (version 3.0):

- superblock versions 0-3 (found at 0, 512, 1024, ... bytes; addresses are
  relative to its base address);
- object headers version 1 and 2 ("OHDR"/"OCHK"), with continuation blocks;
- groups stored as symbol tables (version-1 B-tree of "SNOD" nodes plus a
  local heap) or as link messages in the object header;
- dataspaces (scalar / simple), fixed-point, bitfield, enumeration (as its
  integer base type) and IEEE floating-point datatypes of either byte order,
  fixed-length strings (for attributes);
- compact, contiguous and chunked layouts; chunk indexes: version-1 B-tree
  (layout message 3), and for layout message 4 single chunk, implicit and
  fixed array (unpaged);
- filters deflate (1), shuffle (2) and fletcher32 (3).

Anything else (dense link storage, extensible-array or version-2 B-tree chunk
indexes, compound / variable-length data, other filters, external files) raises
H5Error naming the feature.  Contiguous datasets are returned as read-only
views of a memory map; everything else is assembled into a new array.
"""
from __future__ import annotations

import mmap
import struct
import zlib

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"


class H5Error(Exception):
    """Malformed file or a feature this reader does not implement."""


class Dataset:
    """Metadata of one dataset; ``read()`` returns its values in HDF5 (C) order."""

    def __init__(self, f: "File", name: str, msgs):
        self.file, self.name = f, name
        self.shape = None
        self.dtype = None
        self.layout = None
        self.filters = []
        self.attrs = {}
        self.fill = None
        for mtype, body in msgs:
            if mtype == 0x0001:
                self.shape = f._dataspace(body)
            elif mtype == 0x0003:
                self.dtype = f._datatype(body)[0]
            elif mtype == 0x0008:
                self.layout = body
            elif mtype == 0x000B:
                self.filters = f._filters(body)
            elif mtype == 0x000C:
                try:
                    k, v = f._attribute(body)
                    self.attrs[k] = v
                except H5Error:
                    pass  # attributes of unsupported types are not needed
        if self.shape is None or self.dtype is None or self.layout is None:
            raise H5Error(f"{name}: dataset without dataspace, datatype or layout message")

    def read(self) -> np.ndarray:
        return self.file._read_data(self)


class File:
    """An HDF5 file opened read-only through a memory map."""

    def __init__(self, path: str):
        self._fh = open(path, "rb")
        try:
            self._mm = mmap.mmap(self._fh.fileno(), 0, access=mmap.ACCESS_READ)
        except ValueError as e:  # empty file
            self._fh.close()
            raise H5Error(f"{path}: not an HDF5 file ({e})") from None
        self.path = path
        self._size = len(self._mm)
        self._find_superblock()

    # -- plumbing -----------------------------------------------------------
    def close(self):
        # arrays returned as views keep the map alive through their base
        self._fh.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _bytes(self, off: int, n: int) -> bytes:
        if off < 0 or n < 0 or off + n > self._size:
            raise H5Error(f"{self.path}: read of {n} bytes at {off} outside the file ({self._size} bytes)")
        return self._mm[off:off + n]

    def _uint(self, off: int, n: int) -> int:
        return int.from_bytes(self._bytes(off, n), "little")

    def _addr(self, off: int):
        """An address field (size of offsets) at absolute position off -> (absolute address or None, next)."""
        v = self._uint(off, self.so)
        if v == (1 << (8 * self.so)) - 1:
            return None, off + self.so
        return self.base + v, off + self.so

    def _len(self, off: int):
        return self._uint(off, self.sl), off + self.sl

    def _find_superblock(self):
        pos = 0
        while pos + 8 <= self._size:
            if self._mm[pos:pos + 8] == SIGNATURE:
                break
            pos = 512 if pos == 0 else pos * 2
        else:
            raise H5Error(f"{self.path}: no HDF5 signature (not an HDF5 / MATLAB -v7.3 file)")
        self.sb = pos
        ver = self._bytes(pos + 8, 1)[0]
        self.sb_version = ver
        if ver in (0, 1):
            self.so, self.sl = self._bytes(pos + 13, 1)[0], self._bytes(pos + 14, 1)[0]
            self.group_leaf_k = struct.unpack("<H", self._bytes(pos + 16, 2))[0]
            off = pos + 24 + (4 if ver == 1 else 0)
            self.base = 0
            base = self._uint(off, self.so)
            self.base = base
            off += 4 * self.so  # base, free-space info, end of file, driver info
            # root group symbol table entry: name offset, object header address, cache type, scratch
            off += self.so
            self.root, _ = self._addr(off)
        elif ver in (2, 3):
            self.so, self.sl = self._bytes(pos + 9, 1)[0], self._bytes(pos + 10, 1)[0]
            off = pos + 12
            self.base = 0
            self.base = self._uint(off, self.so)
            off += 3 * self.so  # base, superblock extension, end of file
            self.root, _ = self._addr(off)
        else:
            raise H5Error(f"{self.path}: superblock version {ver} is not supported")
        if self.so not in (2, 4, 8) or self.sl not in (2, 4, 8):
            raise H5Error(f"{self.path}: unusual offset / length sizes {self.so} / {self.sl}")

    # -- object headers -------------------------------------------------------
    def _messages(self, addr: int):
        """All header messages of the object at addr: list of (type, body bytes)."""
        out = []
        if self._bytes(addr, 4) == b"OHDR":
            ver, flags = self._bytes(addr + 4, 2)
            if ver != 2:
                raise H5Error(f"object header version {ver} at {addr}")
            off = addr + 6
            if flags & 0x20:
                off += 16
            if flags & 0x10:
                off += 4
            csize_len = 1 << (flags & 3)
            size0 = self._uint(off, csize_len)
            off += csize_len
            blocks = [(off, size0)]
            hdr = 4 + (2 if flags & 0x04 else 0)   # creation order field when tracked
            while blocks:
                start, size = blocks.pop(0)
                p, end = start, start + size
                while p + hdr <= end:
                    mtype = self._bytes(p, 1)[0]
                    msize = struct.unpack("<H", self._bytes(p + 1, 2))[0]
                    mflags = self._bytes(p + 3, 1)[0]
                    p += hdr
                    body = self._bytes(p, msize)
                    p += msize
                    if mtype == 0x10:
                        caddr = self.base + int.from_bytes(body[:self.so], "little")
                        clen = int.from_bytes(body[self.so:self.so + self.sl], "little")
                        if self._bytes(caddr, 4) != b"OCHK":
                            raise H5Error(f"continuation block without OCHK at {caddr}")
                        blocks.append((caddr + 4, clen - 8))
                    elif mtype != 0:
                        out.append(self._checked(mtype, mflags, body))
            return out
        ver = self._bytes(addr, 1)[0]
        if ver != 1:
            raise H5Error(f"object header version {ver} at {addr}")
        size0 = self._uint(addr + 8, 4)
        blocks = [(addr + 16, size0)]
        while blocks:
            start, size = blocks.pop(0)
            p, end = start, start + size
            while p + 8 <= end:
                mtype, msize = struct.unpack("<HH", self._bytes(p, 4))
                mflags = self._bytes(p + 4, 1)[0]
                body = self._bytes(p + 8, msize)
                p += 8 + msize
                if mtype == 0x10:
                    caddr = self.base + int.from_bytes(body[:self.so], "little")
                    clen = int.from_bytes(body[self.so:self.so + self.sl], "little")
                    blocks.append((caddr, clen))
                elif mtype != 0:
                    out.append(self._checked(mtype, mflags, body))
        return out

    @staticmethod
    def _checked(mtype: int, mflags: int, body: bytes):
        # a shared message holds a reference to the message stored elsewhere
        if mflags & 0x02 and mtype in (0x0001, 0x0003, 0x0005, 0x0008, 0x000B):
            raise H5Error(f"shared header message of type {mtype:#x} is not supported")
        return mtype, body

    # -- message decoders -----------------------------------------------------
    def _dataspace(self, b: bytes):
        ver, rank, flags = b[0], b[1], b[2]
        if ver == 1:
            off = 8
        elif ver == 2:
            off = 4
            if b[3] == 2:   # null dataspace
                return (0,)
        else:
            raise H5Error(f"dataspace version {ver}")
        dims = tuple(int.from_bytes(b[off + i * self.sl:off + (i + 1) * self.sl], "little") for i in range(rank))
        return dims

    def _datatype(self, b: bytes):
        """-> (numpy dtype, bytes consumed)."""
        cls, ver = b[0] & 0x0F, b[0] >> 4
        bits = b[1] | (b[2] << 8) | (b[3] << 16)
        size = struct.unpack("<I", b[4:8])[0]
        if cls == 0:   # fixed point
            order = ">" if bits & 1 else "<"
            kind = "i" if bits & 0x08 else "u"
            if size not in (1, 2, 4, 8):
                raise H5Error(f"integer of {size} bytes")
            return np.dtype(f"{order}{kind}{size}"), 12
        if cls == 1:   # IEEE floating point
            if bits & 0x40:
                raise H5Error("VAX-order floating point")
            order = ">" if bits & 1 else "<"
            if size not in (2, 4, 8):
                raise H5Error(f"floating point of {size} bytes")
            return np.dtype(f"{order}f{size}"), 20
        if cls == 3:   # fixed-length string
            return np.dtype(f"S{size}"), 8
        if cls == 4:   # bitfield: stored as an unsigned integer of its size
            if size not in (1, 2, 4, 8):
                raise H5Error(f"bitfield of {size} bytes")
            return np.dtype(f"{'>' if bits & 1 else '<'}u{size}"), 12
        if cls == 8:   # enumeration: the values are stored in its base integer type
            base, used = self._datatype(b[8:])
            return base, 8 + used
        raise H5Error(f"datatype class {cls} (version {ver}) is not supported")

    def _filters(self, b: bytes):
        ver, n = b[0], b[1]
        off = 8 if ver == 1 else 2
        out = []
        for _ in range(n):
            fid = struct.unpack("<H", b[off:off + 2])[0]
            if ver == 1 or fid >= 256:
                nlen = struct.unpack("<H", b[off + 2:off + 4])[0]
                off += 4
            else:
                nlen = 0
                off += 2
            flags, ncd = struct.unpack("<HH", b[off:off + 4])
            off += 4
            if ver == 1:
                off += (nlen + 7) // 8 * 8
            else:
                off += nlen
            cd = list(struct.unpack(f"<{ncd}I", b[off:off + 4 * ncd]))
            off += 4 * ncd
            if ver == 1 and ncd % 2:
                off += 4
            out.append((fid, flags, cd))
        return out

    def _attribute(self, b: bytes):
        ver = b[0]
        nsize, tsize, ssize = struct.unpack("<HHH", b[2:8])
        off = 8 + (1 if ver == 3 else 0)
        pad = (lambda x: (x + 7) // 8 * 8) if ver == 1 else (lambda x: x)
        if ver not in (1, 2, 3):
            raise H5Error(f"attribute version {ver}")
        if ver >= 2 and b[1] & 0x03:
            raise H5Error("shared attribute datatype / dataspace")
        name = b[off:off + nsize].split(b"\0", 1)[0].decode("utf-8", "replace")
        off += pad(nsize)
        dt = self._datatype(b[off:off + tsize])[0]
        off += pad(tsize)
        shape = self._dataspace(b[off:off + ssize])
        off += pad(ssize)
        n = int(np.prod(shape)) if shape else 1
        arr = np.frombuffer(b, dtype=dt, count=n, offset=off)
        if dt.kind == "S":
            v = arr[0].split(b"\0", 1)[0].decode("utf-8", "replace") if n == 1 else [
                x.split(b"\0", 1)[0].decode("utf-8", "replace") for x in arr]
        else:
            v = arr.reshape(shape).copy() if shape else arr[0]
        return name, v

    # -- groups ---------------------------------------------------------------
    def _group_links(self, addr: int) -> dict:
        """name -> object header address of every link in the group at addr."""
        links = {}
        for mtype, body in self._messages(addr):
            if mtype == 0x0011:   # symbol table: v1 B-tree + local heap
                btree = self.base + int.from_bytes(body[:self.so], "little")
                heap = self.base + int.from_bytes(body[self.so:2 * self.so], "little")
                if self._bytes(heap, 4) != b"HEAP":
                    raise H5Error(f"local heap signature missing at {heap}")
                hdata, _ = self._addr(heap + 8 + 2 * self.sl)
                self._walk_group_btree(btree, hdata, links)
            elif mtype == 0x0006:  # link message
                name, target = self._link(body)
                if target is not None:
                    links[name] = target
            elif mtype == 0x0002:  # link info: dense storage when the fractal heap address is defined
                flags = body[1]
                off = 2 + (8 if flags & 1 else 0)
                heap_addr = int.from_bytes(body[off:off + self.so], "little")
                if heap_addr != (1 << (8 * self.so)) - 1:
                    raise H5Error("groups in dense link storage (fractal heap) are not supported")
        return links

    def _walk_group_btree(self, node: int, hdata: int, links: dict):
        if self._bytes(node, 4) != b"TREE":
            raise H5Error(f"B-tree signature missing at {node}")
        ntype, level = self._bytes(node + 4, 2)
        nused = self._uint(node + 6, 2)
        if ntype != 0:
            raise H5Error("group B-tree of the wrong node type")
        off = node + 8 + 2 * self.so + self.sl   # first key
        for _ in range(nused):
            child, off = self._addr(off)
            off += self.sl   # next key
            if level > 0:
                self._walk_group_btree(child, hdata, links)
            else:
                if self._bytes(child, 4) != b"SNOD":
                    raise H5Error(f"symbol table node signature missing at {child}")
                nsym = self._uint(child + 6, 2)
                e = child + 8
                for _ in range(nsym):
                    name_off = self._uint(e, self.so)
                    obj, _ = self._addr(e + self.so)
                    if obj is not None:   # soft links (cache type 2) carry no object address
                        links[self._cstring(hdata + name_off)] = obj
                    e += 2 * self.so + 24

    def _cstring(self, off: int) -> str:
        end = self._mm.find(b"\0", off)
        if end < 0:
            raise H5Error("unterminated name in a local heap")
        return self._mm[off:end].decode("utf-8", "replace")

    def _link(self, b: bytes):
        flags = b[1]
        off = 2
        ltype = 0
        if flags & 0x08:
            ltype = b[off]
            off += 1
        if flags & 0x04:
            off += 8
        if flags & 0x10:
            off += 1
        nl = 1 << (flags & 3)
        nlen = int.from_bytes(b[off:off + nl], "little")
        off += nl
        name = b[off:off + nlen].decode("utf-8", "replace")
        off += nlen
        if ltype != 0:
            return name, None   # soft / external links are not followed
        return name, self.base + int.from_bytes(b[off:off + self.so], "little")

    def _resolve(self, path: str) -> int:
        addr = self.root
        for part in [p for p in path.split("/") if p]:
            links = self._group_links(addr)
            if part not in links:
                raise KeyError(path)
            addr = links[part]
        return addr

    # -- public ---------------------------------------------------------------
    def keys(self, group: str = "/"):
        return sorted(self._group_links(self._resolve(group)))

    def is_dataset(self, path: str) -> bool:
        return any(t == 0x0008 for t, _ in self._messages(self._resolve(path)))

    def dataset(self, path: str) -> Dataset:
        msgs = self._messages(self._resolve(path))
        return Dataset(self, path, msgs)

    def read(self, path: str) -> np.ndarray:
        return self.dataset(path).read()

    def visit(self, group: str = "/"):
        """Paths of every dataset below group (depth first, sorted by name)."""
        out = []
        for name in self.keys(group):
            p = group.rstrip("/") + "/" + name
            if self.is_dataset(p):
                out.append(p.lstrip("/"))
            else:
                out.extend(self.visit(p))
        return out

    # -- data -----------------------------------------------------------------
    def _read_data(self, ds: Dataset) -> np.ndarray:
        b = ds.layout
        ver = b[0]
        shape, dt = ds.shape, ds.dtype
        n = int(np.prod(shape)) if shape else 1
        if ver in (1, 2):
            return self._read_layout_v12(ds, b)
        if ver not in (3, 4):
            raise H5Error(f"data layout message version {ver}")
        lcls = b[1]
        if lcls == 0:   # compact
            size = struct.unpack("<H", b[2:4])[0]
            if size < n * dt.itemsize:
                raise H5Error(f"{ds.name}: compact data shorter than the dataspace")
            return np.frombuffer(b, dtype=dt, count=n, offset=4).reshape(shape).copy()
        if lcls == 1:   # contiguous
            addr = int.from_bytes(b[2:2 + self.so], "little")
            if addr == (1 << (8 * self.so)) - 1:   # never written: fill value
                return np.zeros(shape, dtype=dt)
            addr += self.base
            if ds.filters:
                raise H5Error(f"{ds.name}: filters on a contiguous dataset")
            self._bytes(addr, n * dt.itemsize)  # bounds check
            return np.frombuffer(self._mm, dtype=dt, count=n, offset=addr).reshape(shape)
        if lcls == 2:
            return self._read_chunked(ds, b)
        raise H5Error(f"{ds.name}: layout class {lcls} (virtual) is not supported")

    def _read_layout_v12(self, ds: Dataset, b: bytes) -> np.ndarray:
        """Layout message versions 1 and 2: dimensionality, class, 5 reserved
        bytes, an address (contiguous / chunked), 32-bit sizes (for chunks the
        last is the element size), and for compact data a size and the data."""
        shape, dt = ds.shape, ds.dtype
        n = int(np.prod(shape)) if shape else 1
        nd, lcls = b[1], b[2]
        off = 8
        addr = None
        if lcls in (1, 2):
            addr = int.from_bytes(b[off:off + self.so], "little")
            off += self.so
        dims = struct.unpack(f"<{nd}I", b[off:off + 4 * nd])
        off += 4 * nd
        if lcls == 0:
            size = struct.unpack("<I", b[off:off + 4])[0]
            if size < n * dt.itemsize:
                raise H5Error(f"{ds.name}: compact data shorter than the dataspace")
            return np.frombuffer(b, dtype=dt, count=n, offset=off + 4).reshape(shape).copy()
        if addr == (1 << (8 * self.so)) - 1:
            return np.zeros(shape, dtype=dt)
        if lcls == 1:
            if ds.filters:
                raise H5Error(f"{ds.name}: filters on a contiguous dataset")
            addr += self.base
            self._bytes(addr, n * dt.itemsize)
            return np.frombuffer(self._mm, dtype=dt, count=n, offset=addr).reshape(shape)
        rank = len(shape)
        cdims = dims[:rank]
        out = np.zeros(shape, dtype=dt)
        for coff, size, mask, caddr in self._chunk_btree(self.base + addr, rank):
            self._place(out, self._decode_chunk(ds, caddr, size, mask, cdims), coff, cdims)
        return out

    def _read_chunked(self, ds: Dataset, b: bytes) -> np.ndarray:
        shape, dt = ds.shape, ds.dtype
        rank = len(shape)
        out = np.zeros(shape, dtype=dt)
        if ds.layout[0] == 3:
            nd = b[2]
            addr = int.from_bytes(b[3:3 + self.so], "little")
            off = 3 + self.so
            cdims = struct.unpack(f"<{nd}I", b[off:off + 4 * nd])[:rank]
            if addr == (1 << (8 * self.so)) - 1:
                return out
            for coff, size, mask, caddr in self._chunk_btree(self.base + addr, rank):
                self._place(out, self._decode_chunk(ds, caddr, size, mask, cdims), coff, cdims)
            return out
        # layout message version 4
        flags, nd, enc = b[2], b[3], b[4]
        off = 5
        cdims = tuple(int.from_bytes(b[off + i * enc:off + (i + 1) * enc], "little") for i in range(nd))[:rank]
        off += nd * enc
        itype = b[off]
        off += 1
        grid = [-(-shape[i] // cdims[i]) for i in range(rank)]
        nchunks = int(np.prod(grid)) if grid else 1
        full = int(np.prod(cdims)) * dt.itemsize
        if itype == 1:   # single chunk
            if flags & 0x02:
                size = int.from_bytes(b[off:off + self.sl], "little")
                mask = struct.unpack("<I", b[off + self.sl:off + self.sl + 4])[0]
                off += self.sl + 4
            else:
                size, mask = full, 0
            caddr, _ = self._addr_in(b, off)
            if caddr is not None:
                self._place(out, self._decode_chunk(ds, caddr, size, mask, cdims), (0,) * rank, cdims)
            return out
        if itype == 2:   # implicit: chunks stored in order, unfiltered
            caddr, _ = self._addr_in(b, off)
            if caddr is None:
                return out
            for k in range(nchunks):
                coff = self._chunk_offset(k, grid, cdims)
                self._place(out, self._decode_chunk(ds, caddr + k * full, full, 0, cdims), coff, cdims)
            return out
        if itype == 3:   # fixed array
            off += 1     # page bits
            haddr, _ = self._addr_in(b, off)
            if haddr is None:
                return out
            for k, (caddr, size, mask) in enumerate(self._fixed_array(haddr, full)):
                if caddr is None:
                    continue
                coff = self._chunk_offset(k, grid, cdims)
                self._place(out, self._decode_chunk(ds, caddr, size, mask, cdims), coff, cdims)
            return out
        raise H5Error(f"{ds.name}: chunk index type {itype} (extensible array / version-2 B-tree) is not supported")

    def _addr_in(self, b: bytes, off: int):
        v = int.from_bytes(b[off:off + self.so], "little")
        if v == (1 << (8 * self.so)) - 1:
            return None, off + self.so
        return self.base + v, off + self.so

    @staticmethod
    def _chunk_offset(k: int, grid, cdims):
        idx = []
        for g in reversed(grid):
            idx.append(k % g)
            k //= g
        return tuple(i * c for i, c in zip(reversed(idx), cdims))

    def _fixed_array(self, haddr: int, full: int):
        if self._bytes(haddr, 4) != b"FAHD":
            raise H5Error(f"fixed array header signature missing at {haddr}")
        client, esize, pbits = self._bytes(haddr + 5, 3)
        nent, off = self._len(haddr + 8)
        daddr, _ = self._addr(off)
        if nent > (1 << pbits):
            raise H5Error("paged fixed-array chunk index is not supported")
        if self._bytes(daddr, 4) != b"FADB":
            raise H5Error(f"fixed array data block signature missing at {daddr}")
        e = daddr + 6 + self.so
        out = []
        for _ in range(nent):
            caddr, p = self._addr(e)
            if client == 1:   # filtered chunks: address, size, filter mask
                slen = esize - self.so - 4
                size = self._uint(p, slen)
                mask = self._uint(p + slen, 4)
            else:
                size, mask = full, 0
            out.append((caddr, size, mask))
            e += esize
        return out

    def _chunk_btree(self, node: int, rank: int):
        """(offsets, stored size, filter mask, address) of every chunk below a v1 B-tree node."""
        if self._bytes(node, 4) != b"TREE":
            raise H5Error(f"chunk B-tree signature missing at {node}")
        ntype, level = self._bytes(node + 4, 2)
        if ntype != 1:
            raise H5Error("chunk B-tree of the wrong node type")
        nused = self._uint(node + 6, 2)
        off = node + 8 + 2 * self.so
        ksize = 8 + 8 * (rank + 1)
        out = []
        for _ in range(nused):
            size, mask = struct.unpack("<II", self._bytes(off, 8))
            coff = struct.unpack(f"<{rank}Q", self._bytes(off + 8, 8 * rank))
            child, _ = self._addr(off + ksize)
            if level > 0:
                out.extend(self._chunk_btree(child, rank))
            else:
                out.append((coff, size, mask, child))
            off += ksize + self.so
        return out

    def _decode_chunk(self, ds: Dataset, addr: int, size: int, mask: int, cdims) -> np.ndarray:
        raw = self._bytes(addr, size)
        for i in reversed(range(len(ds.filters))):
            if mask & (1 << i):
                continue
            fid, _, cd = ds.filters[i]
            if fid == 1:
                try:
                    raw = zlib.decompress(raw)
                except zlib.error as e:
                    raise H5Error(f"{ds.name}: corrupt deflate chunk at {addr} ({e})") from None
            elif fid == 2:
                es = cd[0] if cd else ds.dtype.itemsize
                nel = len(raw) // es
                body = np.frombuffer(raw, dtype=np.uint8, count=nel * es).reshape(es, nel).T.tobytes()
                raw = body + raw[nel * es:]
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise H5Error(f"{ds.name}: filter {fid} is not supported (deflate, shuffle, fletcher32 are)")
        need = int(np.prod(cdims)) * ds.dtype.itemsize
        if len(raw) < need:
            raise H5Error(f"{ds.name}: chunk at {addr} holds {len(raw)} bytes, {need} expected")
        return np.frombuffer(raw, dtype=ds.dtype, count=int(np.prod(cdims))).reshape(cdims)

    @staticmethod
    def _place(out: np.ndarray, chunk: np.ndarray, coff, cdims):
        sl_out, sl_in = [], []
        for o, c, d in zip(coff, cdims, out.shape):
            if o >= d:
                return
            e = min(o + c, d)
            sl_out.append(slice(o, e))
            sl_in.append(slice(0, e - o))
        out[tuple(sl_out)] = chunk[tuple(sl_in)]
