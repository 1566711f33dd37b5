"""Read-only access to the HDF5 files that NetCDF-4 writes (no h5py needed).

Written from the public HDF5 File Format Specification (version 3.0).  It
covers what netCDF-C and xarray produce with HDF5 1.8-or-later format
bounds, which is what the reference's example files are:

* superblock version 2 or 3 (8-byte offsets and lengths);
* version-2 object headers ("OHDR") with their continuation blocks ("OCHK");
* groups whose links are stored compactly in Link messages;
* dataspace, datatype (fixed-point, floating-point, fixed-length string),
  data-layout version 3 (compact, contiguous, chunked with a version-1 B-tree
  index) and filter-pipeline messages; the deflate, shuffle and Fletcher-32
  filters;
* attributes stored in the object header (numeric and fixed-length strings;
  variable-length ones are reported as None).

Anything else (superblock 0/1, dense link or attribute storage, layout
version 4 chunk indexes, other filters) raises NotImplementedError naming
the feature, rather than returning wrong data.
"""
from __future__ import annotations

import mmap
import zlib
from pathlib import Path
from typing import Dict, Optional

import numpy as np

_SIGNATURE = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF

# message types (format spec IV.A.2)
_DATASPACE, _LINK_INFO, _DATATYPE = 0x01, 0x02, 0x03
_LINK, _LAYOUT, _FILTERS, _ATTRIBUTE, _CONTINUATION = 0x06, 0x08, 0x0B, 0x0C, 0x10
_SYMBOL_TABLE, _ATTRIBUTE_INFO = 0x11, 0x15


class HDF5Error(ValueError):
    pass


def _u(buf, off: int, n: int) -> int:
    return int.from_bytes(buf[off:off + n], "little")


class _Datatype:
    """Decoded datatype message: numpy dtype, or None for types this reader does not decode."""

    def __init__(self, buf, off: int):
        b0 = buf[off]
        cls, bits0, size = b0 & 0x0F, buf[off + 1], _u(buf, off + 4, 4)
        self.size = size
        self.dtype: Optional[np.dtype] = None
        self.kind = cls
        order = ">" if bits0 & 0x01 else "<"
        if cls == 0:      # fixed-point
            signed = bool(bits0 & 0x08)
            if size in (1, 2, 4, 8):
                self.dtype = np.dtype(f"{order}{'i' if signed else 'u'}{size}")
        elif cls == 1:    # floating-point (IEEE layouts only)
            if bits0 & 0x40:
                raise NotImplementedError("HDF5 VAX-order floating point")
            if size in (2, 4, 8):
                self.dtype = np.dtype(f"{order}f{size}")
        elif cls == 3:    # fixed-length string
            self.dtype = np.dtype(f"S{size}")


class Dataset:
    """One HDF5 dataset: shape, dtype, attributes; ``read()`` returns a numpy array."""

    def __init__(self, f: "File", name: str, addr: int):
        self.file, self.name, self.addr = f, name, addr
        self.shape = None
        self.datatype: Optional[_Datatype] = None
        self._layout = None
        self._filters = []
        self._attr_msgs = []
        self._dense_attrs = False
        for mtype, off, size in f._messages(addr):
            if mtype == _DATASPACE:
                self.shape = f._dataspace(off)
            elif mtype == _DATATYPE:
                self.datatype = _Datatype(f.buf, off)
            elif mtype == _LAYOUT:
                self._layout = (off, size)
            elif mtype == _FILTERS:
                self._filters = f._filter_pipeline(off)
            elif mtype == _ATTRIBUTE:
                self._attr_msgs.append(off)
            elif mtype == _ATTRIBUTE_INFO:
                self._dense_attrs = f._dense_attribute_storage(off)
        if self.shape is None or self.datatype is None or self._layout is None:
            raise HDF5Error(f"object {name!r} is not a dataset")
        self._attrs = None

    @property
    def dtype(self):
        return self.datatype.dtype

    @property
    def attrs(self) -> Dict[str, object]:
        if self._attrs is None:
            if self._dense_attrs:
                raise NotImplementedError(f"dense attribute storage (dataset {self.name!r})")
            self._attrs = dict(self.file._attribute(off) for off in self._attr_msgs)
        return self._attrs

    def read(self) -> np.ndarray:
        f, dt = self.file, self.datatype
        if dt.dtype is None:
            raise NotImplementedError(f"HDF5 datatype class {dt.kind} (dataset {self.name!r})")
        off, _ = self._layout
        version, cls = f.buf[off], f.buf[off + 1]
        if version != 3:
            raise NotImplementedError(f"data layout message version {version} (dataset {self.name!r})")
        n = int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1
        if cls == 0:       # compact: the raw data sits in the message
            size = _u(f.buf, off + 2, 2)
            raw = bytes(f.buf[off + 4:off + 4 + size])
            return np.frombuffer(raw, dtype=dt.dtype, count=n).reshape(self.shape)
        if cls == 1:       # contiguous
            addr, size = _u(f.buf, off + 2, 8), _u(f.buf, off + 10, 8)
            if addr == _UNDEF:
                return np.zeros(self.shape, dtype=dt.dtype)
            if size < n * dt.size:
                raise HDF5Error(f"contiguous storage of {self.name!r} is shorter than its shape")
            return np.frombuffer(f.buf, dtype=dt.dtype, count=n, offset=addr).reshape(self.shape).copy()
        if cls == 2:       # chunked, version-1 B-tree index
            rank = f.buf[off + 2] - 1
            btree = _u(f.buf, off + 3, 8)
            chunk = tuple(_u(f.buf, off + 11 + 4 * i, 4) for i in range(rank))
            out = np.zeros(self.shape, dtype=dt.dtype)
            if btree == _UNDEF:
                return out
            for caddr, csize, mask, origin in f._chunk_entries(btree, rank):
                data = f._unfilter(bytes(f.buf[caddr:caddr + csize]), self._filters, mask, dt.size)
                block = np.frombuffer(data, dtype=dt.dtype, count=int(np.prod(chunk))).reshape(chunk)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(origin, chunk, self.shape))
                out[sl] = block[tuple(slice(0, s.stop - s.start) for s in sl)]
            return out
        raise NotImplementedError(f"data layout class {cls} (dataset {self.name!r})")


class File:
    """An HDF5 file opened read-only (memory-mapped).

    ``links`` maps the root group's link names to object-header addresses,
    ``dataset(name)`` returns that object as a Dataset, ``root_attrs`` holds the
    root group's attributes."""

    def __init__(self, path):
        self.path = Path(path)
        with open(self.path, "rb") as fh:
            self._map = mmap.mmap(fh.fileno(), 0, access=mmap.ACCESS_READ)
        self.buf = memoryview(self._map)
        self._datasets = {}
        try:
            self._open()
        except BaseException:
            self.close()
            raise

    def _open(self):
        if bytes(self.buf[:8]) != _SIGNATURE:
            raise HDF5Error(f"{self.path} is not an HDF5 file")
        version = self.buf[8]
        if version not in (2, 3):
            raise NotImplementedError(f"HDF5 superblock version {version} ({self.path})")
        if self.buf[9] != 8 or self.buf[10] != 8:
            raise NotImplementedError("HDF5 offsets or lengths other than 8 bytes")
        base = _u(self.buf, 12, 8)
        if base != 0:
            raise NotImplementedError("HDF5 base address other than 0")
        self.root = _u(self.buf, 36, 8)
        self.links = self._group_links(self.root)
        self.root_attrs = dict(self._attribute(off) for t, off, _ in self._messages(self.root) if t == _ATTRIBUTE)

    def close(self):
        if self._map.closed:
            return
        self.buf.release()
        self._map.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def dataset(self, name: str) -> Dataset:
        if name not in self._datasets:
            if name not in self.links:
                raise KeyError(name)
            self._datasets[name] = Dataset(self, name, self.links[name])
        return self._datasets[name]

    # ----------------------------------------------------------------- object headers
    def _messages(self, addr: int):
        """(type, data offset, size) of every message of a version-2 object header."""
        buf = self.buf
        if bytes(buf[addr:addr + 4]) != b"OHDR":
            raise NotImplementedError(f"object header at {addr} is not version 2")
        flags = buf[addr + 5]
        p = addr + 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
        width = 1 << (flags & 0x03)
        size = _u(buf, p, width)
        p += width
        out, blocks = [], [(p, p + size)]
        order = 2 if flags & 0x04 else 0
        while blocks:
            p, end = blocks.pop(0)
            while p + 4 + order <= end:
                mtype, msize = buf[p], _u(buf, p + 1, 2)
                p += 4 + order
                if mtype == _CONTINUATION:
                    caddr, clen = _u(buf, p, 8), _u(buf, p + 8, 8)
                    if bytes(buf[caddr:caddr + 4]) != b"OCHK":
                        raise HDF5Error(f"bad continuation block at {caddr}")
                    blocks.append((caddr + 4, caddr + clen - 4))
                elif mtype != 0:
                    out.append((mtype, p, msize))
                p += msize
        return out

    def _group_links(self, addr: int) -> Dict[str, int]:
        links = {}
        for mtype, off, _ in self._messages(addr):
            if mtype == _LINK:
                name, target = self._link(off)
                if target is not None:
                    links[name] = target
            elif mtype == _LINK_INFO:
                flags = self.buf[off + 1]
                heap = _u(self.buf, off + 2 + (8 if flags & 1 else 0), 8)
                if heap != _UNDEF:
                    raise NotImplementedError("HDF5 dense link storage (fractal heap)")
            elif mtype == _SYMBOL_TABLE:
                raise NotImplementedError("HDF5 symbol-table groups")
        return links

    def _link(self, off: int):
        buf = self.buf
        flags = buf[off + 1]
        p = off + 2
        kind = 0
        if flags & 0x08:
            kind = buf[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        width = 1 << (flags & 0x03)
        n = _u(buf, p, width)
        p += width
        name = bytes(buf[p:p + n]).decode("utf-8")
        p += n
        return name, (_u(buf, p, 8) if kind == 0 else None)

    def _dataspace(self, off: int):
        buf = self.buf
        version, rank, flags = buf[off], buf[off + 1], buf[off + 2]
        if version == 1:
            p = off + 8
        elif version == 2:
            if buf[off + 3] == 2:
                return (0,)
            p = off + 4
        else:
            raise NotImplementedError(f"dataspace message version {version}")
        return tuple(_u(buf, p + 8 * i, 8) for i in range(rank))

    def _filter_pipeline(self, off: int):
        buf = self.buf
        version, count = buf[off], buf[off + 1]
        p = off + (8 if version == 1 else 2)
        filters = []
        for _ in range(count):
            fid = _u(buf, p, 2)
            p += 2
            name_len = 0
            if version == 1 or fid >= 256:
                name_len = _u(buf, p, 2)
                p += 2
            p += 2  # flags
            nvals = _u(buf, p, 2)
            p += 2
            if version == 1:
                name_len = (name_len + 7) // 8 * 8
            p += name_len
            vals = [_u(buf, p + 4 * i, 4) for i in range(nvals)]
            p += 4 * nvals
            if version == 1 and nvals % 2:
                p += 4
            filters.append((fid, vals))
        return filters

    def _dense_attribute_storage(self, off: int) -> bool:
        flags = self.buf[off + 1]
        heap = _u(self.buf, off + 2 + (2 if flags & 1 else 0), 8)
        return heap != _UNDEF

    def _attribute(self, off: int):
        buf = self.buf
        version = buf[off]
        name_size, type_size, space_size = _u(buf, off + 2, 2), _u(buf, off + 4, 2), _u(buf, off + 6, 2)
        if version == 1:
            def pad(n):
                return (n + 7) // 8 * 8
            p = off + 8
        elif version in (2, 3):
            def pad(n):
                return n
            p = off + 8 + (1 if version == 3 else 0)
        else:
            raise NotImplementedError(f"attribute message version {version}")
        name = bytes(buf[p:p + name_size]).split(b"\0", 1)[0].decode("utf-8")
        p += pad(name_size)
        dt = _Datatype(buf, p)
        p += pad(type_size)
        shape = self._dataspace(p)
        p += pad(space_size)
        if dt.dtype is None:
            return name, None
        n = int(np.prod(shape, dtype=np.int64)) if shape else 1
        value = np.frombuffer(bytes(buf[p:p + n * dt.size]), dtype=dt.dtype, count=n)
        if dt.kind == 3:
            text = [v.split(b"\0", 1)[0].decode("utf-8", "replace") for v in value]
            return name, text[0] if len(text) == 1 else text
        return name, value.reshape(shape) if shape else value[0]

    # ----------------------------------------------------------------- chunked storage
    def _chunk_entries(self, addr: int, rank: int):
        """Leaf entries (address, stored size, filter mask, chunk origin) of a raw-data B-tree."""
        buf = self.buf
        if bytes(buf[addr:addr + 4]) != b"TREE" or buf[addr + 4] != 1:
            raise HDF5Error(f"bad chunk B-tree node at {addr}")
        level, used = buf[addr + 5], _u(buf, addr + 6, 2)
        key = 8 + 8 * (rank + 1)
        p = addr + 24
        for _ in range(used):
            size, mask = _u(buf, p, 4), _u(buf, p + 4, 4)
            origin = tuple(_u(buf, p + 8 + 8 * i, 8) for i in range(rank))
            child = _u(buf, p + key, 8)
            if level == 0:
                yield child, size, mask, origin
            else:
                yield from self._chunk_entries(child, rank)
            p += key + 8

    @staticmethod
    def _unfilter(data: bytes, filters, mask: int, elem: int) -> bytes:
        for i in range(len(filters) - 1, -1, -1):
            if mask & (1 << i):
                continue
            fid, vals = filters[i]
            if fid == 1:       # deflate
                data = zlib.decompress(data)
            elif fid == 2:     # shuffle: bytes grouped by position within each element
                size = vals[0] if vals else elem
                n = len(data) // size
                body = np.frombuffer(data[:n * size], dtype=np.uint8).reshape(size, n).T.tobytes()
                data = body + data[n * size:]
            elif fid == 3:     # Fletcher-32: a 4-byte checksum appended
                data = data[:-4]
            else:
                raise NotImplementedError(f"HDF5 filter id {fid}")
        return data
