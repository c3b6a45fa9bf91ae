"""A minimal read-only HDF5 reader for the reference's gold restart files.

TEST INFRASTRUCTURE ONLY (fixture generation).  The image has no h5py; the
reference's gold states (apps/perfect/manual.gold/*/*.h5) are HDF5 1.8 files
with a version 0 superblock, symbol-table groups, version 1 object headers,
contiguous or chunked (optionally deflated) datasets of little-endian
numbers or compounds of them, and attributes.  This reader implements that
subset of the published HDF5 file format specification and refuses anything
else loudly.
"""
import struct
import zlib

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"


class Unsupported(Exception):
    pass


def _elements(buf, dt, count, offset):
    """count elements of dt from buf at offset; an array datatype's
    elements come back as trailing dimensions of its base type."""
    return np.frombuffer(buf, dtype=dt.base, count=count * max(1, int(np.prod(dt.shape))),
                         offset=offset).reshape((count,) + dt.shape)


class File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.d = fh.read()
        if self.d[:8] != SIGNATURE:
            raise Unsupported("not an HDF5 file")
        if self.d[8] != 0:
            raise Unsupported(f"superblock version {self.d[8]}")
        if self.d[13] != 8 or self.d[14] != 8:
            raise Unsupported("offsets and lengths must be 8 bytes")
        self.base = self.u64(24)
        # root group symbol table entry at byte 56: name offset, object
        # header address, cache type, reserved, scratch (B-tree, heap)
        self.root = self.u64(56 + 8)

    # ------------------------------------------------------------ primitives
    def u8(self, o):
        return self.d[o]

    def u16(self, o):
        return struct.unpack_from("<H", self.d, o)[0]

    def u32(self, o):
        return struct.unpack_from("<I", self.d, o)[0]

    def u64(self, o):
        return struct.unpack_from("<Q", self.d, o)[0]

    def cstr(self, o):
        e = self.d.index(b"\0", o)
        return self.d[o:e].decode()

    # --------------------------------------------------------- object headers
    def messages(self, addr):
        """(type, flags, offset, size) of every message of the version 1
        object header at addr, following continuation blocks."""
        if self.u8(addr) != 1:
            raise Unsupported(f"object header version {self.u8(addr)} at {addr}")
        nmsg = self.u16(addr + 2)
        blocks = [(addr + 16, self.u32(addr + 8))]
        out = []
        while blocks and len(out) < nmsg:
            o, size = blocks.pop(0)
            end = o + size
            while o + 8 <= end and len(out) < nmsg:
                mtype, msize, flags = self.u16(o), self.u16(o + 2), self.u8(o + 4)
                body = o + 8
                if mtype == 0x10:                  # continuation
                    blocks.append((self.u64(body), self.u64(body + 8)))
                out.append((mtype, flags, body, msize))
                o = body + msize
        return out

    # ------------------------------------------------------------------ groups
    def _heap_data(self, heap):
        if self.d[heap:heap + 4] != b"HEAP":
            raise Unsupported("local heap signature")
        return self.u64(heap + 24)

    def _btree_group(self, node, heapdata, out):
        if self.d[node:node + 4] != b"TREE" or self.u8(node + 4) != 0:
            raise Unsupported("group B-tree node")
        level, used = self.u8(node + 5), self.u16(node + 6)
        o = node + 24 + 8                           # first key then child
        for _ in range(used):
            child = self.u64(o)
            if level > 0:
                self._btree_group(child, heapdata, out)
            else:
                self._snod(child, heapdata, out)
            o += 16
        return out

    def _snod(self, node, heapdata, out):
        if self.d[node:node + 4] != b"SNOD":
            raise Unsupported("symbol table node")
        n = self.u16(node + 6)
        o = node + 8
        for _ in range(n):
            name = self.cstr(heapdata + self.u64(o))
            out[name] = self.u64(o + 8)
            o += 40

    def members(self, addr):
        """name -> object header address of the group at addr."""
        for mtype, _, body, _ in self.messages(addr):
            if mtype == 0x11:
                btree, heap = self.u64(body), self.u64(body + 8)
                return self._btree_group(btree, self._heap_data(heap), {})
        return None

    def walk(self, addr=None, prefix=""):
        """path -> object header address of every dataset below addr."""
        addr = self.root if addr is None else addr
        out = {}
        for name, child in sorted(self.members(addr).items()):
            path = prefix + "/" + name
            if self.members(child) is not None:
                out.update(self.walk(child, path))
            else:
                out[path] = child
        return out

    # --------------------------------------------------------------- datatypes
    def dtype(self, o):
        """numpy dtype of the datatype message at o, and its byte length."""
        cv = self.u8(o)
        cls, ver = cv & 0x0F, cv >> 4
        bits = self.d[o + 1] | (self.d[o + 2] << 8) | (self.d[o + 3] << 16)
        size = self.u32(o + 4)
        p = o + 8
        if cls in (0, 1) and bits & 1:
            raise Unsupported("big-endian numbers")
        if cls == 0:
            kind = "i" if bits & 0x08 else "u"
            return np.dtype(f"<{kind}{size}"), 12
        if cls == 1:
            return np.dtype(f"<f{size}"), 20
        if cls == 3:
            return np.dtype(f"S{size}"), 8
        if cls == 6:
            names, formats, offsets = [], [], []
            nmemb = bits & 0xFFFF
            for _ in range(nmemb):
                name = self.cstr(p)
                if ver < 3:
                    p += (len(name) + 8) // 8 * 8
                else:
                    p += len(name) + 1
                if ver < 3:
                    off = self.u32(p)
                    p += 4
                else:
                    nb = max(1, (size.bit_length() + 7) // 8)
                    off = int.from_bytes(self.d[p:p + nb], "little")
                    p += nb
                if ver == 1:
                    ndim = self.u8(p)
                    if ndim:
                        raise Unsupported("array members")
                    p += 28
                mt, mlen = self.dtype(p)
                p += mlen
                names.append(name)
                formats.append(mt)
                offsets.append(off)
            dt = np.dtype({"names": names, "formats": formats,
                           "offsets": offsets, "itemsize": size})
            return dt, p - o
        if cls == 10:                                  # array of a base type
            ndim = self.u8(p)
            q = p + (4 if ver == 2 else 1)
            dims = [self.u32(q + 4 * i) for i in range(ndim)]
            q += 4 * ndim
            if ver == 2:
                q += 4 * ndim                          # permutation indices
            base, blen = self.dtype(q)
            return np.dtype((base, tuple(dims))), q + blen - o
        raise Unsupported(f"datatype class {cls}")

    def dataspace(self, o):
        ver, ndim, flags = self.u8(o), self.u8(o + 1), self.u8(o + 2)
        p = o + (8 if ver == 1 else 4)
        if ver == 2 and self.u8(o + 3) == 0:
            return ()                                  # scalar
        return tuple(self.u64(p + 8 * i) for i in range(ndim))

    # ---------------------------------------------------------------- datasets
    def _chunks(self, node, ndim, out):
        if self.d[node:node + 4] != b"TREE" or self.u8(node + 4) != 1:
            raise Unsupported("chunk B-tree node")
        level, used = self.u8(node + 5), self.u16(node + 6)
        ksize = 8 + 8 * ndim
        o = node + 24
        for _ in range(used):
            csize, mask = self.u32(o), self.u32(o + 4)
            offs = tuple(self.u64(o + 8 + 8 * i) for i in range(ndim))
            child = self.u64(o + ksize)
            if level > 0:
                self._chunks(child, ndim, out)
            else:
                out.append((offs, child, csize, mask))
            o += ksize + 8
        return out

    def _filters(self, o):
        ver, n = self.u8(o), self.u8(o + 1)
        p = o + (8 if ver == 1 else 2)
        ids = []
        for _ in range(n):
            fid = self.u16(p)
            if ver == 1 or fid >= 256:
                nlen = self.u16(p + 2)
                p += 4
            else:
                nlen = 0
                p += 2
            nval = self.u16(p + 2)
            p += 4
            p += ((nlen + 7) // 8 * 8) if ver == 1 else nlen
            p += 4 * nval
            if ver == 1 and nval % 2:
                p += 4
            ids.append(fid)
        return ids

    def read(self, addr):
        """The dataset at addr as a numpy array, and its attributes."""
        shape = dt = layout = None
        filters = []
        attrs = {}
        for mtype, _, body, _ in self.messages(addr):
            if mtype == 0x01:
                shape = self.dataspace(body)
            elif mtype == 0x03:
                dt, _ = self.dtype(body)
            elif mtype == 0x08:
                layout = body
            elif mtype == 0x0B:
                filters = self._filters(body)
            elif mtype == 0x0C:
                k, v = self._attribute(body)
                attrs[k] = v
        if shape is None or dt is None or layout is None:
            return None, attrs
        count = int(np.prod(shape)) if shape else 1
        ver, cls = self.u8(layout), self.u8(layout + 1)
        if ver != 3:
            raise Unsupported(f"layout version {ver}")
        if cls == 0:
            size = self.u16(layout + 2)
            raw = self.d[layout + 4:layout + 4 + size]
            arr = _elements(raw, dt, count, 0)
        elif cls == 1:
            a = self.u64(layout + 2)
            if a == 0xFFFFFFFFFFFFFFFF:
                arr = np.zeros((count,) + dt.shape, dtype=dt.base)
            else:
                arr = _elements(self.d, dt, count, a)
        elif cls == 2:
            ndim = self.u8(layout + 2)
            btree = self.u64(layout + 3)
            cdims = [self.u32(layout + 11 + 4 * i) for i in range(ndim)]
            arr = np.zeros(tuple(shape) + dt.shape, dtype=dt.base)
            if set(filters) - {1}:
                raise Unsupported(f"filters {filters}")
            if btree != 0xFFFFFFFFFFFFFFFF:
                for offs, child, csize, mask in self._chunks(btree, ndim, []):
                    raw = self.d[child:child + csize]
                    if 1 in filters and not mask & 1:
                        raw = zlib.decompress(raw)
                    chunk = _elements(raw, dt, int(np.prod(cdims[:-1])), 0).reshape(
                        tuple(cdims[:-1]) + dt.shape)
                    sl = tuple(slice(o, min(o + c, s)) for o, c, s in
                               zip(offs[:-1], cdims[:-1], shape))
                    arr[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]
            return arr, attrs
        else:
            raise Unsupported(f"layout class {cls}")
        return arr.reshape(tuple(shape) + dt.shape).copy(), attrs

    def _attribute(self, o):
        ver = self.u8(o)
        nlen, tlen, slen = self.u16(o + 2), self.u16(o + 4), self.u16(o + 6)
        if ver == 1:
            p = o + 8
            name = self.d[p:p + nlen].rstrip(b"\0").decode()
            p += (nlen + 7) // 8 * 8
            dt, _ = self.dtype(p)
            p += (tlen + 7) // 8 * 8
            shape = self.dataspace(p)
            p += (slen + 7) // 8 * 8
        elif ver in (2, 3):
            p = o + (8 if ver == 2 else 9)
            name = self.d[p:p + nlen].rstrip(b"\0").decode()
            p += nlen
            dt, _ = self.dtype(p)
            p += tlen
            shape = self.dataspace(p)
            p += slen
        else:
            raise Unsupported(f"attribute version {ver}")
        count = int(np.prod(shape)) if shape else 1
        val = _elements(self.d, dt, count, p).reshape(tuple(shape) + dt.shape).copy()
        return name, val

    def attrs(self, addr):
        out = {}
        for mtype, _, body, _ in self.messages(addr):
            if mtype == 0x0C:
                k, v = self._attribute(body)
                out[k] = v
        return out
