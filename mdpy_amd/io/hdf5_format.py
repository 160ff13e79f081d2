"""A small reader and writer of the HDF5 container, for the files mdpy's HDF5 writer and parser
exchange (mdpy/io/__init__.py:21-72: groups of numeric and string datasets), written from the
published HDF5 file-format specification because h5py is not available.

Covered: superblock version 0 with 8-byte offsets and lengths; groups as symbol tables (a
version-1 group B-tree, symbol nodes and a local heap); version-1 object headers with
continuation blocks; datasets with a version-1 or -2 dataspace (scalars and simple arrays),
fixed-point, floating-point, fixed-length string and variable-length string datatypes (the
strings in global heap collections), stored compact or contiguous. Chunked or filtered storage,
attributes and links other than the symbol table are not read (FileFormatError).

The writer lays files out the same way (every group one symbol node, datasets contiguous,
variable-length UTF-8 strings in one global heap collection).
"""
from __future__ import annotations

import struct

import numpy as np

from ..error import FileFormatError

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Reader:
    """Read-only view of an HDF5 file: `keys(path)` of a group, `read(path)` of a dataset."""

    def __init__(self, path):
        with open(path, "rb") as f:
            self.b = f.read()
        self.path = path
        if self.b[:8] != SIGNATURE:
            raise FileFormatError("%s is not an HDF5 file" % path)
        if self.b[8] != 0:
            raise FileFormatError("%s: HDF5 superblock version %d is not supported" % (path, self.b[8]))
        if self.b[13] != 8 or self.b[14] != 8:
            raise FileFormatError("%s: HDF5 offsets/lengths other than 8 bytes" % path)
        self.base = self.u64(24)
        root_header = self.u64(56 + 8)
        self.root = root_header

    # ------------------------------------------------------------------ primitives
    def u8(self, o):
        return self.b[o]

    def u16(self, o):
        return struct.unpack_from("<H", self.b, o)[0]

    def u32(self, o):
        return struct.unpack_from("<I", self.b, o)[0]

    def u64(self, o):
        return struct.unpack_from("<Q", self.b, o)[0]

    def addr(self, a):
        return self.base + a

    # ------------------------------------------------------------------ object headers
    def messages(self, header):
        """[(type, data offset, size)] of a version-1 object header and its continuations."""
        o = self.addr(header)
        if self.u8(o) != 1:
            raise FileFormatError("%s: object header version %d is not supported" % (self.path, self.u8(o)))
        nmsg = self.u16(o + 2)
        blocks = [(o + 16, self.u32(o + 8))]
        out = []
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            p = start
            while p + 8 <= start + size and len(out) < nmsg:
                mtype, msize = self.u16(p), self.u16(p + 2)
                if mtype == 0x10:  # continuation
                    blocks.append((self.addr(self.u64(p + 8)), self.u64(p + 16)))
                out.append((mtype, p + 8, msize))
                p += 8 + msize
        return out

    # ------------------------------------------------------------------ groups
    def _heap_name(self, heap, off):
        o = self.addr(heap)
        if self.b[o:o + 4] != b"HEAP":
            raise FileFormatError("%s: bad local heap" % self.path)
        data = self.addr(self.u64(o + 24))
        end = self.b.index(b"\0", data + off)
        return self.b[data + off:end].decode()

    def _tree_entries(self, node, heap):
        o = self.addr(node)
        if self.b[o:o + 4] != b"TREE" or self.u8(o + 4) != 0:
            raise FileFormatError("%s: bad group B-tree node" % self.path)
        level, used = self.u8(o + 5), self.u16(o + 6)
        p = o + 24 + 8  # past the header and key 0
        out = []
        for _ in range(used):
            child = self.u64(p)
            p += 16
            if level > 0:
                out += self._tree_entries(child, heap)
            else:
                s = self.addr(child)
                if self.b[s:s + 4] != b"SNOD":
                    raise FileFormatError("%s: bad symbol node" % self.path)
                for k in range(self.u16(s + 6)):
                    e = s + 8 + 40 * k
                    out.append((self._heap_name(heap, self.u64(e)), self.u64(e + 8)))
        return out

    def _children(self, header):
        for mtype, p, _ in self.messages(header):
            if mtype == 0x11:
                return self._tree_entries(self.u64(p), self.u64(p + 8))
        raise FileFormatError("%s: object is not a group" % self.path)

    def _find(self, path):
        h = self.root
        for part in [q for q in path.split("/") if q]:
            names = dict(self._children(h))
            if part not in names:
                raise KeyError(path)
            h = names[part]
        return h

    def keys(self, path="/"):
        return [n for n, _ in self._children(self._find(path))]

    def __contains__(self, path):
        try:
            self._find(path)
            return True
        except KeyError:
            return False

    # ------------------------------------------------------------------ datasets
    def _datatype(self, p):
        cls, ver = self.u8(p) & 15, self.u8(p) >> 4
        bits = self.u8(p + 1) | (self.u8(p + 2) << 8) | (self.u8(p + 3) << 16)
        size = self.u32(p + 4)
        order = ">" if bits & 1 else "<"
        if cls == 0:
            return ("num", np.dtype("%s%s%d" % (order, "i" if bits & 8 else "u", size)))
        if cls == 1:
            return ("num", np.dtype("%sf%d" % (order, size)))
        if cls == 3:
            return ("fstr", size)
        if cls == 9 and (bits & 15) == 1:
            return ("vstr", size)
        raise FileFormatError("%s: HDF5 datatype class %d (version %d) is not supported" % (self.path, cls, ver))

    def _dataspace(self, p):
        ver, rank, flags = self.u8(p), self.u8(p + 1), self.u8(p + 2)
        if ver == 1:
            q = p + 8
        elif ver == 2:
            if self.u8(p + 3) == 2:
                return None  # null dataspace
            q = p + 4
        else:
            raise FileFormatError("%s: dataspace version %d is not supported" % (self.path, ver))
        return tuple(self.u64(q + 8 * k) for k in range(rank))

    def _gheap_object(self, coll, index):
        o = self.addr(coll)
        if self.b[o:o + 4] != b"GCOL":
            raise FileFormatError("%s: bad global heap collection" % self.path)
        end = o + self.u64(o + 8)
        p = o + 16
        while p + 16 <= end:
            idx, size = self.u16(p), self.u64(p + 8)
            if idx == index:
                return self.b[p + 16:p + 16 + size]
            if idx == 0:
                break
            p += 16 + ((size + 7) & ~7)
        raise FileFormatError("%s: global heap object %d not found" % (self.path, index))

    def read(self, path):
        """A dataset's value: a numpy array (a numpy scalar for a scalar dataspace); strings come
        back as bytes objects (a numpy object array for arrays)."""
        shape = dtype = raw = None
        for mtype, p, size in self.messages(self._find(path)):
            if mtype == 0x01:
                shape = self._dataspace(p)
            elif mtype == 0x03:
                dtype = self._datatype(p)
            elif mtype == 0x08:
                ver = self.u8(p)
                if ver != 3:
                    raise FileFormatError("%s: data layout version %d is not supported" % (self.path, ver))
                cls = self.u8(p + 1)
                if cls == 0:
                    n = self.u16(p + 2)
                    raw = self.b[p + 4:p + 4 + n]
                elif cls == 1:
                    a, n = self.u64(p + 2), self.u64(p + 10)
                    raw = b"" if a == UNDEF else self.b[self.addr(a):self.addr(a) + n]
                else:
                    raise FileFormatError("%s: chunked HDF5 storage is not supported (%s)" % (self.path, path))
            elif mtype == 0x0C:
                raise FileFormatError("%s: HDF5 attributes are not supported (%s)" % (self.path, path))
        if shape is None or dtype is None or raw is None:
            raise FileFormatError("%s: %s is not a dataset" % (self.path, path))
        count = int(np.prod(shape)) if shape else 1
        kind, t = dtype
        if kind == "num":
            v = np.frombuffer(raw, dtype=t, count=count).astype(t.newbyteorder("="))
        elif kind == "fstr":
            v = np.array([raw[k * t:(k + 1) * t].rstrip(b"\0") for k in range(count)], dtype=object)
        else:
            v = np.empty(count, dtype=object)
            for k in range(count):
                q = 16 * k
                n, coll, idx = struct.unpack_from("<IQI", raw, q)
                v[k] = self._gheap_object(coll, idx)[:n] if n else b""
        return v.reshape(shape) if shape else v[0]


class H5Writer:
    """Build an HDF5 file from a tree of groups (dicts) and datasets (numpy arrays, numpy scalars,
    or lists of str / bytes for variable-length UTF-8 strings): `H5Writer().write(path, tree)`."""

    LEAF_K_MIN, INTERNAL_K = 4, 16

    def __init__(self):
        self.b = bytearray()
        self.strings = []   # global heap objects (UTF-8 bytes), heap indices 1 ..
        self.vstr = []      # (raw data offset, bytes) of string datasets: their heap IDs
        self.leaf_k = self.LEAF_K_MIN

    # ------------------------------------------------------------------ allocation
    def _alloc(self, n):
        o = len(self.b)
        self.b += b"\0" * ((n + 7) & ~7)
        return o

    def _put(self, o, fmt, *v):
        struct.pack_into(fmt, self.b, o, *v)

    # ------------------------------------------------------------------ object headers
    @staticmethod
    def _msg(mtype, data, flags=0):
        data = data + b"\0" * ((-len(data)) % 8)
        return struct.pack("<HHB3x", mtype, len(data), flags) + data

    def _header(self, msgs):
        body = b"".join(msgs)
        o = self._alloc(16 + len(body))
        self._put(o, "<BBHII", 1, 0, len(msgs), 1, len(body))
        self.b[o + 16:o + 16 + len(body)] = body
        return o

    @staticmethod
    def _dataspace(shape):
        return struct.pack("<BBBB4x", 1, len(shape), 0, 0) + b"".join(struct.pack("<Q", d) for d in shape)

    @staticmethod
    def _numeric_type(dt):
        if dt.kind in "iu":  # fixed point: byte 0 bit 3 = signed; bit offset 0, precision
            return struct.pack("<B3sI", 0x10, bytes([8 if dt.kind == "i" else 0, 0, 0]), dt.itemsize) + \
                struct.pack("<HH", 0, 8 * dt.itemsize)
        if dt.kind == "f":  # IEEE: implied mantissa msb, sign bit position, exponent/mantissa fields, bias
            if dt.itemsize == 8:
                return struct.pack("<B3sI", 0x11, bytes([0x20, 63, 0]), 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
            if dt.itemsize == 4:
                return struct.pack("<B3sI", 0x11, bytes([0x20, 31, 0]), 4) + struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
        raise FileFormatError("no HDF5 datatype for numpy dtype %s" % dt)

    @staticmethod
    def _vstr_type():
        # variable-length string (type 1), null-terminated, UTF-8; base type: 1-byte unsigned
        base = struct.pack("<B3sI", 0x10, bytes([0, 0, 0]), 1) + struct.pack("<HH", 0, 8)
        return struct.pack("<B3sI", 0x19, bytes([0x01, 0x01, 0]), 16) + base

    def _dataset(self, value):
        if isinstance(value, (list, tuple)) or (isinstance(value, np.ndarray) and value.dtype == object):
            arr = np.asarray(value, dtype=object)
            shape = arr.shape
            items = [v.encode() if isinstance(v, str) else bytes(v) for v in arr.ravel()]
            raw = bytearray()
            for sv in items:
                self.strings.append(sv)
                raw += struct.pack("<IQI", len(sv), 0, len(self.strings))  # collection address set in write()
            dtype_msg = self._vstr_type()
        else:
            arr = np.asarray(value)
            arr = arr.astype(arr.dtype.newbyteorder("<"))
            shape, raw = arr.shape, arr.tobytes()
            dtype_msg = self._numeric_type(arr.dtype)
        data = self._alloc(max(len(raw), 1))
        self.b[data:data + len(raw)] = raw
        if dtype_msg[0] == 0x19:
            self.vstr.append((data, len(raw)))
        msgs = [self._msg(0x01, self._dataspace(shape)), self._msg(0x03, dtype_msg, 1),
                self._msg(0x05, struct.pack("<BBBB", 2, 2, 2, 0)),
                self._msg(0x08, struct.pack("<BBQQ", 3, 1, data, len(raw)))]
        return self._header(msgs)

    def _group(self, tree):
        """(object header, B-tree node, local heap) of a group: one symbol node holding every
        member (the superblock's leaf K is sized for the largest group)."""
        names = sorted(tree)
        heap_data = bytearray(8)  # offset 0: the empty name
        offsets = []
        for n in names:
            offsets.append(len(heap_data))
            sv = n.encode() + b"\0"
            heap_data += sv + b"\0" * ((-len(sv)) % 8)
        members = [self._group(tree[n])[0] if isinstance(tree[n], dict) else self._dataset(tree[n]) for n in names]
        heap_seg = self._alloc(len(heap_data))
        self.b[heap_seg:heap_seg + len(heap_data)] = heap_data
        heap = self._alloc(32)
        self.b[heap:heap + 4] = b"HEAP"
        self._put(heap + 8, "<QQQ", len(heap_data), UNDEF, heap_seg)
        snod = self._alloc(8 + 40 * 2 * self.leaf_k)
        self.b[snod:snod + 4] = b"SNOD"
        self._put(snod + 4, "<BBH", 1, 0, len(names))
        for k, (off, h) in enumerate(zip(offsets, members)):
            self._put(snod + 8 + 40 * k, "<QQII", off, h, 0, 0)
        node = self._alloc(24 + (2 * self.INTERNAL_K + 1) * 8 + 2 * self.INTERNAL_K * 8)
        self.b[node:node + 4] = b"TREE"
        self._put(node + 4, "<BBHQQ", 0, 0, 1 if names else 0, UNDEF, UNDEF)
        self._put(node + 24, "<QQQ", 0, snod, offsets[-1] if offsets else 0)
        return self._header([self._msg(0x11, struct.pack("<QQ", node, heap))]), node, heap

    @staticmethod
    def _largest_group(tree):
        n = len(tree)
        for v in tree.values():
            if isinstance(v, dict):
                n = max(n, H5Writer._largest_group(v))
        return n

    def write(self, path, tree):
        self.b = bytearray(96)  # superblock with the root symbol table entry
        self.strings, self.vstr = [], []
        self.leaf_k = max(self.LEAF_K_MIN, (self._largest_group(tree) + 1) // 2)
        root, node, heap = self._group(tree)
        if self.strings:  # one global heap collection holding every string
            objs = bytearray()
            for k, sv in enumerate(self.strings):
                objs += struct.pack("<HH4xQ", k + 1, 1, len(sv)) + sv + b"\0" * ((-len(sv)) % 8)
            size = max(16 + len(objs) + 16, 4096)
            coll = self._alloc(size)
            self.b[coll:coll + 4] = b"GCOL"
            self._put(coll + 4, "<B3xQ", 1, size)
            self.b[coll + 16:coll + 16 + len(objs)] = objs
            free = coll + 16 + len(objs)
            self._put(free, "<HH4xQ", 0, 0, size - (free - coll))
            for data, n in self.vstr:
                for q in range(data, data + n, 16):
                    self._put(q + 4, "<Q", coll)
        self.b[:8] = SIGNATURE
        self._put(8, "<8B", 0, 0, 0, 0, 0, 8, 8, 0)
        self._put(16, "<HHI", self.leaf_k, self.INTERNAL_K, 0)
        self._put(24, "<QQQQ", 0, UNDEF, len(self.b), UNDEF)
        self._put(56, "<QQII", 0, root, 1, 0)
        self._put(72, "<QQ", node, heap)  # the root group's symbol table, cached in its entry
        with open(path, "wb") as f:
            f.write(bytes(self.b))
