"""A minimal HDF5 reader for scan files (the reference's src/io/hdf5.zig role): superblock
version 0 or 1, groups as symbol tables (B-tree v1 + local heap), object headers version 1 with
continuation blocks, contiguous or chunked datasets (deflate and shuffle filters), fixed-point,
floating-point and fixed-length string data, and attributes. Enough for kikuchipy's and
EDAX/Bruker's h5ebsd files as written by h5py; anything else is refused with an error naming
what it met.

`File(path)` opens a file; `f["a/b/c"]` is a group or a dataset; `ds.read()` gives a numpy array,
`ds.read_rows(i0, i1)` the rows [i0, i1) of its first axis (reading only the chunks they touch).
"""
import struct
import zlib

import numpy as np


class H5Error(ValueError):
    pass


UNDEF = 0xFFFFFFFFFFFFFFFF


class File:
    def __init__(self, path):
        self.path = path
        self._f = open(path, "rb")
        sb = self._read(0, 96)
        if sb[:8] != b"\x89HDF\r\n\x1a\n":
            raise H5Error(f"{path}: not an HDF5 file")
        ver = sb[8]
        if ver not in (0, 1):
            raise H5Error(f"{path}: superblock version {ver} (this reader takes 0 and 1)")
        self.so, self.sl = sb[13], sb[14]
        if self.so != 8 or self.sl != 8:
            raise H5Error(f"{path}: {self.so}-byte offsets, {self.sl}-byte lengths (8 and 8 taken)")
        p = 24 if ver == 0 else 28
        self.base = self._u(sb, p, 8)
        p += 32  # base, free space, end of file, driver info
        # the root group's symbol table entry
        self.root = self._entry(sb, p)

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ── bytes ──

    def _read(self, addr, n):
        self._f.seek(addr)
        b = self._f.read(n)
        if len(b) < n:
            raise H5Error(f"{self.path}: read past the end at {addr}")
        return b

    @staticmethod
    def _u(b, p, n):
        return int.from_bytes(b[p:p + n], "little")

    def _entry(self, b, p):
        """A symbol table entry: (name offset, object header address, cache type, scratch)."""
        return (self._u(b, p, 8), self._u(b, p + 8, 8), self._u(b, p + 16, 4), b[p + 24:p + 40])

    # ── groups ──

    def _heap_data(self, addr):
        h = self._read(addr, 32)
        if h[:4] != b"HEAP":
            raise H5Error(f"{self.path}: no local heap at {addr}")
        size, data = self._u(h, 8, 8), self._u(h, 24, 8)
        return self._read(data, size)

    def _group_links(self, btree, heap):
        """name -> object header address, from a group's B-tree (type 0) and local heap."""
        names = self._heap_data(heap)
        out = {}
        stack = [btree]
        while stack:
            addr = stack.pop()
            n = self._read(addr, 24)
            if n[:4] != b"TREE" or n[4] != 0:
                raise H5Error(f"{self.path}: no group B-tree node at {addr}")
            level, used = n[5], self._u(n, 6, 2)
            body = self._read(addr + 24, (2 * used + 1) * 8)
            children = [self._u(body, 8 + 16 * i, 8) for i in range(used)]
            if level > 0:
                stack.extend(children)
                continue
            for c in children:
                s = self._read(c, 8)
                if s[:4] != b"SNOD":
                    raise H5Error(f"{self.path}: no symbol node at {c}")
                k = self._u(s, 6, 2)
                ents = self._read(c + 8, 40 * k)
                for i in range(k):
                    off, oh, _, _ = self._entry(ents, 40 * i)
                    end = names.index(b"\0", off)
                    out[names[off:end].decode()] = oh
        return out

    def _object(self, oh, name):
        msgs = self._messages(oh)
        types = {t for t, _ in msgs}
        if 0x11 in types:  # symbol table message: a group
            d = next(d for t, d in msgs if t == 0x11)
            return Group(self, name, self._u(d, 0, 8), self._u(d, 8, 8), msgs)
        if 0x08 in types:
            return Dataset(self, name, msgs)
        raise H5Error(f"{self.path}: {name}: neither a group nor a dataset (messages {sorted(types)})")

    def _messages(self, addr):
        """The header messages (type, data) of a version-1 object header, continuations followed."""
        h = self._read(addr, 16)
        if h[0] != 1:
            raise H5Error(f"{self.path}: object header version {h[0]} at {addr} (this reader takes 1)")
        count = self._u(h, 2, 2)
        blocks = [(addr + 16, self._u(h, 8, 4))]
        out = []
        while blocks and len(out) < count:
            start, size = blocks.pop(0)
            b = self._read(start, size)
            p = 0
            while p + 8 <= size and len(out) < count:
                t, n = self._u(b, p, 2), self._u(b, p + 2, 2)
                d = b[p + 8:p + 8 + n]
                if t == 0x10:
                    blocks.append((self._u(d, 0, 8), self._u(d, 8, 8)))
                out.append((t, d))
                p += 8 + n
        return out

    def __getitem__(self, path):
        _, oh, ctype, scratch = self.root
        obj = self._object(oh, "/")
        for part in [p for p in path.split("/") if p]:
            if not isinstance(obj, Group):
                raise H5Error(f"{self.path}: {obj.name} is not a group")
            obj = obj[part]
        return obj


class Group:
    def __init__(self, f, name, btree, heap, msgs):
        self.f, self.name = f, name
        self._links = f._group_links(btree, heap)
        self.attrs = _attributes(f, msgs)

    def keys(self):
        return list(self._links)

    def __contains__(self, k):
        return k in self._links

    def __getitem__(self, k):
        if "/" in k:
            obj = self
            for part in [p for p in k.split("/") if p]:
                obj = obj[part]
            return obj
        if k not in self._links:
            raise KeyError(f"{self.name}/{k}")
        return self.f._object(self._links[k], f"{self.name.rstrip('/')}/{k}")


def _dtype(d):
    """numpy dtype of a datatype message."""
    cls, ver = d[0] & 0x0F, d[0] >> 4
    bits = d[1] | (d[2] << 8) | (d[3] << 16)
    size = int.from_bytes(d[4:8], "little")
    order = ">" if bits & 1 else "<"
    if cls == 0:
        return np.dtype(f"{order}{'i' if bits & 8 else 'u'}{size}")
    if cls == 1:
        return np.dtype(f"{order}f{size}")
    if cls == 3:
        return np.dtype(f"S{size}")
    if cls == 8:  # an enumeration (h5py's booleans): its base type
        return _dtype(d[8:])
    raise H5Error(f"datatype class {cls} (fixed-point, floating-point, enumerations and fixed strings are read)")


def _shape(d):
    ver, rank, flags = d[0], d[1], d[2]
    p = 8 if ver == 1 else 4
    return tuple(int.from_bytes(d[p + 8 * i:p + 8 * i + 8], "little") for i in range(rank))


def _pad8(n):
    return (n + 7) & ~7


def _attributes(f, msgs):
    out = {}
    for t, d in msgs:
        if t != 0x0C:
            continue
        if d[0] != 1:
            continue  # versions 2 and 3 of the attribute message are not read
        nn, tn, sn = int.from_bytes(d[2:4], "little"), int.from_bytes(d[4:6], "little"), int.from_bytes(d[6:8], "little")
        p = 8
        name = d[p:p + nn].split(b"\0")[0].decode()
        p += _pad8(nn)
        try:
            dt = _dtype(d[p:p + tn])
        except H5Error:
            continue
        p += _pad8(tn)
        shape = _shape(d[p:p + sn])
        p += _pad8(sn)
        count = int(np.prod(shape)) if shape else 1
        v = np.frombuffer(d[p:p + count * dt.itemsize], dt).reshape(shape)
        out[name] = v
    return out


class Dataset:
    def __init__(self, f, name, msgs):
        self.f, self.name = f, name
        m = dict()
        for t, d in msgs:
            m.setdefault(t, d)
        self.shape = _shape(m[0x01]) if 0x01 in m else ()
        try:
            self.dtype = _dtype(m[0x03])
            self._dtype_error = None
        except H5Error as e:  # kept for read(): the rest of the file stays readable
            self.dtype, self._dtype_error = None, e
        self.attrs = _attributes(f, msgs)
        lay = m[0x08]
        if lay[0] != 3:
            raise H5Error(f"{name}: data layout version {lay[0]} (3 is read)")
        self.layout = lay[1]
        self.filters = []
        if 0x0B in m:
            d = m[0x0B]
            if d[0] != 1:
                raise H5Error(f"{name}: filter pipeline version {d[0]}")
            p = 8
            for _ in range(d[1]):
                fid, nlen, _flags, ncv = struct.unpack_from("<HHHH", d, p)
                p += 8 + _pad8(nlen)
                cv = struct.unpack_from(f"<{ncv}I", d, p)
                p += 4 * ncv + (4 if ncv % 2 else 0)
                self.filters.append((fid, cv))
            for fid, _ in self.filters:
                if fid not in (1, 2):
                    raise H5Error(f"{name}: filter {fid} (deflate and shuffle are read)")
        if self.layout == 0:
            n = int.from_bytes(lay[2:4], "little")
            self._compact = lay[4:4 + n]
        elif self.layout == 1:
            self._addr = int.from_bytes(lay[2:10], "little")
            self._size = int.from_bytes(lay[10:18], "little")
        elif self.layout == 2:
            rank = lay[2]
            self._btree = int.from_bytes(lay[3:11], "little")
            self._chunk = tuple(int.from_bytes(lay[11 + 4 * i:15 + 4 * i], "little") for i in range(rank - 1))
        else:
            raise H5Error(f"{name}: layout class {self.layout}")

    def read(self):
        if self._dtype_error is not None:
            raise H5Error(f"{self.name}: {self._dtype_error}")
        return self.read_rows(0, self.shape[0] if self.shape else 1) if self.shape else self._whole().reshape(())

    def _whole(self):
        n = int(np.prod(self.shape)) if self.shape else 1
        if self.layout == 0:
            return np.frombuffer(self._compact[:n * self.dtype.itemsize], self.dtype)
        if self._addr == UNDEF:
            return np.zeros(n, self.dtype)
        return np.frombuffer(self.f._read(self.f.base + self._addr, n * self.dtype.itemsize), self.dtype)

    def _decode(self, raw, nbytes):
        for fid, cv in reversed(self.filters):
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                sz = self.dtype.itemsize
                a = np.frombuffer(raw, np.uint8)
                raw = a.reshape(sz, -1).T.copy().tobytes()
        return raw[:nbytes]

    def _chunks(self):
        """(offsets, address, size, filter mask) of every chunk, from the chunk B-tree (type 1)."""
        rank = len(self._chunk)
        ksize = 8 + 8 * (rank + 1)
        out = []
        stack = [self._btree]
        while stack:
            addr = stack.pop()
            n = self.f._read(addr, 24)
            if n[:4] != b"TREE" or n[4] != 1:
                raise H5Error(f"{self.name}: no chunk B-tree node at {addr}")
            level, used = n[5], int.from_bytes(n[6:8], "little")
            body = self.f._read(addr + 24, used * (ksize + 8) + ksize)
            for i in range(used):
                k = i * (ksize + 8)
                size, mask = struct.unpack_from("<II", body, k)
                offs = struct.unpack_from(f"<{rank}Q", body, k + 8)
                child = int.from_bytes(body[k + ksize:k + ksize + 8], "little")
                if level > 0:
                    stack.append(child)
                else:
                    out.append((offs, child, size, mask))
        return out

    def read_rows(self, i0, i1):
        """Rows [i0, i1) of the first axis."""
        if self._dtype_error is not None:
            raise H5Error(f"{self.name}: {self._dtype_error}")
        shape = self.shape
        if self.layout != 2:
            a = self._whole().reshape(shape)
            return a[i0:i1].copy()
        out = np.zeros((i1 - i0,) + tuple(shape[1:]), self.dtype)
        csh = self._chunk
        nbytes = int(np.prod(csh)) * self.dtype.itemsize
        for offs, addr, size, mask in self._chunks():
            if offs[0] + csh[0] <= i0 or offs[0] >= i1:
                continue
            raw = self.f._read(self.f.base + addr, size)
            if mask == 0:
                raw = self._decode(raw, nbytes)
            elif self.filters:
                raise H5Error(f"{self.name}: chunk with filters skipped (mask {mask})")
            blk = np.frombuffer(raw[:nbytes], self.dtype).reshape(csh)
            src = tuple(slice(max(i0 - offs[0], 0) if d == 0 else 0, min(csh[d], (i1 - offs[0]) if d == 0 else shape[d] - offs[d]))
                        for d in range(len(shape)))
            dst = tuple(slice(offs[0] + src[0].start - i0, offs[0] + src[0].stop - i0) if d == 0 else
                        slice(offs[d], offs[d] + src[d].stop) for d in range(len(shape)))
            out[dst] = blk[src]
        return out
