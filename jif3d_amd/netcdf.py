"""Reading and writing the netCDF files jif3D keeps its data and models in.

jif3D writes through netCDF-C++ (`NcFile::replace`), i.e. netCDF-4 files, which are HDF5
files; netCDF-C also reads classic files.  This module has no dependency beyond numpy
and zlib:

* `Dataset(path)` reads classic netCDF (CDF-1, CDF-2, CDF-5) and the subset of HDF5 that
  netCDF-4 writes for jif3D's variables: superblock versions 0-3, object headers 1 and 2,
  groups as symbol tables or compact links, contiguous / compact / chunked storage
  (version-1 B-tree chunk index, deflate and shuffle filters), fixed-point, floating-point
  and string data, and the DIMENSION_LIST references that name a variable's dimensions.
  Groups in dense (fractal-heap) storage and the version-4 chunk indexes other than a
  single chunk are refused with an error.
* `write_classic(path, dims, variables, attrs)` writes a CDF-2 classic file (readable by
  netCDF-C, hence by jif3D's readers).

The layouts follow the published netCDF classic format specification and the HDF5 file
format specification.
"""
import struct
import zlib

import numpy as np

__all__ = ["Dataset", "Variable", "write_classic"]


class Variable:
    def __init__(self, name, dims, shape, attrs, reader):
        self.name = name
        self.dims = tuple(dims)
        self.shape = tuple(int(s) for s in shape)
        self.attrs = attrs
        self._reader = reader

    def read(self):
        return self._reader()

    def __repr__(self):
        return f"Variable({self.name!r}, dims={self.dims}, shape={self.shape})"


class Dataset:
    """Variables (name -> Variable), dimensions (name -> size) and global attributes."""

    def __init__(self, path):
        with open(path, "rb") as f:
            buf = f.read()
        if buf[:3] == b"CDF":
            self.dimensions, self.variables, self.attrs = _read_classic(buf)
            self.format = "classic"
        else:
            self.dimensions, self.variables, self.attrs = _H5File(buf).netcdf()
            self.format = "netcdf4"

    def __contains__(self, name):
        return name in self.variables

    def __getitem__(self, name):
        return self.variables[name]


# ---------------------------------------------------------------------------
# classic format
# ---------------------------------------------------------------------------
_NC_TYPES = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8", 7: ">u1", 8: ">u2",
             9: ">u4", 10: ">i8", 11: ">u8"}
_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 10, 11, 12


def _read_classic(buf):
    version = buf[3]
    if version not in (1, 2, 5):
        raise ValueError(f"unknown classic netCDF version {version}")
    p = [4]
    big = version == 5

    def u32():
        v = struct.unpack_from(">I", buf, p[0])[0]
        p[0] += 4
        return v

    def u64():
        v = struct.unpack_from(">Q", buf, p[0])[0]
        p[0] += 8
        return v

    count = u64 if big else u32

    def name():
        n = count()
        s = buf[p[0]:p[0] + n].decode("utf-8")
        p[0] += (n + 3) // 4 * 4
        return s

    def values(nctype, n):
        dt = np.dtype(_NC_TYPES[nctype])
        nbytes = dt.itemsize * n
        raw = buf[p[0]:p[0] + nbytes]
        p[0] += (nbytes + 3) // 4 * 4
        if nctype == 2:
            return raw.rstrip(b"\0").decode("utf-8", "replace")
        return np.frombuffer(raw, dtype=dt).astype(dt.newbyteorder("="))

    def attr_list():
        tag, n = u32(), count()
        out = {}
        if tag == 0 and n == 0:
            return out
        if tag != _NC_ATTRIBUTE:
            raise ValueError("corrupt classic netCDF header (attributes)")
        for _ in range(n):
            nm = name()
            t = u32()
            out[nm] = values(t, count())
        return out

    numrecs = count()
    tag, n = u32(), count()
    dims = []
    if not (tag == 0 and n == 0):
        if tag != _NC_DIMENSION:
            raise ValueError("corrupt classic netCDF header (dimensions)")
        for _ in range(n):
            nm = name()
            dims.append((nm, count()))
    gattrs = attr_list()
    tag, n = u32(), count()
    variables, recvars = {}, []
    if not (tag == 0 and n == 0):
        if tag != _NC_VARIABLE:
            raise ValueError("corrupt classic netCDF header (variables)")
        for _ in range(n):
            nm = name()
            ids = [count() for _ in range(count())]
            attrs = attr_list()
            t = u32()
            vsize = count()
            begin = u64() if version in (2, 5) else u32()
            vdims = [dims[i][0] for i in ids]
            shape = [dims[i][1] for i in ids]
            is_rec = bool(ids) and dims[ids[0]][1] == 0
            if is_rec:
                shape[0] = numrecs
                recvars.append(nm)
            variables[nm] = dict(dims=vdims, shape=shape, attrs=attrs, type=t, begin=begin,
                                 vsize=vsize, rec=is_rec)
    recsize = sum(variables[v]["vsize"] for v in recvars)
    if len(recvars) == 1:
        recsize = int(np.prod(variables[recvars[0]]["shape"][1:], dtype=np.int64)) * \
            np.dtype(_NC_TYPES[variables[recvars[0]]["type"]]).itemsize
    out = {}
    for nm, v in variables.items():
        out[nm] = Variable(nm, v["dims"], v["shape"], v["attrs"],
                           _classic_reader(buf, v, recsize))
    return {d: s if s else numrecs for d, s in dims}, out, gattrs


def _classic_reader(buf, v, recsize):
    def read():
        dt = np.dtype(_NC_TYPES[v["type"]])
        if v["type"] == 2:
            dt = np.dtype("S1")
        shape = v["shape"]
        if not v["rec"]:
            n = int(np.prod(shape, dtype=np.int64))
            a = np.frombuffer(buf, dtype=dt, count=n, offset=v["begin"])
            return a.astype(dt.newbyteorder("=")).reshape(shape)
        per = int(np.prod(shape[1:], dtype=np.int64))
        recs = [np.frombuffer(buf, dtype=dt, count=per, offset=v["begin"] + r * recsize)
                for r in range(shape[0])]
        a = np.concatenate(recs) if recs else np.zeros(0, dtype=dt)
        return a.astype(dt.newbyteorder("=")).reshape(shape)
    return read


def write_classic(path, dims, variables, attrs=None):
    """Write a CDF-2 (64-bit offset) classic file.

    dims: ordered {name: size}; variables: ordered {name: (dim names, array, attrs)}
    with float64 / float32 / int32 / int16 / int8 arrays; attrs: global attributes
    (str or numeric)."""
    dim_names = list(dims)
    out = bytearray(b"CDF\x02")

    def u32(v):
        out.extend(struct.pack(">I", v))

    def name(s):
        b = s.encode("utf-8")
        u32(len(b))
        out.extend(b + b"\0" * ((4 - len(b) % 4) % 4))

    def attr_block(a):
        if not a:
            u32(0)
            u32(0)
            return
        u32(_NC_ATTRIBUTE)
        u32(len(a))
        for k, v in a.items():
            name(k)
            if isinstance(v, str):
                b = v.encode("utf-8")
                u32(2)
                u32(len(b))
                out.extend(b + b"\0" * ((4 - len(b) % 4) % 4))
            else:
                arr = np.atleast_1d(np.asarray(v))
                t = _nc_type(arr.dtype)
                u32(t)
                u32(arr.size)
                raw = arr.astype(np.dtype(_NC_TYPES[t])).tobytes()
                out.extend(raw + b"\0" * ((4 - len(raw) % 4) % 4))

    u32(0)  # numrecs: no record dimension
    if dims:
        u32(_NC_DIMENSION)
        u32(len(dims))
        for d in dim_names:
            name(d)
            u32(int(dims[d]))
    else:
        u32(0)
        u32(0)
    attr_block(attrs or {})
    # header size first (begin offsets depend on it), then the data
    entries = []
    for vname, (vdims, arr, vattrs) in variables.items():
        arr = np.asarray(arr)
        t = _nc_type(arr.dtype)
        shape = tuple(int(dims[d]) for d in vdims)
        if arr.size != int(np.prod(shape, dtype=np.int64)):
            raise ValueError(f"variable {vname}: {arr.size} values for shape {shape}")
        raw = np.ascontiguousarray(arr).astype(np.dtype(_NC_TYPES[t])).tobytes()
        raw += b"\0" * ((4 - len(raw) % 4) % 4)
        entries.append((vname, vdims, vattrs or {}, t, raw))
    header_tail = bytearray()
    saved = out
    out = header_tail
    if entries:
        u32(_NC_VARIABLE)
        u32(len(entries))
    else:
        u32(0)
        u32(0)
    begins_at = []
    for vname, vdims, vattrs, t, raw in entries:
        name(vname)
        u32(len(vdims))
        for d in vdims:
            u32(dim_names.index(d))
        attr_block(vattrs)
        u32(t)
        u32(min(len(raw), 0xFFFFFFFF))
        begins_at.append(len(out))
        out.extend(b"\0" * 8)
    out = saved
    header_len = len(out) + len(header_tail)
    pos = header_len
    for (vname, vdims, vattrs, t, raw), at in zip(entries, begins_at):
        struct.pack_into(">Q", header_tail, at, pos)
        pos += len(raw)
    out.extend(header_tail)
    for e in entries:
        out.extend(e[4])
    with open(path, "wb") as f:
        f.write(bytes(out))


def _nc_type(dt):
    dt = np.dtype(dt)
    for code, s in ((6, "f8"), (5, "f4"), (4, "i4"), (3, "i2"), (1, "i1")):
        if dt == np.dtype(s):
            return code
    if dt.kind in "iu":
        return 4 if dt.itemsize <= 4 else 10
    if dt.kind == "f":
        return 6
    raise ValueError(f"no classic netCDF type for {dt}")


# ---------------------------------------------------------------------------
# HDF5 (netCDF-4)
# ---------------------------------------------------------------------------
_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = None


class _H5File:
    def __init__(self, buf):
        self.b = buf
        base = 0
        while buf[base:base + 8] != _SIG:
            base = 512 if base == 0 else 2 * base
            if base + 8 > len(buf):
                raise ValueError("not a netCDF file (no classic or HDF5 signature)")
        v = buf[base + 8]
        if v in (0, 1):
            self.so, self.sl = buf[base + 13], buf[base + 14]
            p = base + 24 + (4 if v == 1 else 0)
            self.base = self._o(p)
            p += 4 * self.so
            # root group symbol table entry: link name offset, object header address
            self.root = self._o(p + self.so)
        elif v in (2, 3):
            self.so, self.sl = buf[base + 9], buf[base + 10]
            p = base + 12
            self.base = self._o(p)
            self.root = self._o(p + 3 * self.so)
        else:
            raise ValueError(f"HDF5 superblock version {v} not supported")
        self._names = {}

    # primitive readers ------------------------------------------------------
    def _u(self, p, n):
        return int.from_bytes(self.b[p:p + n], "little")

    def _o(self, p):
        v = self._u(p, self.so)
        return None if v == (1 << (8 * self.so)) - 1 else v

    def _l(self, p):
        return self._u(p, self.sl)

    def _addr(self, a):
        return a + (self.base or 0)

    # object headers -----------------------------------------------------------
    def messages(self, addr):
        a = self._addr(addr)
        b = self.b
        out = []
        if b[a:a + 4] == b"OHDR":
            flags = b[a + 5]
            p = a + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            nsz = 1 << (flags & 3)
            size = self._u(p, nsz)
            p += nsz
            blocks = [(p, p + size)]
            crt = bool(flags & 0x04)
            while blocks:
                s, e = blocks.pop(0)
                q = s
                while q + 4 <= e:
                    t, n, mf = b[q], self._u(q + 1, 2), b[q + 3]
                    q += 4 + (2 if crt else 0)
                    if q + n > e:
                        break
                    out.append((t, q, n, mf))
                    if t == 0x10:
                        ca, cl = self._o(q), self._l(q + self.so)
                        ca = self._addr(ca)
                        if b[ca:ca + 4] != b"OCHK":
                            raise ValueError("corrupt HDF5 continuation block")
                        blocks.append((ca + 4, ca + cl - 4))
                    q += n
            return out
        if b[a] != 1:
            raise ValueError(f"HDF5 object header version {b[a]} not supported")
        nmsg, size = self._u(a + 2, 2), self._u(a + 8, 4)
        blocks = [(a + 16, a + 16 + size)]
        while blocks and len(out) < nmsg:
            s, e = blocks.pop(0)
            q = s
            while q + 8 <= e and len(out) < nmsg:
                t, n, mf = self._u(q, 2), self._u(q + 2, 2), b[q + 4]
                q += 8
                out.append((t, q, n, mf))
                if t == 0x10:
                    ca, cl = self._addr(self._o(q)), self._l(q + self.so)
                    blocks.append((ca, ca + cl))
                q += n
        return out

    # groups -------------------------------------------------------------------
    def links(self, addr):
        """{name: object header address} of a group."""
        out = {}
        for t, p, n, mf in self.messages(addr):
            if t == 0x11:  # symbol table: v1 B-tree of symbol nodes + local heap
                self._symbol_table(self._o(p), self._o(p + self.so), out)
            elif t == 0x06:
                self._link(p, out)
            elif t == 0x02:
                flags = self.b[p + 1]
                q = p + 2 + (8 if flags & 1 else 0)
                if self._o(q) is not None:
                    raise NotImplementedError("HDF5 groups in dense link storage")
        return out

    def _symbol_table(self, btree, heap, out):
        h = self._addr(heap)
        if self.b[h:h + 4] != b"HEAP":
            raise ValueError("corrupt HDF5 local heap")
        data = self._addr(self._o(h + 8 + 2 * self.sl))

        def walk(node):
            a = self._addr(node)
            if self.b[a:a + 4] == b"SNOD":
                nsym = self._u(a + 6, 2)
                q = a + 8
                for _ in range(nsym):
                    off, oh = self._l(q), self._o(q + self.so)
                    end = self.b.index(b"\0", data + off)
                    out[self.b[data + off:end].decode("utf-8")] = oh
                    q += 2 * self.so + 24
                return
            if self.b[a:a + 4] != b"TREE" or self.b[a + 4] != 0:
                raise ValueError("corrupt HDF5 group B-tree")
            used = self._u(a + 6, 2)
            q = a + 8 + 2 * self.so
            for i in range(used):
                q += self.sl  # key
                walk(self._o(q))
                q += self.so
        walk(btree)

    def _link(self, p, out):
        b = self.b
        flags = b[p + 1]
        q = p + 2
        ltype = 0
        if flags & 0x08:
            ltype = b[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        nl = 1 << (flags & 3)
        ln = self._u(q, nl)
        q += nl
        nm = b[q:q + ln].decode("utf-8")
        q += ln
        if ltype == 0:
            out[nm] = self._o(q)

    # datasets -----------------------------------------------------------------
    def dataset(self, addr):
        info = dict(shape=None, dtype=None, layout=None, filters=[], attrs={})
        for t, p, n, mf in self.messages(addr):
            if mf & 0x02 and t in (0x03, 0x01):
                raise NotImplementedError("shared HDF5 datatype or dataspace messages")
            if t == 0x01:
                info["shape"] = self._dataspace(p)
            elif t == 0x03:
                info["dtype"] = self._datatype(p)
            elif t == 0x08:
                info["layout"] = self._layout(p)
            elif t == 0x0B:
                info["filters"] = self._filters(p)
            elif t == 0x0C:
                k, v = self._attribute(p)
                info["attrs"][k] = v
            elif t == 0x15:
                flags = self.b[p + 1]
                q = p + 2 + (2 if flags & 1 else 0)
                if self._o(q) is not None:
                    raise NotImplementedError("HDF5 attributes in dense storage")
        return info

    def _dataspace(self, p):
        b = self.b
        ver, rank, flags = b[p], b[p + 1], b[p + 2]
        if ver == 1:
            q = p + 8
        elif ver == 2:
            if b[p + 3] == 2:
                return None  # null dataspace
            q = p + 4
        else:
            raise ValueError(f"HDF5 dataspace version {ver}")
        return tuple(self._l(q + i * self.sl) for i in range(rank))

    def _datatype(self, p):
        b = self.b
        cls = b[p] & 0x0F
        bf = b[p + 1]
        size = self._u(p + 4, 4)
        order = ">" if bf & 1 else "<"
        if cls == 0:
            signed = bool(bf & 0x08)
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}")
        if cls == 1:
            return np.dtype(f"{order}f{size}")
        if cls == 3:
            return ("str", size)
        if cls == 9:
            base = self._datatype(p + 8)
            return ("vlen", bf & 0x0F, base)
        if cls == 7:
            return ("ref", size)
        return ("unsupported", cls, size)

    def _layout(self, p):
        b = self.b
        ver = b[p]
        if ver not in (3, 4):
            raise NotImplementedError(f"HDF5 data layout version {ver}")
        cls = b[p + 1]
        q = p + 2
        if cls == 0:
            n = self._u(q, 2)
            return ("compact", q + 2, n)
        if cls == 1:
            return ("contiguous", self._o(q), self._l(q + self.so))
        if cls != 2:
            raise NotImplementedError(f"HDF5 layout class {cls}")
        if ver == 3:
            rank = b[q]
            addr = self._o(q + 1)
            q += 1 + self.so
            dims = [self._u(q + 4 * i, 4) for i in range(rank)]
            return ("chunked_btree", addr, dims)
        flags, rank, enc = b[q], b[q + 1], b[q + 2]
        q += 3
        dims = [self._u(q + enc * i, enc) for i in range(rank)]
        q += enc * rank
        itype = b[q]
        q += 1
        if itype == 1:
            fsize = fmask = None
            if flags & 0x02:
                fsize, fmask = self._l(q), self._u(q + self.sl, 4)
                q += self.sl + 4
            return ("single_chunk", self._o(q), dims, fsize, fmask)
        if itype == 2:
            return ("implicit", self._o(q), dims)
        raise NotImplementedError(f"HDF5 chunk index type {itype}")

    def _filters(self, p):
        b = self.b
        ver, nf = b[p], b[p + 1]
        q = p + (8 if ver == 1 else 2)
        out = []
        for _ in range(nf):
            fid = self._u(q, 2)
            q += 2
            nlen = 0
            if ver == 1 or fid >= 256:
                nlen = self._u(q, 2)
                q += 2
            q += 2  # flags
            ncd = self._u(q, 2)
            q += 2
            if ver == 1:
                q += (nlen + 7) // 8 * 8
            else:
                q += nlen
            cd = [self._u(q + 4 * i, 4) for i in range(ncd)]
            q += 4 * ncd
            if ver == 1 and ncd % 2:
                q += 4
            out.append((fid, cd))
        return out

    def _attribute(self, p):
        b = self.b
        ver = b[p]
        nsz, tsz, ssz = self._u(p + 2, 2), self._u(p + 4, 2), self._u(p + 6, 2)
        q = p + 8 + (1 if ver == 3 else 0)
        pad = (lambda n: (n + 7) // 8 * 8) if ver == 1 else (lambda n: n)
        name = b[q:q + nsz].split(b"\0")[0].decode("utf-8")
        q += pad(nsz)
        dt = self._datatype(q)
        q += pad(tsz)
        shape = self._dataspace(q)
        q += pad(ssz)
        n = int(np.prod(shape, dtype=np.int64)) if shape else (0 if shape is None else 1)
        return name, self._values(dt, q, n, shape)

    def _values(self, dt, q, n, shape):
        b = self.b
        if isinstance(dt, np.dtype):
            a = np.frombuffer(b, dtype=dt, count=n, offset=q).astype(dt.newbyteorder("="))
            return a.reshape(shape) if shape else (a[0] if n else a)
        if dt[0] == "str":
            size = dt[1]
            vals = [b[q + i * size:q + (i + 1) * size].split(b"\0")[0].decode("utf-8", "replace")
                    for i in range(n)]
            return vals[0] if n == 1 else vals
        if dt[0] == "vlen":
            out = []
            step = 4 + self.so + 4
            for i in range(n):
                e = q + i * step
                length, coll, idx = self._u(e, 4), self._o(e + 4), self._u(e + 4 + self.so, 4)
                data = self._global_heap(coll, idx) if coll is not None else b""
                base = dt[2]
                if dt[1] == 1:  # string
                    out.append(data.decode("utf-8", "replace"))
                elif isinstance(base, tuple) and base[0] == "ref":
                    out.append([int.from_bytes(data[k * self.so:(k + 1) * self.so], "little")
                                for k in range(length)])
                elif isinstance(base, np.dtype):
                    out.append(np.frombuffer(data, dtype=base, count=length))
                else:
                    out.append(data)
            return out
        return None

    def _global_heap(self, coll, idx):
        a = self._addr(coll)
        if self.b[a:a + 4] != b"GCOL":
            raise ValueError("corrupt HDF5 global heap")
        size = self._l(a + 8)
        q, end = a + 8 + self.sl, a + size
        while q + 8 + self.sl <= end:
            oid, osz = self._u(q, 2), self._l(q + 8)
            if oid == 0:
                break
            if oid == idx:
                return self.b[q + 8 + self.sl:q + 8 + self.sl + osz]
            q += 8 + self.sl + (osz + 7) // 8 * 8
        raise ValueError("HDF5 global heap object not found")

    def read(self, info):
        dt, shape = info["dtype"], info["shape"]
        if not isinstance(dt, np.dtype):
            raise NotImplementedError(f"HDF5 dataset type {dt}")
        if shape is None:
            return np.zeros(0, dtype=dt.newbyteorder("="))
        n = int(np.prod(shape, dtype=np.int64)) if shape else 1
        lay = info["layout"]
        if lay[0] == "compact":
            raw = self.b[lay[1]:lay[1] + n * dt.itemsize]
        elif lay[0] == "contiguous":
            if lay[1] is None:
                raw = b"\0" * (n * dt.itemsize)
            else:
                a = self._addr(lay[1])
                raw = self.b[a:a + n * dt.itemsize]
        elif lay[0] in ("single_chunk", "implicit", "chunked_btree"):
            return self._read_chunked(info, dt, shape)
        else:
            raise NotImplementedError(lay[0])
        a = np.frombuffer(raw, dtype=dt, count=n).astype(dt.newbyteorder("="))
        return a.reshape(shape) if shape else a

    def _unfilter(self, raw, filters, mask):
        for i, (fid, cd) in reversed(list(enumerate(filters))):
            if mask & (1 << i):
                continue
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                es = cd[0] if cd else 1
                a = np.frombuffer(raw, dtype=np.uint8)
                nel = a.size // es
                body = a[:nel * es].reshape(es, nel).T.reshape(-1)
                raw = body.tobytes() + raw[nel * es:]
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise NotImplementedError(f"HDF5 filter {fid}")
        return raw

    def _read_chunked(self, info, dt, shape):
        lay = info["layout"]
        rank = len(shape)
        cdims = list(lay[2][:rank])
        out = np.zeros(shape, dtype=dt.newbyteorder("="))
        csize = int(np.prod(cdims, dtype=np.int64)) * dt.itemsize

        def place(offsets, raw):
            a = np.frombuffer(raw[:csize], dtype=dt).reshape(cdims).astype(out.dtype)
            sl_out = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offsets, cdims, shape))
            sl_in = tuple(slice(0, x.stop - x.start) for x in sl_out)
            out[sl_out] = a[sl_in]

        if lay[0] == "single_chunk":
            a = self._addr(lay[1])
            size = lay[3] if lay[3] is not None else csize
            raw = self._unfilter(self.b[a:a + size], info["filters"], lay[4] or 0)
            place([0] * rank, raw)
            return out
        if lay[0] == "implicit":
            a = self._addr(lay[1])
            grid = [(s + c - 1) // c for s, c in zip(shape, cdims)]
            for k, idx in enumerate(np.ndindex(*grid)):
                place([i * c for i, c in zip(idx, cdims)], self.b[a + k * csize:a + (k + 1) * csize])
            return out

        def walk(node):
            a = self._addr(node)
            if self.b[a:a + 4] != b"TREE" or self.b[a + 4] != 1:
                raise ValueError("corrupt HDF5 chunk B-tree")
            level, used = self.b[a + 5], self._u(a + 6, 2)
            q = a + 8 + 2 * self.so
            ksz = 8 + 8 * (rank + 1)
            for _ in range(used):
                size, mask = self._u(q, 4), self._u(q + 4, 4)
                offs = [self._u(q + 8 + 8 * d, 8) for d in range(rank)]
                child = self._o(q + ksz)
                if level > 0:
                    walk(child)
                else:
                    c = self._addr(child)
                    place(offs, self._unfilter(self.b[c:c + size], info["filters"], mask))
                q += ksz + self.so
        walk(lay[1])
        return out

    # netCDF view ---------------------------------------------------------------
    def netcdf(self):
        root_msgs_attrs = {}
        for t, p, n, mf in self.messages(self.root):
            if t == 0x0C:
                k, v = self._attribute(p)
                root_msgs_attrs[k] = v
        entries = self.links(self.root)
        infos = {nm: self.dataset(a) for nm, a in entries.items()}
        addr_name = {a: nm for nm, a in entries.items()}
        dims, variables = {}, {}
        for nm, info in infos.items():
            attrs = info["attrs"]
            if attrs.get("CLASS") == "DIMENSION_SCALE":
                dims[nm] = int(info["shape"][0]) if info["shape"] else 0
        for nm, info in infos.items():
            attrs = info["attrs"]
            label = attrs.get("NAME", "")
            if isinstance(label, str) and label.startswith(
                    "This is a netCDF dimension but not a netCDF variable"):
                continue
            shape = info["shape"] or ()
            if attrs.get("CLASS") == "DIMENSION_SCALE":
                vdims = [nm]
            else:
                refs = attrs.get("DIMENSION_LIST")
                vdims = [addr_name.get(r[0], "") if r else "" for r in refs] if refs else []
            public = {k: v for k, v in attrs.items()
                      if k not in ("CLASS", "NAME", "DIMENSION_LIST", "REFERENCE_LIST")
                      and not k.startswith("_Netcdf4")}
            variables[nm] = Variable(nm, vdims, shape, public,
                                     (lambda i=info: self.read(i)))
        gattrs = {k: v for k, v in root_msgs_attrs.items() if not k.startswith("_")}
        return dims, variables, gattrs
