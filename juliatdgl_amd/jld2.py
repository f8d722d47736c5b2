"""A minimal reader for the JLD2 0.2 files that hold the reference's golden states.

JLD2 files are HDF5 files (superblock v2/v3, version-2 object headers, a user block of 512
bytes).  This reader understands exactly what those files use: compact link storage in the
root group, contiguous / compact / single-chunk (deflate) data layouts, committed compound
datatypes (JLD2's Julia types), variable-length strings in global heap collections, object
references, and JLD2's ``SerializedDict`` encoding of ``Dict{String,Any}`` (a vector of
``Pair{String,Any}`` whose second field is a reference to the value's dataset).

juliatdgl_amd.checkpoint reads the reference's checkpoints with it (save_state, gtdgl/src/io.jl:
131-150: top-level datasets, deflate-compressed), and tests/golden/convert_goldens.py turned the
reference's golden files into ``.npz`` fixtures with it.  Written from the public HDF5 file format
specification (version 3.0 of the spec: superblock, object header, message and global-heap
layouts).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

SIG = b"\x89HDF\r\n\x1a\n"


class DType:
    def __init__(self, cls, size, **kw):
        self.cls = cls
        self.size = size
        self.__dict__.update(kw)

    def __repr__(self):
        extra = {k: v for k, v in self.__dict__.items() if k not in ("cls", "size")}
        return f"DType(cls={self.cls}, size={self.size}, {extra})"


class JLD2File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        pos = self.buf.find(SIG)
        if pos < 0:
            raise ValueError("no HDF5 signature")
        b = self.buf
        ver = b[pos + 8]
        if ver not in (2, 3):
            raise ValueError(f"superblock version {ver} unsupported")
        self.so, self.sl = b[pos + 9], b[pos + 10]
        assert self.so == 8 and self.sl == 8
        self.base = struct.unpack_from("<Q", b, pos + 12)[0]
        root = struct.unpack_from("<Q", b, pos + 36)[0]
        self.root = self.read_object(root)
        self._heaps = {}

    # -------------------------------------------------------------- low level
    def addr(self, rel):
        return self.base + rel

    def read_object(self, rel):
        """Parse the object header at relative address `rel` -> list of (type, flags, data bytes)."""
        b = self.buf
        p = self.addr(rel)
        if b[p:p + 4] != b"OHDR":
            raise ValueError(f"no OHDR at {rel}")
        flags = b[p + 5]
        q = p + 6
        if flags & 0x20:
            q += 16
        if flags & 0x10:
            q += 4
        csz = 1 << (flags & 3)
        chunk = int.from_bytes(b[q:q + csz], "little")
        q += csz
        msgs = []
        self._parse_msgs(q, q + chunk, flags, msgs)
        return msgs

    def _parse_msgs(self, q, end, hflags, msgs):
        b = self.buf
        while q + 4 <= end:
            mtype = b[q]
            msize = struct.unpack_from("<H", b, q + 1)[0]
            mflags = b[q + 3]
            q += 4
            if hflags & 0x04:
                q += 2
            data = b[q:q + msize]
            q += msize
            if mtype == 0x10:           # continuation
                off, ln = struct.unpack_from("<QQ", data, 0)
                cp = self.addr(off)
                assert b[cp:cp + 4] == b"OCHK"
                self._parse_msgs(cp + 4, cp + ln - 4, hflags, msgs)
            elif mtype != 0:
                msgs.append((mtype, mflags, data))

    # -------------------------------------------------------------- datatypes
    def parse_dtype(self, d, off=0):
        """Datatype message at d[off:] -> (DType, bytes consumed)."""
        cv = d[off]
        cls, ver = cv & 0x0F, cv >> 4
        bits = d[off + 1:off + 4]
        size = struct.unpack_from("<I", d, off + 4)[0]
        p = off + 8
        if cls == 0:    # fixed point
            signed = bool(bits[0] & 0x08)
            p += 4
            return DType(0, size, signed=signed), p - off
        if cls == 1:    # float
            p += 12
            return DType(1, size), p - off
        if cls == 3:    # fixed string
            return DType(3, size), p - off
        if cls == 4:    # bitfield
            p += 4
            return DType(4, size), p - off
        if cls == 5:    # opaque
            taglen = bits[0]
            p += taglen
            return DType(5, size), p - off
        if cls == 6:    # compound
            nm = bits[0] | (bits[1] << 8)
            members = []
            for _ in range(nm):
                e = d.index(b"\0", p)
                name = d[p:e].decode()
                if ver < 3:
                    ln = e - p + 1
                    p += (ln + 7) // 8 * 8
                    moff = struct.unpack_from("<I", d, p)[0]
                    p += 4
                    if ver == 1:
                        p += 28
                else:
                    p = e + 1
                    nb = 1 if size < 256 else 2 if size < 65536 else 4 if size < 2**32 else 8
                    moff = int.from_bytes(d[p:p + nb], "little")
                    p += nb
                mt, used = self.parse_dtype(d, p)
                p += used
                members.append((name, moff, mt))
            return DType(6, size, members=members), p - off
        if cls == 7:    # reference
            return DType(7, size), p - off
        if cls == 9:    # variable length
            vtype = bits[0] & 0x0F
            base, used = self.parse_dtype(d, p)
            p += used
            return DType(9, size, vlen_string=(vtype == 1), base=base), p - off
        if cls == 10:   # array
            nd = d[p]
            p += 1
            if ver < 3:
                p += 3
            dims = struct.unpack_from(f"<{nd}I", d, p)
            p += 4 * nd
            if ver < 3:
                p += 4 * nd
            base, used = self.parse_dtype(d, p)
            p += used
            return DType(10, size, dims=dims, base=base), p - off
        raise ValueError(f"datatype class {cls} unsupported")

    def dtype_of(self, mflags, data):
        if mflags & 0x02:   # shared (committed) datatype
            ver = data[0]
            if ver == 3:
                ref = struct.unpack_from("<Q", data, 2)[0]
            else:
                ref = struct.unpack_from("<Q", data, 8 if ver == 1 else 2)[0]
            for t, f, d in self.read_object(ref):
                if t == 3:
                    return self.dtype_of(f, d)
            raise ValueError("committed datatype without a datatype message")
        return self.parse_dtype(data)[0]

    # -------------------------------------------------------------- dataspace / layout
    @staticmethod
    def parse_space(d):
        ver = d[0]
        nd = d[1]
        if ver == 1:
            p = 8
        else:
            p = 4
            if d[3] == 2:   # null
                return None
        dims = struct.unpack_from(f"<{nd}Q", d, p) if nd else ()
        return tuple(dims)

    def layout_bytes(self, d, nbytes, filters):
        ver = d[0]
        cls = d[1]
        if ver not in (3, 4):
            raise ValueError(f"layout version {ver}")
        if cls == 0:
            sz = struct.unpack_from("<H", d, 2)[0]
            return d[4:4 + sz]
        if cls == 1:
            a, sz = struct.unpack_from("<QQ", d, 2)
            if a == 0xFFFFFFFFFFFFFFFF:
                return b"\0" * nbytes
            return self.buf[self.addr(a):self.addr(a) + sz]
        if cls == 2:
            if ver == 3:
                nd = d[2]
                a = struct.unpack_from("<Q", d, 3)[0]
                return self._btree_v1_chunks(a, nd, nbytes, filters)
            flags = d[2]
            nd = d[3]
            enc = d[4]
            p = 5 + nd * enc
            idx = d[p]
            p += 1
            if idx == 1:    # single chunk
                if flags & 0x02:
                    fsz = int.from_bytes(d[p:p + 8], "little")
                    p += 8 + 4
                    a = struct.unpack_from("<Q", d, p)[0]
                else:
                    a = struct.unpack_from("<Q", d, p)[0]
                    fsz = None
                raw = self.buf[self.addr(a):self.addr(a) + (fsz if fsz else nbytes)]
                return self._unfilter(raw, filters)
            raise ValueError(f"chunk index {idx} unsupported")
        raise ValueError(f"layout class {cls}")

    def _btree_v1_chunks(self, a, nd, nbytes, filters):
        b = self.buf
        p = self.addr(a)
        assert b[p:p + 4] == b"TREE"
        level = b[p + 5]
        assert level == 0
        nent = struct.unpack_from("<H", b, p + 6)[0]
        q = p + 8 + 16
        out = bytearray(nbytes)
        for _ in range(nent):
            csz = struct.unpack_from("<I", b, q)[0]
            q += 8 + 8 * nd
            ca = struct.unpack_from("<Q", b, q)[0]
            q += 8
            raw = self._unfilter(b[self.addr(ca):self.addr(ca) + csz], filters)
            out[:len(raw)] = raw
        q += 8 + 8 * nd
        return bytes(out)

    @staticmethod
    def _unfilter(raw, filters):
        for fid in reversed(filters):
            if fid == 1:
                raw = zlib.decompress(raw)
            else:
                raise ValueError(f"filter {fid} unsupported")
        return raw

    @staticmethod
    def parse_filters(d):
        ver = d[0]
        n = d[1]
        p = 8 if ver == 1 else 2
        ids = []
        for _ in range(n):
            fid = struct.unpack_from("<H", d, p)[0]
            p += 2
            if ver == 1 or fid >= 256:
                nl = struct.unpack_from("<H", d, p)[0]
                p += 2
            else:
                nl = 0
            p += 2   # flags
            nv = struct.unpack_from("<H", d, p)[0]
            p += 2
            if nl:
                p += nl if ver != 1 else (nl + 7) // 8 * 8
            p += 4 * nv
            if ver == 1 and nv % 2:
                p += 4
            ids.append(fid)
        return ids

    # -------------------------------------------------------------- global heap
    def heap_object(self, caddr, index):
        if caddr not in self._heaps:
            b = self.buf
            p = self.addr(caddr)
            assert b[p:p + 4] == b"GCOL"
            csize = struct.unpack_from("<Q", b, p + 8)[0]
            q = p + 16
            objs = {}
            while q + 16 <= p + csize:
                idx, _rc = struct.unpack_from("<HH", b, q)
                osz = struct.unpack_from("<Q", b, q + 8)[0]
                if idx == 0:
                    break
                objs[idx] = b[q + 16:q + 16 + osz]
                q += 16 + (osz + 7) // 8 * 8
            self._heaps[caddr] = objs
        return self._heaps[caddr][index]

    # -------------------------------------------------------------- values
    def decode(self, dt, raw, off=0):
        """One element of type dt at raw[off:]."""
        if dt.cls == 0:
            fmt = {1: "b", 2: "h", 4: "i", 8: "q"}[dt.size]
            if not dt.signed:
                fmt = fmt.upper()
            return struct.unpack_from("<" + fmt, raw, off)[0]
        if dt.cls == 1:
            return struct.unpack_from("<d" if dt.size == 8 else "<f", raw, off)[0]
        if dt.cls == 3:
            return raw[off:off + dt.size].split(b"\0")[0].decode()
        if dt.cls == 7:
            return ("ref", struct.unpack_from("<Q", raw, off)[0])
        if dt.cls == 9:
            ln = struct.unpack_from("<I", raw, off)[0]
            caddr, idx = struct.unpack_from("<QI", raw, off + 4)
            if ln == 0 or caddr == 0:
                data = b""
            else:
                data = self.heap_object(caddr, idx)
            if dt.vlen_string:
                return data[:ln].decode()
            return data
        if dt.cls == 6:
            return {name: self.decode(mt, raw, off + moff) for name, moff, mt in dt.members}
        if dt.cls == 4:
            return int.from_bytes(raw[off:off + dt.size], "little")
        raise ValueError(f"decode class {dt.cls}")

    def read_dataset(self, rel):
        """The value of the dataset at `rel`: numpy arrays for numeric arrays (Julia shape,
        column-major), Python scalars/strings otherwise, dicts for SerializedDicts."""
        msgs = self.read_object(rel)
        space = dt = lay = None
        filters = []
        jtype = None
        for t, f, d in msgs:
            if t == 1:
                space = self.parse_space(d)
            elif t == 3:
                dt = self.dtype_of(f, d)
                if f & 0x02:
                    jtype = self._committed_name(d)
            elif t == 8:
                lay = d
            elif t == 11:
                filters = self.parse_filters(d)
        if dt is None:
            raise ValueError("not a dataset")
        if space is None:
            return None
        n = int(np.prod(space)) if space else 1
        raw = self.layout_bytes(lay, n * dt.size, filters) if lay is not None else b""
        if dt.cls == 1 or (dt.cls == 6 and _is_complex(dt)) or dt.cls == 0:
            if dt.cls == 1:
                arr = np.frombuffer(raw, dtype="<f8" if dt.size == 8 else "<f4", count=n)
            elif dt.cls == 0:
                code = {1: "i1", 2: "i2", 4: "i4", 8: "i8"}[dt.size]
                if not dt.signed:
                    code = "u" + code[1:]
                arr = np.frombuffer(raw, dtype="<" + code, count=n)
            else:
                arr = np.frombuffer(raw, dtype="<c16", count=n)
            if not space:
                v = arr[0]
                if jtype == "Core.Bool" or (dt.cls == 0 and dt.size == 1 and jtype and "Bool" in jtype):
                    return bool(v)
                return v.item()
            shape = tuple(reversed(space))
            return np.array(arr.reshape(shape, order="F"))
        vals = [self.decode(dt, raw, i * dt.size) for i in range(n)]
        if jtype == "JLD2.SerializedDict" or (dt.cls == 6 and [m[0] for m in dt.members] == ["kvvec"]):
            ref = vals[0]["kvvec"]
            pairs = self.read_dataset(ref[1])
            return {p["first"]: self._resolve(p["second"]) for p in pairs}
        if not space:
            return self._resolve(vals[0])
        return [self._resolve(v) for v in vals]

    def _resolve(self, v):
        if isinstance(v, tuple) and v and v[0] == "ref":
            return self.read_dataset(v[1]) if v[1] else None
        if isinstance(v, dict):
            return {k: self._resolve(x) for k, x in v.items()}
        return v

    def _committed_name(self, d):
        ver = d[0]
        ref = struct.unpack_from("<Q", d, 2 if ver == 3 else 8)[0]
        for t, f, dd in self.read_object(ref):
            if t == 0x0C:
                name, val = self._attribute(dd)
                if name == "julia_type":
                    return _type_name(val)
        return None

    def _attribute(self, d):
        ver = d[0]
        if ver in (2, 3):
            nsz, tsz, ssz = struct.unpack_from("<HHH", d, 2)
            p = 9 if ver == 3 else 8
            name = d[p:p + nsz].split(b"\0")[0].decode()
            p += nsz
            tdata = d[p:p + tsz]
            p += tsz
            sdata = d[p:p + ssz]
            p += ssz
        else:
            nsz, tsz, ssz = struct.unpack_from("<HHH", d, 2)
            p = 8
            name = d[p:p + nsz].split(b"\0")[0].decode()
            p += (nsz + 7) // 8 * 8
            tdata = d[p:p + tsz]
            p += (tsz + 7) // 8 * 8
            sdata = d[p:p + ssz]
            p += (ssz + 7) // 8 * 8
        mflags = d[1]
        dt = self.dtype_of(0x02 if mflags & 0x01 else 0, tdata)
        space = self.parse_space(sdata)
        raw = d[p:]
        if space:
            n = int(np.prod(space))
            return name, [self.decode(dt, raw, i * dt.size) for i in range(n)]
        return name, self.decode(dt, raw, 0)

    # -------------------------------------------------------------- top level
    def keys(self):
        return list(self.links().keys())

    def links(self):
        out = {}
        for t, f, d in self.root:
            if t == 6:
                flags = d[1]
                p = 2
                ltype = 0
                if flags & 0x08:
                    ltype = d[p]
                    p += 1
                if flags & 0x04:
                    p += 8
                if flags & 0x10:
                    p += 1
                nb = 1 << (flags & 3)
                ln = int.from_bytes(d[p:p + nb], "little")
                p += nb
                name = d[p:p + ln].decode()
                p += ln
                if ltype == 0:
                    out[name] = struct.unpack_from("<Q", d, p)[0]
        return out

    def __getitem__(self, key):
        return self.read_dataset(self.links()[key])


def _is_complex(dt):
    return (dt.cls == 6 and len(dt.members) == 2 and [m[0] for m in dt.members] == ["re", "im"]
            and all(m[2].cls == 1 and m[2].size == 8 for m in dt.members))


def _type_name(val):
    if isinstance(val, dict):
        n = val.get("name")
        if isinstance(n, str):
            return n
    if isinstance(val, str):
        return val
    return None
