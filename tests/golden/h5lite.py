"""A small, independent HDF5 writer and reader for the tests of the native
file contract (no HDF5 library is installed).

write_input(path, variables, attrs, compression) writes a file the way the
reference's MATLAB side does (+kwave/+legacy/writeMatrix.m through MATLAB's
HDF5 library): superblock version 0, the root group as a symbol table,
every matrix CHUNKED with writeMatrix.m's chunk shapes ([Nx Ny 1] for 3-D
matrices, [Nx 1 1] for 2-D ones, the 1-D rules for vectors), indexed by a
version-1 B-tree, deflate-compressed when compression > 0 (chunks of
vectors of at most 1 MB are never compressed, as writeMatrix.m decides),
with the domain_type / data_type attributes and the root attributes.

read(path) returns {name: array} for every dataset of a file written by
the HDF5 library's default layout (version-0 superblock, symbol-table
groups, version-1 object headers, contiguous or chunked storage), with
arrays in MATLAB order (the reverse of the HDF5 dimensions, column-major),
groups as nested dicts, and read_attrs(path, name) the string attributes.

Structures follow the HDF5 File Format Specification; this module shares no
code with the library it tests (k-wave-ii_amd/csrc/kw_h5.cpp).
"""

import struct
import zlib

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIG = b"\x89HDF\r\n\x1a\n"


def _p8(b):
    return b + b"\0" * ((-len(b)) % 8)


# ---------------------------------------------------------------------------
# writer (MATLAB style: chunked datasets)
# ---------------------------------------------------------------------------

def _dtype_msg(kind, size):
    if kind == "f":
        if size == 4:
            return struct.pack("<BBBBI HHBBBBI", 0x11, 0x20, 31, 0, 4, 0, 32, 23, 8, 0, 23, 127)
        return struct.pack("<BBBBI HHBBBBI", 0x11, 0x20, 63, 0, 8, 0, 64, 52, 11, 0, 52, 1023)
    if kind == "u":
        return struct.pack("<BBBBI HH", 0x10, 0, 0, 0, size, 0, size * 8)
    if kind == "s":
        return struct.pack("<BBBBI", 0x13, 0x00, 0, 0, size)
    raise ValueError(kind)


def _space_msg(dims):
    return struct.pack("<BBBB I", 1, len(dims), 0, 0, 0) + b"".join(struct.pack("<Q", d) for d in dims)


def _msg(mtype, body):
    body = _p8(body)
    return struct.pack("<HHB3x", mtype, len(body), 0) + body


def _attr_msg(name, value):
    v = value.encode() + b"\0"
    nm = name.encode() + b"\0"
    dt = _dtype_msg("s", len(v))
    ds = struct.pack("<BBBB I", 1, 0, 0, 0, 0)
    body = struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(ds)) + _p8(nm) + _p8(dt) + _p8(ds) + v
    return _msg(0x000C, body)


def _matlab_chunk(shape, compression):
    """writeMatrix.m: chunk size in MATLAB order and the compression level."""
    Nx, Ny, Nz = (list(shape) + [1, 1, 1])[:3]
    ndim = sum(1 for n in (Nx, Ny, Nz) if n > 1)
    if ndim == 3:
        return [Nx, Ny, 1], compression
    if ndim == 2:
        return [Nx, 1, 1], compression
    one_mb = 1024 ** 2 // 8
    n = Nx * Ny * Nz
    if n > one_mb:
        if Nx > Ny:
            return [one_mb, 1, 1], compression
        if Ny > Nz:
            return [1, one_mb, 1], compression
        return [1, 1, one_mb], compression
    if n == 1:
        return [1, 1, 1], 0
    if Nx > Ny:
        return [Nx, 1, 1], 0
    if Ny > Nz:
        return [1, Ny, 1], 0
    return [1, 1, Nz], 0


class _File:
    def __init__(self):
        self.buf = bytearray()

    def alloc(self, n):
        a = len(self.buf)
        self.buf += b"\0" * n
        return a

    def put(self, addr, data):
        self.buf[addr:addr + len(data)] = data

    def append(self, data):
        a = len(self.buf)
        self.buf += data
        self.buf += b"\0" * ((-len(self.buf)) % 8)
        return a


def write_input(path, variables, attrs=None, compression=0):
    """variables: {name: (array in MATLAB order, 'single' | 'uint64')}; every
    array is stored with MATLAB's size (at least 3 dimensions, [N 1 1] for a
    column), chunked as writeMatrix.m does."""
    f = _File()
    sb = f.alloc(96)
    names = sorted(variables)
    objs = {}
    chunk_k = 32  # the library's default for chunk B-trees
    for name in names:
        arr, cls = variables[name]
        a = np.asarray(arr, dtype=np.float32 if cls == "single" else np.uint64)
        mshape = list(a.shape) + [1] * (3 - a.ndim)
        if a.ndim > 3:
            raise ValueError("at most 3 dimensions")
        hd = mshape[::-1]  # HDF5 dims, slowest first
        data = np.asfortranarray(a.reshape(mshape, order="F"))
        flat = data.ravel(order="F")  # x fastest == C order of hd
        cm, level = _matlab_chunk(mshape, compression)
        ch = cm[::-1]
        es = a.dtype.itemsize
        # chunks in C order of their grid coordinates
        grid = [(hd[i] + ch[i] - 1) // ch[i] for i in range(3)]
        cube = flat.reshape(hd)
        chunks = []
        for c0 in range(grid[0]):
            for c1 in range(grid[1]):
                for c2 in range(grid[2]):
                    off = [c0 * ch[0], c1 * ch[1], c2 * ch[2]]
                    block = np.zeros(ch, dtype=a.dtype)
                    part = cube[off[0]:off[0] + ch[0], off[1]:off[1] + ch[1], off[2]:off[2] + ch[2]]
                    block[:part.shape[0], :part.shape[1], :part.shape[2]] = part
                    raw = block.tobytes(order="C")
                    if level > 0:
                        raw = zlib.compress(raw, level)
                    chunks.append((off, raw))
        if len(chunks) > 2 * chunk_k:
            raise ValueError("more chunks than one B-tree leaf holds")
        caddr = [f.append(raw) for _, raw in chunks]
        # B-tree (type 1) leaf: key i = (chunk bytes, filter mask, offsets + 0)
        ksz = 8 + 8 * 4
        node_size = 24 + 2 * chunk_k * 8 + (2 * chunk_k + 1) * ksz
        bt = f.alloc(node_size)
        body = b"TREE" + struct.pack("<BBH", 1, 0, len(chunks)) + struct.pack("<QQ", UNDEF, UNDEF)
        for (off, raw), ad in zip(chunks, caddr):
            body += struct.pack("<II", len(raw), 0) + b"".join(struct.pack("<Q", o) for o in off + [0])
            body += struct.pack("<Q", ad)
        body += struct.pack("<II", 0, 0) + b"".join(struct.pack("<Q", d) for d in hd + [0])
        f.put(bt, body)
        msgs = [
            _msg(0x0001, _space_msg(hd)),
            _msg(0x0003, _dtype_msg("f" if cls == "single" else "u", es)),
            _msg(0x0005, struct.pack("<BBBB", 2, 2, 2, 0)),
            _msg(0x0008, struct.pack("<BBBQ", 3, 2, 4, bt) + b"".join(struct.pack("<I", c) for c in ch + [es])),
        ]
        if level > 0:
            # deflate filter, version-1 pipeline: id 1, name "deflate", one value
            nm = _p8(b"deflate\0")
            flt = struct.pack("<HHHH", 1, len(b"deflate\0"), 1, 1) + nm + struct.pack("<I", level) + b"\0" * 4
            msgs.append(_msg(0x000B, struct.pack("<BB6x", 1, 1) + flt))
        msgs.append(_attr_msg("domain_type", "real"))
        msgs.append(_attr_msg("data_type", "float" if cls == "single" else "long"))
        objs[name] = _ohdr(f, msgs)
    root_msgs_attrs = [_attr_msg(k, v) for k, v in (attrs or {}).items()]
    root = _group(f, names, objs, root_msgs_attrs, leaf_k=max(4, (len(names) + 1) // 2))
    eof = len(f.buf)
    sbb = SIG + struct.pack("<BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0)
    sbb += struct.pack("<HHI", max(4, (len(names) + 1) // 2), 16, 0)
    sbb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sbb += struct.pack("<QQII QQ", 0, root[0], 1, 0, root[1], root[2])
    f.put(sb, sbb)
    with open(path, "wb") as fp:
        fp.write(bytes(f.buf))


def _ohdr(f, msgs):
    body = b"".join(msgs)
    hdr = struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\0" * 4
    return f.append(hdr + body)


def _group(f, names, objs, extra_msgs, leaf_k):
    # local heap with the names, a one-leaf B-tree, a symbol table node
    heap_data = bytearray(b"\0" * 8)
    offs = {}
    for n in names:
        offs[n] = len(heap_data)
        heap_data += _p8(n.encode() + b"\0")
    hd_addr = f.append(bytes(heap_data))
    heap = f.append(b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), 1, hd_addr))
    snod = f.alloc(8 + 2 * leaf_k * 40)
    ent = b"".join(struct.pack("<QQII16x", offs[n], objs[n], 0, 0) for n in names)
    f.put(snod, b"SNOD" + struct.pack("<BBH", 1, 0, len(names)) + ent)
    k = 16
    bt = f.alloc(24 + 2 * k * 8 + (2 * k + 1) * 8)
    body = b"TREE" + struct.pack("<BBH", 0, 0, 1 if names else 0) + struct.pack("<QQ", UNDEF, UNDEF)
    if names:
        body += struct.pack("<QQQ", 0, snod, offs[names[-1]])
    f.put(bt, body)
    oh = _ohdr(f, [_msg(0x0011, struct.pack("<QQ", bt, heap))] + extra_msgs)
    return oh, bt, heap


# ---------------------------------------------------------------------------
# reader
# ---------------------------------------------------------------------------

class _R:
    def __init__(self, path):
        with open(path, "rb") as fp:
            self.b = fp.read()
        if self.b[:8] != SIG:
            raise ValueError("not an HDF5 file")
        ver = self.b[8]
        if ver not in (0, 1):
            raise ValueError("superblock version %d" % ver)
        p = 24 + (4 if ver == 1 else 0)
        self.root = struct.unpack_from("<Q", self.b, p + 32 + 8)[0]

    def q(self, o):
        return struct.unpack_from("<Q", self.b, o)[0]

    def messages(self, oh):
        ver, _, n, _, size = struct.unpack_from("<BBHII", self.b, oh)
        if ver != 1:
            raise ValueError("object header version %d" % ver)
        out, blocks = [], [(oh + 16, size)]
        while blocks and len(out) < n:
            p, size = blocks.pop(0)
            end = p + size
            while p + 8 <= end and len(out) < n:
                t, sz = struct.unpack_from("<HH", self.b, p)
                data = self.b[p + 8:p + 8 + sz]
                if t == 0x0010:
                    blocks.append(struct.unpack_from("<QQ", data))
                out.append((t, data))
                p += 8 + sz
        return out

    def heap_name(self, heap, off):
        da = struct.unpack_from("<Q", self.b, heap + 24)[0]
        s = self.b[da + off:]
        return s[:s.index(b"\0")].decode()

    def group(self, oh):
        st = [d for t, d in self.messages(oh) if t == 0x0011]
        if not st:
            return None
        bt, heap = struct.unpack_from("<QQ", st[0])
        out = {}
        self._btree(bt, heap, out)
        return out

    def _btree(self, node, heap, out):
        assert self.b[node:node + 4] == b"TREE" and self.b[node + 4] == 0
        level, n = self.b[node + 5], struct.unpack_from("<H", self.b, node + 6)[0]
        p = node + 24
        for _ in range(n):
            child = self.q(p + 8)
            p += 16
            if level:
                self._btree(child, heap, out)
                continue
            assert self.b[child:child + 4] == b"SNOD"
            ns = struct.unpack_from("<H", self.b, child + 6)[0]
            for k in range(ns):
                e = child + 8 + 40 * k
                out[self.heap_name(heap, self.q(e))] = self.q(e + 8)

    def dataset(self, oh):
        ms = dict((t, d) for t, d in self.messages(oh) if t in (1, 3, 8, 0x0B))
        sp, dt, ly = ms[1], ms[3], ms[8]
        rank = sp[1]
        off = 8 if sp[0] == 1 else 4
        dims = [struct.unpack_from("<Q", sp, off + 8 * i)[0] for i in range(rank)]
        cls, size = dt[0] & 0x0F, struct.unpack_from("<I", dt, 4)[0]
        dtype = {(1, 4): "<f4", (1, 8): "<f8", (0, 8): "<u8" if not dt[1] & 8 else "<i8",
                 (0, 4): "<u4" if not dt[1] & 8 else "<i4"}[(cls, size)]
        n = int(np.prod(dims)) if dims else 1
        assert ly[0] == 3
        if ly[1] == 1:
            addr, nbytes = struct.unpack_from("<QQ", ly, 2)
            raw = self.b[addr:addr + nbytes] if addr != UNDEF else b"\0" * (n * size)
            a = np.frombuffer(raw, dtype=dtype, count=n)
        elif ly[1] == 2:
            rk1 = ly[2]
            bt = struct.unpack_from("<Q", ly, 3)[0]
            cd = [struct.unpack_from("<I", ly, 11 + 4 * i)[0] for i in range(rk1)][:-1]
            filt = 0x0B in ms
            a = np.zeros(dims, dtype=dtype)
            self._chunks(bt, dims, cd, dtype, filt, a)
            a = a.ravel()
        else:
            raise ValueError("layout class %d" % ly[1])
        return a.reshape(dims[::-1] if dims else (), order="F")

    def _chunks(self, node, dims, cd, dtype, filt, out):
        level, n = self.b[node + 5], struct.unpack_from("<H", self.b, node + 6)[0]
        rk1 = len(cd) + 1
        ksz = 8 + 8 * rk1
        p = node + 24
        for _ in range(n):
            csz = struct.unpack_from("<I", self.b, p)[0]
            offs = [self.q(p + 8 + 8 * i) for i in range(len(cd))]
            child = self.q(p + ksz)
            p += ksz + 8
            if level:
                self._chunks(child, dims, cd, dtype, filt, out)
                continue
            raw = self.b[child:child + csz]
            if filt:
                raw = zlib.decompress(raw)
            block = np.frombuffer(raw, dtype=dtype).reshape(cd)
            sl = tuple(slice(o, min(o + c, d)) for o, c, d in zip(offs, cd, dims))
            out[sl] = block[tuple(slice(0, s.stop - s.start) for s in sl)]

    def attrs(self, oh):
        out = {}
        for t, d in self.messages(oh):
            if t != 0x000C:
                continue
            ver, _, nl, tl, sl = struct.unpack_from("<BBHHH", d)
            pad = (lambda x: x + (-x) % 8) if ver == 1 else (lambda x: x)
            o = 8
            name = d[o:o + nl].split(b"\0")[0].decode()
            o += pad(nl)
            size = struct.unpack_from("<I", d, o + 4)[0]
            o += pad(tl) + pad(sl)
            out[name] = d[o:o + size].split(b"\0")[0].decode()
        return out


def read(path):
    r = _R(path)

    def walk(oh):
        g = r.group(oh)
        if g is None:
            return r.dataset(oh)
        return {k: walk(v) for k, v in g.items()}

    return walk(r.root)


def read_attrs(path, name=""):
    r = _R(path)
    oh = r.root
    for part in [p for p in name.split("/") if p]:
        oh = r.group(oh)[part]
    return r.attrs(oh)
