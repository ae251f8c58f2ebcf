"""The HDF5 files REST exchanges guesses and results through, read and written
without an HDF5 library (none is installed in this image).

REST's formats (via the hdf5 crate):
  - guess file: dataset "init_guess" at the root, spin_channel × nbf² FP64 values
    (initial_guess_from_hdf5guess, src/initial_guess/mod.rs:75-87);
  - chkfile: group "scf" with "e_tot" (scalar), "mo_coeff" (spins back to
    back, each the data of the transposed column-major nbf × nstate eigenvector
    matrix, i.e. C[ao][state] row-major — the bytes PySCF writes too) and
    "mo_energy" (spins back to back)
    (save_chkfile, src/post_scf_analysis/mod.rs:47-103; read back by
    initial_guess_from_hdf5chk, src/initial_guess/mod.rs:89-128).

Reader: the file-format subset those files use — superblock version 0, version-1
object headers (with continuation blocks), symbol-table groups (version-1
B-trees, symbol nodes, local heaps), datasets with compact, contiguous or
chunked layout (version-1 chunk B-trees), deflate / shuffle / Fletcher-32
filters, IEEE float and integer element types.  Attributes and
variable-length data are skipped.

Writer: superblock version 0, symbol-table groups, contiguous little-endian
FP64 datasets (dataspace, datatype, fill-value and layout messages), the
structures an HDF5 1.8+ library reads.
"""
from __future__ import annotations

import struct
import zlib
from pathlib import Path
from typing import Dict, Union

import numpy as np

SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(RuntimeError):
    pass


# ---------------------------------------------------------------------------
# reader
# ---------------------------------------------------------------------------
class H5Reader:
    def __init__(self, path: Union[str, Path]):
        self.b = Path(path).read_bytes()
        b = self.b
        if b[:8] != SIG:
            raise H5Error("not an HDF5 file")
        if b[8] != 0:
            raise H5Error(f"superblock version {b[8]} is not supported (version 0 only)")
        self.so, self.sl = b[13], b[14]
        if (self.so, self.sl) != (8, 8):
            raise H5Error("only 8-byte offsets and lengths are supported")
        self.leaf_k, self.node_k = struct.unpack_from("<HH", b, 16)
        self.base = struct.unpack_from("<Q", b, 24)[0]
        root = self._entry(56)
        self.root = root

    # -- primitives ----------------------------------------------------------
    def _q(self, off):
        return struct.unpack_from("<Q", self.b, off)[0]

    def _entry(self, off):
        name_off, hdr, cache = struct.unpack_from("<QQI", self.b, off)
        scratch = self.b[off + 24:off + 40]
        return {"name_off": name_off, "header": hdr, "cache": cache, "scratch": scratch}

    def _messages(self, addr):
        """(type, data bytes) of the version-1 object header at addr, continuation blocks followed."""
        b = self.b
        a = self.base + addr
        if b[a] != 1:
            raise H5Error(f"object header version {b[a]} at {addr} is not supported")
        nmsg = struct.unpack_from("<H", b, a + 2)[0]
        size = struct.unpack_from("<I", b, a + 8)[0]
        blocks = [(a + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            p, n = blocks.pop(0)
            end = p + n
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize, flags = struct.unpack_from("<HHB", b, p)
                data = b[p + 8:p + 8 + msize]
                if mtype == 0x10:
                    coff, clen = struct.unpack_from("<QQ", data, 0)
                    blocks.append((self.base + coff, clen))
                out.append((mtype, data))
                p += 8 + msize
        return out

    def _heap_name(self, heap_addr, off):
        b = self.b
        a = self.base + heap_addr
        if b[a:a + 4] != b"HEAP":
            raise H5Error("bad local heap")
        data = self._q(a + 24)
        s = self.base + data + off
        e = b.index(b"\0", s)
        return b[s:e].decode()

    def _group_members(self, btree, heap) -> Dict[str, int]:
        out = {}
        stack = [btree]
        while stack:
            a = self.base + stack.pop()
            b = self.b
            if b[a:a + 4] != b"TREE" or b[a + 4] != 0:
                raise H5Error("bad group B-tree node")
            level, used = b[a + 5], struct.unpack_from("<H", b, a + 6)[0]
            p = a + 24 + 8   # first key
            for _ in range(used):
                child = self._q(p)
                if level > 0:
                    stack.append(child)
                else:
                    s = self.base + child
                    if b[s:s + 4] != b"SNOD":
                        raise H5Error("bad symbol node")
                    nsym = struct.unpack_from("<H", b, s + 6)[0]
                    for k in range(nsym):
                        e = self._entry(s + 8 + 40 * k)
                        out[self._heap_name(heap, e["name_off"])] = e["header"]
                p += 16
        return out

    def _group_of(self, header_addr):
        for t, d in self._messages(header_addr):
            if t == 0x11:
                bt, hp = struct.unpack_from("<QQ", d, 0)
                return bt, hp
        return None

    def _resolve(self, path: str) -> int:
        addr = self.root["header"]
        for part in [p for p in path.strip("/").split("/") if p]:
            g = self._group_of(addr)
            if g is None:
                raise H5Error(f"{part!r}: parent is not a group")
            members = self._group_members(*g)
            if part not in members:
                raise KeyError(path)
            addr = members[part]
        return addr

    # -- public --------------------------------------------------------------
    def keys(self, path: str = "/") -> list:
        g = self._group_of(self._resolve(path))
        if g is None:
            raise H5Error(f"{path} is not a group")
        return sorted(self._group_members(*g))

    def is_group(self, path: str) -> bool:
        return self._group_of(self._resolve(path)) is not None

    def read(self, path: str) -> np.ndarray:
        msgs = self._messages(self._resolve(path))
        shape, dtype, layout, filters = None, None, None, []
        for t, d in msgs:
            if t == 0x01:
                shape = self._dataspace(d)
            elif t == 0x03:
                dtype = self._datatype(d)
            elif t == 0x08:
                layout = d
            elif t == 0x0B:
                filters = self._filters(d)
        if shape is None or dtype is None or layout is None:
            raise H5Error(f"{path} is not a dataset")
        n = int(np.prod(shape)) if shape else 1
        raw = self._layout_data(layout, shape, dtype, filters, n)
        return np.frombuffer(raw, dtype=dtype, count=n).reshape(shape).astype(dtype.newbyteorder("="))

    # -- messages -------------------------------------------------------------
    def _dataspace(self, d):
        ver, rank, flags = d[0], d[1], d[2]
        off = 8 if ver == 1 else 4
        return tuple(struct.unpack_from("<Q", d, off + 8 * i)[0] for i in range(rank))

    def _datatype(self, d):
        cls, ver = d[0] & 0x0F, d[0] >> 4
        bits = d[1] | (d[2] << 8) | (d[3] << 16)
        size = struct.unpack_from("<I", d, 4)[0]
        order = ">" if bits & 1 else "<"
        if cls == 1:
            return np.dtype(f"{order}f{size}")
        if cls == 0:
            return np.dtype(f"{order}{'i' if bits & 8 else 'u'}{size}")
        raise H5Error(f"datatype class {cls} is not supported")

    def _filters(self, d):
        ver, nf = d[0], d[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(nf):
            fid = struct.unpack_from("<H", d, p)[0]
            if ver == 1 or fid >= 256:
                nlen, flags, ncd = struct.unpack_from("<HHH", d, p + 2)
                p += 8
            else:
                nlen = 0
                flags, ncd = struct.unpack_from("<HH", d, p + 2)
                p += 6
            if ver == 1:
                p += (nlen + 7) // 8 * 8
            else:
                p += nlen
            cd = struct.unpack_from(f"<{ncd}I", d, p) if ncd else ()
            p += 4 * ncd
            if ver == 1 and ncd % 2:
                p += 4
            out.append((fid, cd))
        return out

    def _unfilter(self, raw, filters, mask, esize):
        for k in range(len(filters) - 1, -1, -1):
            fid, cd = filters[k]
            if mask & (1 << k):
                continue
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                a = np.frombuffer(raw, dtype=np.uint8)
                raw = a.reshape(esize, -1).T.tobytes()
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise H5Error(f"filter {fid} is not supported")
        return raw

    def _layout_data(self, d, shape, dtype, filters, n):
        ver = d[0]
        es = dtype.itemsize
        if ver == 3:
            cls = d[1]
            if cls == 0:
                size = struct.unpack_from("<H", d, 2)[0]
                return d[4:4 + size]
            if cls == 1:
                addr, size = struct.unpack_from("<QQ", d, 2)
                if addr == UNDEF:
                    return b"\0" * (n * es)
                return self.b[self.base + addr:self.base + addr + size]
            if cls == 2:
                rank = d[2]
                addr = struct.unpack_from("<Q", d, 3)[0]
                cdims = struct.unpack_from(f"<{rank}I", d, 11)
                return self._chunked(addr, shape, cdims[:len(shape)], es, filters, rank)
        elif ver in (1, 2):
            rank, cls = d[1], d[2]
            p = 8
            addr = None
            if cls != 0:
                addr = struct.unpack_from("<Q", d, p)[0]
                p += 8
            dims = struct.unpack_from(f"<{rank}I", d, p)
            p += 4 * rank
            if cls == 0:
                size = struct.unpack_from("<I", d, p)[0]
                return d[p + 4:p + 4 + size]
            if cls == 1:
                return self.b[self.base + addr:self.base + addr + n * es]
            return self._chunked(addr, shape, dims[:len(shape)], es, filters, rank)
        raise H5Error(f"data layout version {ver} is not supported")

    def _chunked(self, btree, shape, cdims, es, filters, krank):
        out = np.zeros(int(np.prod(shape)) * es, dtype=np.uint8).reshape(tuple(shape) + (es,)) if shape else None
        stack = [btree]
        b = self.b
        ksize = 8 + 8 * krank
        while stack:
            a = self.base + stack.pop()
            if b[a:a + 4] != b"TREE" or b[a + 4] != 1:
                raise H5Error("bad chunk B-tree node")
            level, used = b[a + 5], struct.unpack_from("<H", b, a + 6)[0]
            p = a + 24
            for _ in range(used):
                csize, mask = struct.unpack_from("<II", b, p)
                offs = struct.unpack_from(f"<{krank}Q", b, p + 8)
                child = self._q(p + ksize)
                if level > 0:
                    stack.append(child)
                else:
                    raw = b[self.base + child:self.base + child + csize]
                    raw = self._unfilter(raw, filters, mask, es)
                    chunk = np.frombuffer(raw, dtype=np.uint8).reshape(tuple(cdims) + (es,))
                    sel_out, sel_in = [], []
                    for o, c, s in zip(offs, cdims, shape):
                        m = min(c, s - o)
                        sel_out.append(slice(o, o + m))
                        sel_in.append(slice(0, m))
                    out[tuple(sel_out)] = chunk[tuple(sel_in)]
                p += ksize + 8
        return out.tobytes()


# ---------------------------------------------------------------------------
# writer
# ---------------------------------------------------------------------------
def _pad8(x: bytes) -> bytes:
    return x + b"\0" * (-len(x) % 8)


class H5Writer:
    """Collects a tree {name: ndarray | dict} and writes it in one pass."""

    LEAF_K, NODE_K = 4, 16

    def __init__(self):
        self.tree: dict = {}

    def put(self, path: str, value):
        parts = [p for p in path.strip("/").split("/") if p]
        node = self.tree
        for p in parts[:-1]:
            node = node.setdefault(p, {})
            if not isinstance(node, dict):
                raise H5Error(f"{p} is a dataset")
        node[parts[-1]] = np.ascontiguousarray(value, dtype=np.float64)

    def save(self, path: Union[str, Path]):
        self.out = bytearray(b"\0" * 96)   # superblock + root symbol table entry, filled last
        root = self._group(self.tree)
        sb = bytearray(SIG)
        sb += bytes([0, 0, 0, 0, 0, 8, 8, 0])
        sb += struct.pack("<HHI", self.LEAF_K, self.NODE_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.out), UNDEF)
        sb += struct.pack("<QQII", 0, root["header"], 1, 0) + struct.pack("<QQ", root["btree"], root["heap"])
        self.out[:96] = sb
        Path(path).write_bytes(bytes(self.out))

    def _alloc(self, data: bytes) -> int:
        addr = len(self.out)
        self.out += _pad8(data)
        return addr

    def _object_header(self, messages) -> int:
        body = b""
        for mtype, data in messages:
            data = _pad8(data)
            body += struct.pack("<HHB3x", mtype, len(data), 0) + data
        hdr = struct.pack("<BBHII", 1, 0, len(messages), 1, len(body)) + b"\0" * 4
        return self._alloc(hdr + body)

    def _dataset(self, a: np.ndarray) -> int:
        data_addr = self._alloc(a.astype("<f8").tobytes()) if a.size else UNDEF
        rank = a.ndim
        space = struct.pack("<BBB5x", 1, rank, 0) + b"".join(struct.pack("<Q", d) for d in a.shape)
        dtype = bytes([0x11, 0x20, 63, 0]) + struct.pack("<I", 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
        fill = bytes([2, 2, 2, 0])
        layout = struct.pack("<BBQQ", 3, 1, data_addr, a.size * 8)
        return self._object_header([(0x01, space), (0x03, dtype), (0x05, fill), (0x08, layout)])

    def _group(self, node: dict) -> dict:
        names = sorted(node)
        if len(names) > 2 * self.LEAF_K * 2 * self.NODE_K:
            raise H5Error("too many members in one group")
        headers = {}
        for n in names:
            v = node[n]
            headers[n] = self._group(v)["header"] if isinstance(v, dict) else self._dataset(v)
        # local heap: offset 0 holds the empty string
        heap_data = bytearray(b"\0" * 8)
        name_off = {}
        for n in names:
            name_off[n] = len(heap_data)
            heap_data += _pad8(n.encode() + b"\0")
        heap_data += b"\0" * 8
        data_addr = self._alloc(bytes(heap_data))
        heap = self._alloc(b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(heap_data), UNDEF, data_addr))
        # symbol nodes of at most 2·LEAF_K entries, full-size
        cap = 2 * self.LEAF_K
        chunks = [names[i:i + cap] for i in range(0, len(names), cap)] or [[]]
        snods = []
        for ch in chunks:
            body = b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(ch))
            for n in ch:
                body += struct.pack("<QQII", name_off[n], headers[n], 0, 0) + b"\0" * 16
            body += b"\0" * (40 * (cap - len(ch)))
            snods.append(self._alloc(body))
        # one leaf B-tree node (type 0), allocated at full capacity
        used = len(snods)
        node_b = b"TREE" + bytes([0, 0]) + struct.pack("<HQQ", used, UNDEF, UNDEF)
        node_b += struct.pack("<Q", 0)
        for ch, sn in zip(chunks, snods):
            node_b += struct.pack("<QQ", sn, name_off[ch[-1]] if ch else 0)
        node_b += b"\0" * (16 * (2 * self.NODE_K - used))
        btree = self._alloc(node_b)
        header = self._object_header([(0x11, struct.pack("<QQ", btree, heap))])
        return {"header": header, "btree": btree, "heap": heap}


# ---------------------------------------------------------------------------
# REST's two files
# ---------------------------------------------------------------------------
def read_init_guess(path, nbf: int, spin_channel: int = 1) -> list:
    """Density matrices of REST's guess file (initial_guess_from_hdf5guess): spin_channel nbf×nbf
    matrices taken in order from the flat "init_guess" dataset (REST reads them into its
    column-major MatrixFull; every density is symmetric)."""
    raw = H5Reader(path).read("init_guess").ravel()
    if raw.size < spin_channel * nbf * nbf:
        raise H5Error(f"init_guess holds {raw.size} values, {spin_channel}×{nbf}² needed")
    return [raw[s * nbf * nbf:(s + 1) * nbf * nbf].reshape(nbf, nbf).T.copy() for s in range(spin_channel)]


def write_init_guess(path, dms: list):
    w = H5Writer()
    w.put("init_guess", np.concatenate([np.asarray(d, dtype=np.float64).T.ravel() for d in dms]))
    w.save(path)


def read_chkfile(path, nbf: int, nstate: int, spin_channel: int = 1):
    """(e_tot, [C_s nbf×nstate], [ε_s]) of REST's chkfile group "scf" (initial_guess_from_hdf5chk)."""
    r = H5Reader(path)
    e = float(np.asarray(r.read("scf/e_tot")).ravel()[0])
    c = r.read("scf/mo_coeff").ravel()
    eps = r.read("scf/mo_energy").ravel()
    cs, es = [], []
    for s in range(spin_channel):
        # REST reads the block as a column-major [num_state, num_basis] matrix and transposes it
        # (src/initial_guess/mod.rs:105-108): the bytes are C[ao][state] row-major, PySCF's layout too
        cs.append(c[s * nstate * nbf:(s + 1) * nstate * nbf].reshape(nbf, nstate).copy())
        es.append(eps[s * nstate:(s + 1) * nstate].copy())
    return e, cs, es


def write_chkfile(path, e_tot: float, mo_coeff: list, mo_energy: list):
    """REST's save_chkfile layout (src/post_scf_analysis/mod.rs:77-81): scf/e_tot, scf/mo_coeff (each
    spin's C[ao][state] row-major: the data of the transposed column-major matrix), scf/mo_energy."""
    w = H5Writer()
    w.put("scf/e_tot", np.float64(e_tot))
    w.put("scf/mo_coeff", np.concatenate([np.asarray(c, dtype=np.float64).ravel() for c in mo_coeff]))
    w.put("scf/mo_energy", np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in mo_energy]))
    w.save(path)
