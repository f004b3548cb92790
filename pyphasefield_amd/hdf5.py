"""Numeric datasets of HDF5 files, read in pure Python (numpy + zlib).

XDMF files keep their heavy data in HDF5 by default: meshio's XDMF writers,
which the reference's ``CSVtoXDMF`` uses (ppf_utils.py:63-110, its
``TimeSeriesWriter`` default ``data_format="HDF"``), store every array as a
dataset of a ``.h5`` file beside the ``.xdmf`` and name it there as
``file.h5:/dataK``.  No HDF5 library is available to this package, so this
module reads the parts of the format such files use, following the HDF5
File Format Specification (version 3.0, The HDF Group):

* superblock versions 0-3; version-1 and version-2 object headers with their
  continuation blocks;
* groups: symbol tables (version-1 B-tree, symbol-table nodes, local heap),
  link messages in the object header, and dense link storage (fractal heap
  with a root direct or indirect block, version-2 B-tree name index of depth
  0 or 1); hard links only;
* datasets of fixed-point and IEEE floating-point types of either byte
  order; dataspace versions 1-2; fill-value messages; data layouts compact,
  contiguous and chunked -- version-1 B-tree chunk index (layout versions
  1-3), single-chunk, implicit and unpaged fixed-array indexes (layout
  version 4); filters deflate, shuffle and fletcher32 (the checksum is
  dropped, not verified).

Anything else (extensible-array and version-2 B-tree chunk indexes, shared
messages, variable-length or compound types, other filters, external links)
raises ``HDF5Error`` naming what was found.  Checksums of version-2
structures are not verified.
"""

import os
import zlib

import numpy as np

__all__ = ["HDF5Error", "File", "read_dataset", "write_datasets"]


class HDF5Error(ValueError):
    pass


_SIGNATURE = b"\x89HDF\r\n\x1a\n"

# object header message types
_DATASPACE, _LINK_INFO, _DATATYPE, _FILL_OLD, _FILL = 0x01, 0x02, 0x03, 0x04, 0x05
_LINK, _LAYOUT, _FILTERS, _CONTINUATION, _SYMBOL_TABLE = 0x06, 0x08, 0x0B, 0x10, 0x11

_FILTER_DEFLATE, _FILTER_SHUFFLE, _FILTER_FLETCHER32 = 1, 2, 3


def _bytes_for(value):
    """Bytes needed to encode a non-negative integer."""
    return max(1, (int(value).bit_length() + 7) // 8)


class _Cursor:
    """Little-endian reader over a bytes object."""

    def __init__(self, data, pos=0, so=8, sl=8):
        self.data, self.pos, self.so, self.sl = data, pos, so, sl

    def u(self, n):
        if self.pos + n > len(self.data):
            raise HDF5Error("truncated structure")
        v = int.from_bytes(self.data[self.pos:self.pos + n], "little")
        self.pos += n
        return v

    def raw(self, n):
        if self.pos + n > len(self.data):
            raise HDF5Error("truncated structure")
        b = self.data[self.pos:self.pos + n]
        self.pos += n
        return b

    def offset(self):
        return self.u(self.so)

    def length(self):
        return self.u(self.sl)

    def skip(self, n):
        self.pos += n


class _Dataset:
    __slots__ = ("shape", "dtype", "layout", "filters", "fill")


class File:
    """An HDF5 file opened for reading datasets by path."""

    def __init__(self, path):
        self.path = path
        self._f = open(path, "rb")
        try:
            self._size = os.fstat(self._f.fileno()).st_size
            self._superblock()
        except Exception:
            self._f.close()
            raise

    def close(self):
        self._f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ raw access --
    def _read(self, addr, n):
        self._f.seek(self._base + addr)
        b = self._f.read(n)
        if len(b) != n:
            raise HDF5Error("%s: read of %d bytes at %d runs past the end of the file" % (self.path, n, addr))
        return b

    def _cursor(self, addr, n):
        return _Cursor(self._read(addr, n), 0, self._so, self._sl)

    def _undefined(self, addr):
        return addr == (1 << (8 * self._so)) - 1

    # ------------------------------------------------------------ superblock --
    def _superblock(self):
        at = 0
        while True:
            self._f.seek(at)
            if self._f.read(8) == _SIGNATURE:
                break
            at = 512 if at == 0 else 2 * at
            if at >= self._size:
                raise HDF5Error("%s is not an HDF5 file (no superblock signature)" % self.path)
        self._f.seek(at + 8)
        head = self._f.read(200)
        version = head[0]
        if version in (0, 1):
            self._so, self._sl = head[5], head[6]
            c = _Cursor(head, 8 + 4 + 4 + (4 if version == 1 else 0), self._so, self._sl)
            self._base = c.offset()
            c.offset()  # free-space info
            c.offset()  # end of file
            c.offset()  # driver info
            c.offset()  # root entry: link name offset
            self._root = c.offset()
        elif version in (2, 3):
            self._so, self._sl = head[1], head[2]
            c = _Cursor(head, 4, self._so, self._sl)
            self._base = c.offset()
            c.offset()  # superblock extension
            c.offset()  # end of file
            self._root = c.offset()
        else:
            raise HDF5Error("%s: superblock version %d is not supported" % (self.path, version))

    # --------------------------------------------------------- object header --
    def _messages(self, addr):
        """[(type, flags, data)] of the object header at addr."""
        first = self._read(addr, 4)
        if first == b"OHDR":
            return self._messages_v2(addr)
        if first[0] != 1:
            raise HDF5Error("object header version %d at %d is not supported" % (first[0], addr))
        c = self._cursor(addr, 16)
        c.skip(2)
        nmsgs = c.u(2)
        c.skip(4)
        size = c.u(4)
        out = []
        blocks = [(addr + 16, size)]
        while blocks and len(out) < nmsgs:
            baddr, bsize = blocks.pop(0)
            c = self._cursor(baddr, bsize)
            while c.pos + 8 <= bsize and len(out) < nmsgs:
                mtype, msize, mflags = c.u(2), c.u(2), c.u(1)
                c.skip(3)
                data = c.raw(msize)
                if mtype == _CONTINUATION:
                    d = _Cursor(data, 0, self._so, self._sl)
                    blocks.append((d.offset(), d.length()))
                out.append((mtype, mflags, data))
        return out

    def _messages_v2(self, addr):
        c = self._cursor(addr, 6 + 16 + 4 + 8)
        c.skip(4)
        if c.u(1) != 2:
            raise HDF5Error("object header at %d: unknown OHDR version" % addr)
        flags = c.u(1)
        if flags & 0x20:
            c.skip(16)
        if flags & 0x10:
            c.skip(4)
        size = c.u(1 << (flags & 3))
        order = bool(flags & 0x04)
        out = []
        blocks = [(addr + c.pos, size, False)]
        while blocks:
            baddr, bsize, is_cont = blocks.pop(0)
            if is_cont:
                if self._read(baddr, 4) != b"OCHK":
                    raise HDF5Error("object header continuation at %d: bad signature" % baddr)
                baddr, bsize = baddr + 4, bsize - 8
            c = self._cursor(baddr, bsize)
            hdr = 4 + (2 if order else 0)
            while c.pos + hdr <= bsize:
                mtype, msize, mflags = c.u(1), c.u(2), c.u(1)
                if order:
                    c.skip(2)
                data = c.raw(msize)
                if mtype == _CONTINUATION:
                    d = _Cursor(data, 0, self._so, self._sl)
                    blocks.append((d.offset(), d.length(), True))
                out.append((mtype, mflags, data))
        return out

    # ---------------------------------------------------------------- groups --
    def _links(self, addr):
        """{name: object header address} of the group at addr."""
        links = {}
        for mtype, _, data in self._messages(addr):
            if mtype == _SYMBOL_TABLE:
                c = _Cursor(data, 0, self._so, self._sl)
                links.update(self._symbol_table(c.offset(), c.offset()))
            elif mtype == _LINK:
                name, target = self._link_message(data)
                links[name] = target
            elif mtype == _LINK_INFO:
                c = _Cursor(data, 0, self._so, self._sl)
                c.skip(1)
                flags = c.u(1)
                if flags & 0x01:
                    c.skip(8)
                heap, names = c.offset(), c.offset()
                if not self._undefined(heap):
                    links.update(self._dense_links(heap, names))
        return links

    def _link_message(self, data):
        c = _Cursor(data, 0, self._so, self._sl)
        if c.u(1) != 1:
            raise HDF5Error("unknown link message version")
        flags = c.u(1)
        kind = c.u(1) if flags & 0x08 else 0
        if flags & 0x04:
            c.skip(8)
        if flags & 0x10:
            c.skip(1)
        name = c.raw(c.u(1 << (flags & 3))).decode("utf-8")
        if kind != 0:
            raise HDF5Error("link %r is a %s link; only hard links are supported"
                            % (name, {1: "soft", 64: "external"}.get(kind, "type-%d" % kind)))
        return name, c.offset()

    def _symbol_table(self, btree, heap):
        h = self._cursor(heap, 8 + 2 * self._sl + self._so)
        if h.raw(4) != b"HEAP":
            raise HDF5Error("local heap at %d: bad signature" % heap)
        h.skip(4)
        seg_size = h.length()
        h.length()
        names = self._read(h.offset(), seg_size)
        links = {}
        for snod in self._btree1(btree, 0, None):
            c = self._cursor(snod, 8)
            if c.raw(4) != b"SNOD":
                raise HDF5Error("symbol table node at %d: bad signature" % snod)
            c.skip(2)
            n = c.u(2)
            entry = 2 * self._so + 24
            e = self._cursor(snod + 8, n * entry)
            for _ in range(n):
                name_off, target = e.offset(), e.offset()
                e.skip(24)
                end = names.index(b"\0", name_off)
                links[names[name_off:end].decode("utf-8")] = target
        return links

    def _btree1(self, addr, node_type, ndims):
        """Children of a version-1 B-tree: symbol-table node addresses (type
        0) or (chunk address, size, filter mask, offsets) (type 1)."""
        key_size = self._sl if node_type == 0 else 8 + 8 * ndims
        c = self._cursor(addr, 8 + 2 * self._so)
        if c.raw(4) != b"TREE":
            raise HDF5Error("B-tree node at %d: bad signature" % addr)
        if c.u(1) != node_type:
            raise HDF5Error("B-tree node at %d: unexpected node type" % addr)
        level, n = c.u(1), c.u(2)
        body = self._cursor(addr + 8 + 2 * self._so, (n + 1) * key_size + n * self._so)
        out = []
        for _ in range(n):
            if node_type == 1:
                size, mask = body.u(4), body.u(4)
                offsets = [body.u(8) for _ in range(ndims)]
            else:
                body.skip(key_size)
            child = body.offset()
            if level > 0:
                out.extend(self._btree1(child, node_type, ndims))
            elif node_type == 1:
                out.append((child, size, mask, offsets))
            else:
                out.append(child)
        return out

    # ---------------------------------------------------- dense link storage --
    def _dense_links(self, heap_addr, btree_addr):
        heap = self._fractal_heap(heap_addr)
        links = {}
        for rec in self._btree2_records(btree_addr, 5):
            heap_id = rec[4:]
            name, target = self._link_message(heap["object"](heap_id))
            links[name] = target
        return links

    def _fractal_heap(self, addr):
        c = self._cursor(addr, 256)
        if c.raw(4) != b"FRHP":
            raise HDF5Error("fractal heap at %d: bad signature" % addr)
        c.skip(1)
        id_len, filter_len, flags = c.u(2), c.u(2), c.u(1)
        max_managed = c.u(4)
        c.length()  # next huge object ID
        c.offset()  # huge-object B-tree
        c.length()  # free space in managed blocks
        c.offset()  # free-space manager
        for _ in range(8):
            c.length()  # managed space, allocated, iterator, object counts and sizes
        width = c.u(2)
        start_block, max_dblock = c.length(), c.length()
        max_heap_bits = c.u(2)
        c.u(2)  # starting rows of the root indirect block
        root = c.offset()
        rows = c.u(2)
        if filter_len:
            raise HDF5Error("fractal heap at %d has I/O filters, which are not supported" % addr)
        off_bytes = (max_heap_bits + 7) // 8
        len_bytes = _bytes_for(min(max_dblock, max_managed))
        checksummed = bool(flags & 0x02)
        dblock_header = 5 + self._so + off_bytes + (4 if checksummed else 0)
        blocks = []  # (heap offset, address, size)
        if rows == 0:
            blocks.append((0, root, start_block))
        else:
            max_drows = (max_dblock.bit_length() - 1) - (start_block.bit_length() - 1) + 2
            if rows > max_drows:
                raise HDF5Error("fractal heap at %d: indirect blocks below the root are not supported" % addr)
            ib = self._cursor(root, 5 + self._so + off_bytes + rows * width * self._so + 4)
            if ib.raw(4) != b"FHIB":
                raise HDF5Error("fractal heap indirect block at %d: bad signature" % root)
            ib.skip(1 + self._so + off_bytes)
            heap_off = 0
            for r in range(rows):
                size = start_block if r < 2 else start_block << (r - 1)
                for _ in range(width):
                    child = ib.offset()
                    if not self._undefined(child):
                        blocks.append((heap_off, child, size))
                    heap_off += size
        cache = {}

        def obj(heap_id):
            kind = (heap_id[0] >> 4) & 0x3
            if kind != 0:
                raise HDF5Error("fractal heap object of type %d (huge or tiny) is not supported" % kind)
            if 1 + off_bytes + len_bytes > id_len:
                raise HDF5Error("fractal heap ID layout not understood")
            off = int.from_bytes(heap_id[1:1 + off_bytes], "little")
            n = int.from_bytes(heap_id[1 + off_bytes:1 + off_bytes + len_bytes], "little")
            for boff, baddr, bsize in blocks:
                if boff <= off < boff + bsize:
                    if baddr not in cache:
                        data = self._read(baddr, bsize)
                        if data[:4] != b"FHDB":
                            raise HDF5Error("fractal heap direct block at %d: bad signature" % baddr)
                        cache[baddr] = data
                    pos = off - boff
                    if pos < dblock_header or pos + n > bsize:
                        raise HDF5Error("fractal heap object outside its block")
                    return cache[baddr][pos:pos + n]
            raise HDF5Error("fractal heap object at offset %d not found" % off)

        return {"object": obj}

    def _btree2_records(self, addr, want_type):
        c = self._cursor(addr, 16 + 2 * self._so + self._sl)
        if c.raw(4) != b"BTHD":
            raise HDF5Error("version-2 B-tree at %d: bad signature" % addr)
        c.skip(1)
        if c.u(1) != want_type:
            raise HDF5Error("version-2 B-tree at %d: unexpected record type" % addr)
        node_size, rec_size, depth = c.u(4), c.u(2), c.u(2)
        c.skip(2)
        root, nroot = c.offset(), c.u(2)
        if depth > 1:
            raise HDF5Error("version-2 B-tree at %d: depth %d is not supported" % (addr, depth))
        if self._undefined(root) or nroot == 0:
            return []
        max_leaf = (node_size - 10) // rec_size
        count_bytes = _bytes_for(max_leaf)

        def leaf(a, n):
            d = self._cursor(a, 6 + n * rec_size)
            if d.raw(4) != b"BTLF":
                raise HDF5Error("version-2 B-tree leaf at %d: bad signature" % a)
            d.skip(2)
            return [d.raw(rec_size) for _ in range(n)]

        if depth == 0:
            return leaf(root, nroot)
        d = self._cursor(root, 6 + nroot * rec_size + (nroot + 1) * (self._so + count_bytes))
        if d.raw(4) != b"BTIN":
            raise HDF5Error("version-2 B-tree internal node at %d: bad signature" % root)
        d.skip(2)
        records = [d.raw(rec_size) for _ in range(nroot)]
        for _ in range(nroot + 1):
            child, n = d.offset(), d.u(count_bytes)
            records.extend(leaf(child, n))
        return records

    # -------------------------------------------------------------- datasets --
    def _resolve(self, path):
        addr = self._root
        for part in [p for p in path.split("/") if p]:
            links = self._links(addr)
            if part not in links:
                raise KeyError("%s: no object %r in %s" % (self.path, part, path))
            addr = links[part]
        return addr

    def keys(self, group="/"):
        return sorted(self._links(self._resolve(group)))

    def _dataset(self, addr, path):
        ds = _Dataset()
        ds.shape = ds.dtype = ds.layout = None
        ds.filters, ds.fill = [], None
        for mtype, mflags, data in self._messages(addr):
            if mflags & 0x02 and mtype in (_DATASPACE, _DATATYPE, _LAYOUT, _FILTERS):
                raise HDF5Error("%s: %s uses shared object header messages, which are not supported"
                                % (self.path, path))
            if mtype == _DATASPACE:
                ds.shape = self._dataspace(data)
            elif mtype == _DATATYPE:
                ds.dtype = _datatype(data, path)
            elif mtype == _LAYOUT:
                ds.layout = data
            elif mtype == _FILTERS:
                ds.filters = _filters(data)
            elif mtype == _FILL:
                ds.fill = _fill_value(data)
        if ds.shape is None or ds.dtype is None or ds.layout is None:
            raise HDF5Error("%s: %s is not a dataset" % (self.path, path))
        return ds

    def _dataspace(self, data):
        c = _Cursor(data, 0, self._so, self._sl)
        version, rank, flags = c.u(1), c.u(1), c.u(1)
        if version == 1:
            c.skip(5)
        elif version == 2:
            kind = c.u(1)
            if kind == 2:
                return None
            if kind == 0:
                return ()
        else:
            raise HDF5Error("dataspace version %d is not supported" % version)
        return tuple(c.length() for _ in range(rank))

    def read(self, path):
        """The dataset at ``path`` as a numpy array (native byte order)."""
        ds = self._dataset(self._resolve(path), path)
        shape, dtype = ds.shape, ds.dtype
        if shape is None:
            raise HDF5Error("%s: %s has a null dataspace" % (self.path, path))
        count = int(np.prod(shape)) if shape else 1
        fill = np.zeros((), dtype)
        if ds.fill is not None and len(ds.fill) == dtype.itemsize:
            fill = np.frombuffer(ds.fill, dtype)[0]
        lay = _Cursor(ds.layout, 0, self._so, self._sl)
        version = lay.u(1)
        if version in (1, 2):
            ndims, cls = lay.u(1), lay.u(1)
            lay.skip(5)
            address = lay.offset() if cls != 0 else None
            dims = [lay.u(4) for _ in range(ndims)]
            if cls == 0:
                raw = lay.raw(lay.u(4))
                return _values(raw, dtype, shape, count)
            if cls == 1:
                return self._contiguous(address, dtype, shape, count, fill)
            lay.u(4)  # element size
            chunks = self._btree1(address, 1, ndims) if not self._undefined(address) else []
            return self._assemble(chunks, dims[:len(shape)], ds, shape, fill, path)
        if version == 3 or version == 4:
            cls = lay.u(1)
            if cls == 0:
                return _values(lay.raw(lay.u(2)), dtype, shape, count)
            if cls == 1:
                address = lay.offset()
                lay.length()
                return self._contiguous(address, dtype, shape, count, fill)
            if cls != 2:
                raise HDF5Error("%s: %s: layout class %d is not supported" % (self.path, path, cls))
            if version == 3:
                ndims = lay.u(1)
                address = lay.offset()
                dims = [lay.u(4) for _ in range(ndims)]
                chunks = self._btree1(address, 1, ndims) if not self._undefined(address) else []
                return self._assemble(chunks, dims[:len(shape)], ds, shape, fill, path)
            return self._chunked_v4(lay, ds, shape, fill, path)
        raise HDF5Error("%s: %s: data layout version %d is not supported" % (self.path, path, version))

    def _contiguous(self, address, dtype, shape, count, fill):
        if self._undefined(address):
            return np.full(shape, fill, dtype)
        return _values(self._read(address, count * dtype.itemsize), dtype, shape, count)

    def _chunked_v4(self, lay, ds, shape, fill, path):
        flags, ndims = lay.u(1), lay.u(1)
        enc = lay.u(1)
        dims = [lay.u(enc) for _ in range(ndims)]
        index = lay.u(1)
        cdims = dims[:len(shape)]
        itemsize = ds.dtype.itemsize
        chunk_bytes = int(np.prod(cdims)) * itemsize
        grid = [-(-n // c) for n, c in zip(shape, cdims)]
        if index == 1:  # single chunk
            size, mask = chunk_bytes, 0
            if flags & 0x02:
                size, mask = lay.length(), lay.u(4)
            address = lay.offset()
            chunks = [] if self._undefined(address) else [(address, size, mask, [0] * len(shape))]
        elif index == 2:  # implicit: every chunk, unfiltered, in row-major order
            address = lay.offset()
            chunks = []
            if not self._undefined(address):
                for k, idx in enumerate(np.ndindex(*grid)):
                    chunks.append((address + k * chunk_bytes, chunk_bytes, 0,
                                   [i * c for i, c in zip(idx, cdims)]))
        elif index == 3:  # fixed array
            lay.u(1)  # page bits
            address = lay.offset()
            chunks = [] if self._undefined(address) else self._fixed_array(address, grid, cdims, path)
        else:
            raise HDF5Error("%s: %s: chunk index type %d (%s) is not supported" % (
                self.path, path, index, {4: "extensible array", 5: "version-2 B-tree"}.get(index, "unknown")))
        return self._assemble(chunks, cdims, ds, shape, fill, path)

    def _fixed_array(self, addr, grid, cdims, path):
        c = self._cursor(addr, 12 + self._sl + self._so)
        if c.raw(4) != b"FAHD":
            raise HDF5Error("fixed array header at %d: bad signature" % addr)
        c.skip(1)
        client, entry_size, page_bits = c.u(1), c.u(1), c.u(1)
        n = c.length()
        block = c.offset()
        if n > (1 << page_bits):
            raise HDF5Error("%s: %s: paged fixed-array chunk index is not supported" % (self.path, path))
        d = self._cursor(block, 6 + self._so + n * entry_size)
        if d.raw(4) != b"FADB":
            raise HDF5Error("fixed array data block at %d: bad signature" % block)
        d.skip(2 + self._so)
        out = []
        for k, idx in enumerate(np.ndindex(*grid)):
            if k >= n:
                break
            caddr = d.offset()
            if client == 1:
                size = d.u(entry_size - self._so - 4)
                mask = d.u(4)
            else:
                size, mask = None, 0
            if not self._undefined(caddr):
                out.append((caddr, size, mask, [i * c for i, c in zip(idx, cdims)]))
        return out

    def _assemble(self, chunks, cdims, ds, shape, fill, path):
        dtype = ds.dtype
        out = np.full(shape, fill, dtype)
        chunk_count = int(np.prod(cdims))
        for address, size, mask, offsets in chunks:
            if size is None:
                size = chunk_count * dtype.itemsize
            raw = self._read(address, size)
            raw = _unfilter(raw, ds.filters, mask, dtype.itemsize, path)
            block = _values(raw, dtype, tuple(cdims), chunk_count)
            dst = tuple(slice(o, min(o + c, n)) for o, c, n in zip(offsets, cdims, shape))
            src = tuple(slice(0, s.stop - s.start) for s in dst)
            out[dst] = block[src]
        return out


def _values(raw, dtype, shape, count):
    if len(raw) < count * dtype.itemsize:
        raise HDF5Error("dataset storage holds %d bytes, %d needed" % (len(raw), count * dtype.itemsize))
    a = np.frombuffer(raw, dtype, count).reshape(shape)
    return a.astype(dtype.newbyteorder("="))


def _datatype(data, path):
    cls, version = data[0] & 0x0F, data[0] >> 4
    b0 = data[1]
    size = int.from_bytes(data[4:8], "little")
    order = ">" if b0 & 0x01 else "<"
    if cls == 0:
        if size not in (1, 2, 4, 8):
            raise HDF5Error("%s: integer size %d is not supported" % (path, size))
        return np.dtype("%s%s%d" % (order, "i" if b0 & 0x08 else "u", size))
    if cls == 1:
        if b0 & 0x40:
            raise HDF5Error("%s: VAX floating point is not supported" % path)
        if size not in (2, 4, 8):
            raise HDF5Error("%s: float size %d is not supported" % (path, size))
        return np.dtype("%sf%d" % (order, size))
    raise HDF5Error("%s: datatype class %d (version %d) is not supported; only integers and floats are"
                    % (path, cls, version))


def _filters(data):
    c = _Cursor(data)
    version, n = c.u(1), c.u(1)
    if version == 1:
        c.skip(6)
    out = []
    for _ in range(n):
        fid = c.u(2)
        name_len = c.u(2) if (version == 1 or fid >= 256) else 0
        c.u(2)  # flags
        nvals = c.u(2)
        if name_len:
            c.skip(((name_len + 7) // 8) * 8 if version == 1 else name_len)
        vals = [c.u(4) for _ in range(nvals)]
        if version == 1 and nvals % 2:
            c.skip(4)
        out.append((fid, vals))
    return out


def _fill_value(data):
    version = data[0]
    if version in (1, 2):
        defined = data[3]
        if version == 2 and not defined:
            return None
        size = int.from_bytes(data[4:8], "little")
        return data[8:8 + size] if size else None
    if version == 3:
        flags = data[1]
        if not flags & 0x20:
            return None
        size = int.from_bytes(data[2:6], "little")
        return data[6:6 + size]
    return None


def _unfilter(raw, filters, mask, itemsize, path):
    for k in range(len(filters) - 1, -1, -1):
        if mask & (1 << k):
            continue
        fid, vals = filters[k]
        if fid == _FILTER_DEFLATE:
            raw = zlib.decompress(raw)
        elif fid == _FILTER_SHUFFLE:
            es = vals[0] if vals else itemsize
            n = len(raw) // es
            body = np.frombuffer(raw, np.uint8, n * es).reshape(es, n).T.tobytes()
            raw = body + raw[n * es:]
        elif fid == _FILTER_FLETCHER32:
            raw = raw[:-4]
        else:
            raise HDF5Error("%s: filter %d is not supported (deflate, shuffle and fletcher32 are)" % (path, fid))
    return raw


def read_dataset(path, dataset):
    """One dataset of the HDF5 file at ``path``."""
    with File(path) as f:
        return f.read(dataset)


# ---------------------------------------------------------------------------
# Writer: contiguous datasets in the root group, in the file format a default
# libhdf5 file uses (superblock version 0, version-1 object headers, a
# symbol-table root group).  Enough for the heavy data of XDMF series.
# ---------------------------------------------------------------------------

_UNDEF = (1 << 64) - 1
_LEAF_K_MIN, _INTERNAL_K = 4, 16
_SNOD_ENTRY = 40


def _le(v, n):
    return int(v).to_bytes(n, "little")


def _pad8(b):
    return b + b"\0" * (-len(b) % 8)


def _message(mtype, body):
    body = _pad8(body)
    return _le(mtype, 2) + _le(len(body), 2) + b"\0" * 4 + body


def _object_header(messages):
    body = b"".join(messages)
    return bytes([1, 0]) + _le(len(messages), 2) + _le(1, 4) + _le(len(body), 4) + b"\0" * 4 + body


def _datatype_message(dtype):
    dt = np.dtype(dtype)
    order = 0x01 if dt.byteorder == ">" or (dt.byteorder == "=" and not np.little_endian) else 0x00
    n = dt.itemsize
    if dt.kind == "f" and n in (4, 8):
        exp_bits, man_bits = (8, 23) if n == 4 else (11, 52)
        bits = bytes([0x20 | order, 8 * n - 1, 0])
        props = (_le(0, 2) + _le(8 * n, 2) + bytes([man_bits, exp_bits, 0, man_bits])
                 + _le((1 << (exp_bits - 1)) - 1, 4))
        return bytes([0x11]) + bits + _le(n, 4) + props
    if dt.kind in "iu" and n in (1, 2, 4, 8):
        bits = bytes([order | (0x08 if dt.kind == "i" else 0), 0, 0])
        return bytes([0x10]) + bits + _le(n, 4) + _le(0, 2) + _le(8 * n, 2)
    raise HDF5Error("write: dtype %s is not supported (floats of 4 or 8 bytes, integers)" % dt)


def write_datasets(path, datasets):
    """Write ``{name: array}`` as contiguous datasets of the root group of a
    new HDF5 file (names without '/'); arrays keep their dtype (floats of 4
    or 8 bytes, integers)."""
    names = sorted(datasets)
    for name in names:
        if not name or "/" in name or "\0" in name:
            raise HDF5Error("write: bad dataset name %r" % (name,))
    leaf_k = max(_LEAF_K_MIN, -(-len(names) // 2))
    # local heap data segment: "" at offset 0, then every name, 8-aligned
    heap = bytearray(b"\0" * 8)
    name_off = {}
    for name in names:
        name_off[name] = len(heap)
        heap += _pad8(name.encode("utf-8") + b"\0")
    heap = bytes(heap)

    superblock_size = 96
    btree_size = 24 + (2 * _INTERNAL_K + 1) * 8 + 2 * _INTERNAL_K * 8
    snod_size = 8 + 2 * leaf_k * _SNOD_ENTRY
    root_hdr = _object_header([_message(_SYMBOL_TABLE, b"\0" * 16)])
    a_root = superblock_size
    a_btree = a_root + len(root_hdr)
    a_snod = a_btree + btree_size
    a_heap = a_snod + snod_size
    a_heap_data = a_heap + 32
    pos = a_heap_data + len(heap)

    headers, blobs = [], []
    for name in names:
        a = np.ascontiguousarray(datasets[name])
        space = bytes([1, a.ndim, 0, 0]) + b"\0" * 4 + b"".join(_le(n, 8) for n in a.shape)
        fill = bytes([2, 2, 2, 0])
        hdr_len = len(_object_header([_message(_DATASPACE, space), _message(_DATATYPE, _datatype_message(a.dtype)),
                                      _message(_FILL, fill), _message(_LAYOUT, b"\0" * 18)]))
        headers.append((name, a, space, fill, pos))
        pos += hdr_len
    for name, a, space, fill, hdr_at in headers:
        blobs.append((hdr_at, pos, a))
        pos += a.nbytes
    eof = pos

    out = bytearray(eof)
    sb = (_SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + _le(leaf_k, 2) + _le(_INTERNAL_K, 2) + _le(0, 4)
          + _le(0, 8) + _le(_UNDEF, 8) + _le(eof, 8) + _le(_UNDEF, 8)
          + _le(0, 8) + _le(a_root, 8) + _le(1, 4) + _le(0, 4) + _le(a_btree, 8) + _le(a_heap, 8))
    assert len(sb) == superblock_size
    out[0:superblock_size] = sb
    out[a_root:a_root + len(root_hdr)] = _object_header(
        [_message(_SYMBOL_TABLE, _le(a_btree, 8) + _le(a_heap, 8))])
    last = name_off[names[-1]] if names else 0
    btree = (b"TREE" + bytes([0, 0]) + _le(1 if names else 0, 2) + _le(_UNDEF, 8) + _le(_UNDEF, 8)
             + _le(0, 8) + _le(a_snod, 8) + _le(last, 8))
    out[a_btree:a_btree + len(btree)] = btree
    snod = bytearray(b"SNOD" + bytes([1, 0]) + _le(len(names), 2))
    for (name, a, space, fill, hdr_at) in headers:
        snod += _le(name_off[name], 8) + _le(hdr_at, 8) + _le(0, 4) + _le(0, 4) + b"\0" * 16
    out[a_snod:a_snod + len(snod)] = snod
    out[a_heap:a_heap + 32] = b"HEAP" + bytes([0, 0, 0, 0]) + _le(len(heap), 8) + _le(1, 8) + _le(a_heap_data, 8)
    out[a_heap_data:a_heap_data + len(heap)] = heap
    for (name, a, space, fill, hdr_at), (_, data_at, _) in zip(headers, blobs):
        layout = bytes([3, 1]) + _le(data_at, 8) + _le(a.nbytes, 8)
        hdr = _object_header([_message(_DATASPACE, space), _message(_DATATYPE, _datatype_message(a.dtype)),
                              _message(_FILL, fill), _message(_LAYOUT, layout)])
        out[hdr_at:hdr_at + len(hdr)] = hdr
        out[data_at:data_at + a.nbytes] = a.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))
