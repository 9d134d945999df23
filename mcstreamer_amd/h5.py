"""A small HDF5 writer, enough for the reference's HDF5 field export
(MCStreamer/src/plot.jl:210-240 `h5export`): nested groups of contiguous
datasets of little-endian float64 or int64, scalars and arrays.

The file is plain HDF5 as the HDF5 format specification lays it out, in its
oldest forms, which every HDF5 reader accepts: a version-0 superblock (8-byte
offsets and lengths), version-1 object headers, and groups as symbol tables
(a version-1 B-tree with one symbol-table node, names in a local heap).  An
array's dimensions are written as numpy holds them (C order, last index
fastest): a Julia matrix of size (M, N) is stored by HDF5.jl with dimensions
(N, M), which is the numpy array of shape (N, M) that holds the same memory.

    write("fields.h5", {"grid": {"R": 1e-3, "M": np.int64(10), ...},
                        "u": np.zeros((N, M)), ...})
"""
import struct

import numpy as np

__all__ = ["write"]

_UNDEF = 0xFFFFFFFFFFFFFFFF
_FREE_NULL = 1          # "no free block" in a local heap's free-list offset
_LEAF_K = 16            # a symbol-table node holds up to 2 * 16 entries
_NODE_K = 16            # a group B-tree node up to 2 * 16 children
_SIG = b"\x89HDF\r\n\x1a\n"

_F64 = bytes([0x11, 0x20, 0x3F, 0x00]) + struct.pack("<I", 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
_I64 = bytes([0x10, 0x08, 0x00, 0x00]) + struct.pack("<I", 8) + struct.pack("<HH", 0, 64)


def _pad8(b):
    return b + b"\0" * ((-len(b)) % 8)


def _message(mtype, body):
    body = _pad8(bytes(body))
    return struct.pack("<HHB3x", mtype, len(body), 0) + body


def _object_header(messages):
    data = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(data)) + data


class _File:
    def __init__(self):
        self.buf = bytearray(b"\0" * 96)   # the superblock, filled in last

    def append(self, data):
        self.buf += b"\0" * ((-len(self.buf)) % 8)
        at = len(self.buf)
        self.buf += data
        return at

    def dataset(self, value):
        a = np.asarray(value)
        if a.dtype.kind == "f":
            a, dtype = a.astype("<f8"), _F64
        elif a.dtype.kind in "iub":
            a, dtype = a.astype("<i8"), _I64
        else:
            raise TypeError(f"h5: no HDF5 type for {a.dtype}")
        raw = np.ascontiguousarray(a).tobytes()
        addr = self.append(raw) if raw else _UNDEF
        space = struct.pack("<BBBB4x", 1, a.ndim, 0, 0) + b"".join(struct.pack("<Q", d) for d in a.shape)
        msgs = [
            _message(0x0001, space),
            _message(0x0003, dtype),
            _message(0x0005, bytes([2, 1, 2, 0])),   # fill value: early allocation, none defined
            _message(0x0008, struct.pack("<BBQQ", 3, 1, addr, len(raw))),   # contiguous layout
        ]
        return self.append(_object_header(msgs))

    def group(self, items):
        """The group of `items` (name -> value or dict); returns (object header,
        B-tree, heap) addresses."""
        names = sorted(items, key=lambda k: k.encode())
        if len(names) > 2 * _LEAF_K:
            raise ValueError(f"h5: at most {2 * _LEAF_K} members per group")
        entries = []
        for name in names:
            v = items[name]
            if isinstance(v, dict):
                oh, bt, hp = self.group(v)
                entries.append((name, oh, 1, struct.pack("<QQ", bt, hp)))
            else:
                entries.append((name, self.dataset(v), 0, b"\0" * 16))
        # local heap: the empty name at offset 0, then each name, 8-byte aligned
        seg = bytearray(b"\0" * 8)
        offs = {}
        for name in names:
            b = name.encode()
            if b"\0" in b or not b or b"/" in b:
                raise ValueError(f"h5: bad member name {name!r}")
            offs[name] = len(seg)
            seg += _pad8(b + b"\0")
        seg_addr = self.append(bytes(seg))
        heap = self.append(b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack("<QQQ", len(seg), _FREE_NULL, seg_addr))
        # one symbol-table node with every entry, in name order
        snod = bytearray(b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(entries)))
        for name, oh, cache, scratch in entries:
            snod += struct.pack("<QQII", offs[name], oh, cache, 0) + scratch
        snod += b"\0" * (8 + 2 * _LEAF_K * 40 - len(snod))
        snod_addr = self.append(bytes(snod))
        # the group's B-tree: one level-0 node pointing at it; key 0 is the empty
        # name, key 1 the last name in the node
        tree = bytearray(b"TREE" + bytes([0, 0]) + struct.pack("<H", 1 if entries else 0))
        tree += struct.pack("<QQ", _UNDEF, _UNDEF)
        if entries:
            tree += struct.pack("<QQQ", 0, snod_addr, offs[names[-1]])
        tree += b"\0" * (24 + (2 * _NODE_K + 1) * 8 + 2 * _NODE_K * 8 - len(tree))
        btree = self.append(bytes(tree))
        oh = self.append(_object_header([_message(0x0011, struct.pack("<QQ", btree, heap))]))
        return oh, btree, heap

    def finish(self, root):
        oh, bt, hp = root
        eof = len(self.buf)
        sb = _SIG + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", _LEAF_K, _NODE_K, 0)
        sb += struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF)
        sb += struct.pack("<QQII", 0, oh, 1, 0) + struct.pack("<QQ", bt, hp)
        assert len(sb) == 96
        self.buf[:96] = sb
        return bytes(self.buf)


def write(path, tree):
    """Write `tree` (name -> numpy array, number or nested dict) as an HDF5 file."""
    f = _File()
    data = f.finish(f.group(tree))
    with open(path, "wb") as fh:
        fh.write(data)
