"""Minimal MATLAB MAT-file (level 5) reader for the reconstruction inputs.

Reads the numeric variables of a version-5 MAT-file: real double, single and
integer arrays (and logical), stored plain or zlib-compressed (miCOMPRESSED),
in little- or big-endian files.  Cell, struct, object, sparse, complex and
character arrays are skipped with their names listed in `skipped`.  Nothing
in the file is executed: the format is parsed field by field with `struct`
and numpy.

Layout (MATLAB "MAT-File Format", level 5): a 128-byte header (text, then a
version word 0x0100 and the endian indicator "IM"/"MI"), then data elements.
An element is an 8-byte tag (data type, byte count) followed by its data,
padded to 8 bytes; a "small element" packs a type (low 16 bits), a count
(high 16 bits) and up to 4 bytes of data into the tag itself.  A variable is
an miMATRIX element holding the array flags, the dimensions, the name and the
real part.

    arrays = read_mat("testdata/swine_2d.mat")   # {"prjSwine": ndarray, ...}
"""
import struct
import zlib

import numpy as np

# data types (element tags)
MI_INT8, MI_UINT8, MI_INT16, MI_UINT16, MI_INT32, MI_UINT32 = 1, 2, 3, 4, 5, 6
MI_SINGLE, MI_DOUBLE, MI_INT64, MI_UINT64 = 7, 9, 12, 13
MI_MATRIX, MI_COMPRESSED, MI_UTF8 = 14, 15, 16

_MI_DTYPE = {MI_INT8: "i1", MI_UINT8: "u1", MI_INT16: "i2", MI_UINT16: "u2", MI_INT32: "i4",
             MI_UINT32: "u4", MI_SINGLE: "f4", MI_DOUBLE: "f8", MI_INT64: "i8", MI_UINT64: "u8",
             MI_UTF8: "u1"}

# array classes (array flags, low byte)
MX_CELL, MX_STRUCT, MX_OBJECT, MX_CHAR, MX_SPARSE = 1, 2, 3, 4, 5
MX_DOUBLE, MX_SINGLE = 6, 7
MX_INT8, MX_UINT8, MX_INT16, MX_UINT16, MX_INT32, MX_UINT32, MX_INT64, MX_UINT64 = range(8, 16)

_MX_DTYPE = {MX_DOUBLE: "f8", MX_SINGLE: "f4", MX_INT8: "i1", MX_UINT8: "u1", MX_INT16: "i2",
             MX_UINT16: "u2", MX_INT32: "i4", MX_UINT32: "u4", MX_INT64: "i8", MX_UINT64: "u8"}
_FLAG_COMPLEX, _FLAG_LOGICAL = 0x0800, 0x0200


class MatError(ValueError):
    pass


class _Reader:
    def __init__(self, buf, order):
        self.buf, self.pos, self.o = buf, 0, order

    def done(self):
        return self.pos >= len(self.buf)

    def element(self):
        """(type, payload bytes) of the next element; advances past padding."""
        if self.pos + 8 > len(self.buf):
            raise MatError("truncated element tag")
        word, = struct.unpack_from(self.o + "I", self.buf, self.pos)
        if word >> 16:  # small element: count in the high half, data in the tag
            typ, n = word & 0xFFFF, word >> 16
            if n > 4:
                raise MatError("bad small element")
            data = self.buf[self.pos + 4:self.pos + 4 + n]
            self.pos += 8
            return typ, data
        typ, n = struct.unpack_from(self.o + "II", self.buf, self.pos)
        start = self.pos + 8
        if start + n > len(self.buf):
            raise MatError("truncated element data")
        data = self.buf[start:start + n]
        # compressed elements are not padded; everything else pads to 8 bytes
        self.pos = start + (n if typ == MI_COMPRESSED else (n + 7) // 8 * 8)
        return typ, data


def _numeric(typ, data, order):
    if typ not in _MI_DTYPE:
        raise MatError(f"unsupported numeric data type {typ}")
    return np.frombuffer(data, dtype=np.dtype(order + _MI_DTYPE[typ]))


def _matrix(payload, order, out, skipped):
    r = _Reader(payload, order)
    typ, flags = r.element()
    if typ != MI_UINT32 or len(flags) < 8:
        raise MatError("bad array flags")
    f0, = struct.unpack_from(order + "I", flags, 0)
    cls, complex_, logical = f0 & 0xFF, bool(f0 & _FLAG_COMPLEX), bool(f0 & _FLAG_LOGICAL)
    typ, dims = r.element()
    dims = tuple(int(d) for d in _numeric(typ, dims, order))
    typ, name = r.element()
    name = bytes(name).decode("ascii", "replace")
    if cls not in _MX_DTYPE or complex_:
        skipped.append(name)
        return
    typ, real = r.element()
    a = _numeric(typ, real, order).astype(np.dtype(order + _MX_DTYPE[cls]).newbyteorder("="))
    if a.size != int(np.prod(dims)):
        raise MatError(f"{name}: {a.size} values for dimensions {dims}")
    a = a.reshape(dims, order="F")  # MATLAB stores column-major
    out[name] = a.astype(bool) if logical else a


def read_mat(path_or_bytes):
    """Numeric variables of a level-5 MAT-file as {name: ndarray} (MATLAB
    dimensions, column-major order preserved).  The names of variables of
    other classes are in the result's `skipped` attribute."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        buf = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as fh:
            buf = fh.read()
    if len(buf) < 128:
        raise MatError("not a MAT-file (short header)")
    endian = buf[126:128]
    if endian == b"IM":
        order = "<"
    elif endian == b"MI":
        order = ">"
    else:
        raise MatError("not a level-5 MAT-file (endian indicator)")
    version, = struct.unpack_from(order + "H", buf, 124)
    if version != 0x0100:
        raise MatError(f"unsupported MAT-file version {version:#x}")
    out, skipped = _MatDict(), []
    r = _Reader(buf[128:], order)
    while not r.done():
        typ, data = r.element()
        if typ == MI_COMPRESSED:
            inner = _Reader(zlib.decompress(data), order)
            typ, data = inner.element()
        if typ == MI_MATRIX:
            _matrix(data, order, out, skipped)
    out.skipped = skipped
    return out


class _MatDict(dict):
    skipped = ()
