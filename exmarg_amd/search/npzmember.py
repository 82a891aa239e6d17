"""
Where an uncompressed .npy member of an .npz archive lies inside the file, so that it can be read
sequentially, memory-mapped or streamed by the HIP library's reader threads without unpacking the
archive (the file format the reference's store and segment files use: src/exmarg/search/store.py:23-52,
src/exmarg/search/segments.py:40-54).

A zip entry's data starts after its 30-byte local file header (signature PK\\x03\\x04, then fixed
fields; the file-name and extra-field lengths are the little-endian uint16 at bytes 26 and 28) and
those two variable fields; an .npy member's data is its .npy header followed by the raw array.
"""

import struct
import zipfile
from dataclasses import dataclass

import numpy as np

_LOCAL = struct.Struct("<4s22sHH")          # signature, fixed fields, name length, extra length


@dataclass(frozen=True)
class Member:
    """the array of one .npy member: file, byte offset of its data, shape, dtype, Fortran order"""
    path: str
    offset: int
    shape: tuple
    dtype: np.dtype
    fortran: bool

    @property
    def count(self):
        return int(np.prod(self.shape))

    def read(self):
        """the whole array, read sequentially into host memory"""
        with open(self.path, "rb") as fh:
            fh.seek(self.offset)
            flat = np.fromfile(fh, dtype=self.dtype, count=self.count)
        return flat.reshape(self.shape, order="F" if self.fortran else "C")

    def mapped(self):
        """a read-only memory map of the array (nothing is read until used)"""
        return np.memmap(self.path, dtype=self.dtype, mode="r", offset=self.offset, shape=self.shape,
                         order="F" if self.fortran else "C")


def locate(path, name, compressed_ok=False):
    """the Member of `name` (with its .npy suffix) in the archive at path.  A compressed member raises
    ValueError, or returns None when compressed_ok (the caller then loads it through numpy)."""
    with zipfile.ZipFile(path) as zf:
        info = zf.getinfo(name)
    if info.compress_type != zipfile.ZIP_STORED:
        if compressed_ok:
            return None
        raise ValueError("member %s of %s is compressed" % (name, path))
    with open(path, "rb") as fh:
        fh.seek(info.header_offset)
        sig, _, name_len, extra_len = _LOCAL.unpack(fh.read(_LOCAL.size))
        if sig != b"PK\x03\x04":
            raise ValueError("%s: no local header for member %s" % (path, name))
        fh.seek(name_len + extra_len, 1)
        major, minor = np.lib.format.read_magic(fh)
        header = np.lib.format.read_array_header_1_0 if (major, minor) == (1, 0) else \
            np.lib.format.read_array_header_2_0
        shape, fortran, dtype = header(fh)
        return Member(path=str(path), offset=fh.tell(), shape=tuple(shape), dtype=np.dtype(dtype),
                      fortran=bool(fortran))
