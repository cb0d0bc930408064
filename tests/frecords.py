"""Fortran unformatted sequential files in GNU Fortran's convention, written and read from the
published description of its format (the gfortran manual, "File format of unformatted
sequential files"):

  a logical record is one or more subrecords; a subrecord is a 4-byte little-endian int32
  marker, at most 2147483639 bytes of data, and a closing marker; both markers hold the
  subrecord's data length, the leading one negated when another subrecord follows, the trailing
  one negated when one precedes.

Used by the tests as the independent reader / writer of the driver's restart and snapshot files.
"""
import struct

MAX_SUBRECORD = 2147483639


def write_records(path, records, max_sub=MAX_SUBRECORD):
    """records: a list of bytes-like objects, one per logical record."""
    with open(path, "wb") as fh:
        for rec in records:
            data = memoryview(bytes(rec))
            n, pos, first = len(data), 0, True
            while True:
                k = min(max_sub, n - pos)
                last = pos + k >= n
                fh.write(struct.pack("<i", k if last else -k))
                fh.write(data[pos:pos + k])
                fh.write(struct.pack("<i", k if first else -k))
                pos += k
                first = False
                if last:
                    break


def read_records(path, with_subrecords=False):
    """A list of bytes objects, one per logical record (and, with with_subrecords, the list of
    subrecord lengths of each)."""
    out, subs = [], []
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        parts, lens, first = [], [], True
        while True:
            (lead,) = struct.unpack_from("<i", data, pos)
            k = abs(lead)
            parts.append(data[pos + 4:pos + 4 + k])
            (trail,) = struct.unpack_from("<i", data, pos + 4 + k)
            if abs(trail) != k or (trail < 0) == first:
                raise ValueError(f"bad trailing marker {trail} for a subrecord of {k} bytes at {pos}")
            lens.append(k)
            pos += k + 8
            first = False
            if lead >= 0:
                break
        out.append(b"".join(parts))
        subs.append(lens)
    return (out, subs) if with_subrecords else out
