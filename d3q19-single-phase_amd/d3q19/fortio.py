"""Fortran unformatted sequential files, as the channel checkpoints use them.

Channel-Flow/saveload.f90 writes `form='unformatted'` sequential files:
finit.NNN holds the records (istat) and (f) (saveload.f90:119-120);
endrunflow2D16x8.SSSSSSS.NNN holds (istep, istat, imovie) and (f) (:226-227).
Each record is framed by 4-byte little-endian length markers.  A record
longer than 2^31 - 1 bytes is split into subrecords whose leading marker is
negated when more subrecords follow (the gfortran convention for large
records); reading accepts both forms.
"""
import os
import struct

import numpy as np

_MAXREC = 2**31 - 1


def write_records(path, records):
    """Write a list of numpy arrays / bytes, one Fortran record each."""
    with open(path, "wb") as fh:
        for rec in records:
            data = rec.tobytes() if isinstance(rec, np.ndarray) else bytes(rec)
            pos = 0
            n = len(data)
            if n == 0:
                fh.write(struct.pack("<ii", 0, 0))
                continue
            while pos < n:
                chunk = data[pos:pos + _MAXREC]
                more = pos + len(chunk) < n
                first = pos == 0
                lead = -len(chunk) if more else len(chunk)
                trail = -len(chunk) if not first else len(chunk)
                fh.write(struct.pack("<i", lead))
                fh.write(chunk)
                fh.write(struct.pack("<i", trail))
                pos += len(chunk)


def read_records(path):
    """Return the list of records of a Fortran unformatted sequential file as bytes."""
    out = []
    size = os.path.getsize(path)
    with open(path, "rb") as fh:
        while fh.tell() < size:
            parts = []
            while True:
                (lead,) = struct.unpack("<i", fh.read(4))
                n = abs(lead)
                parts.append(fh.read(n))
                (trail,) = struct.unpack("<i", fh.read(4))
                if abs(trail) != n:
                    raise ValueError(f"{path}: record markers {lead} / {trail} disagree")
                if lead >= 0:
                    break
            out.append(b"".join(parts))
    return out


def write_finit(path, f, istat=0):
    """finit.NNN: (istat) then f(0:18,lx,ly,lz) in FP64 (saveload.f90:119-120)."""
    write_records(path, [np.array([istat], dtype="<i4"), np.ascontiguousarray(f, dtype="<f8")])


def read_finit(path, shape):
    recs = read_records(path)
    istat = int(np.frombuffer(recs[0], dtype="<i4")[0])
    return istat, np.frombuffer(recs[1], dtype="<f8").reshape(shape).copy()


def write_cntd(path, f, istep, istat=0, imovie=0):
    """endrunflow2D16x8.SSSSSSS.NNN: (istep, istat, imovie) then f (saveload.f90:226-227)."""
    write_records(path, [np.array([istep, istat, imovie], dtype="<i4"), np.ascontiguousarray(f, dtype="<f8")])


def read_cntd(path, shape):
    recs = read_records(path)
    istep, istat, imovie = (int(v) for v in np.frombuffer(recs[0], dtype="<i4")[:3])
    return (istep, istat, imovie), np.frombuffer(recs[1], dtype="<f8").reshape(shape).copy()
