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


# --- checkpoint sets of an nprocY x nprocZ run --------------------------------

def rank_blocks(ny, nz, nprocY, nprocZ):
    """The y/z block of every rank of the reference's 2-D decomposition
    (Channel-Flow/para.f90:218-262): rank myid has indy = myid mod nprocY,
    indz = myid div nprocY; the first ny mod nprocY columns (and
    nz mod nprocZ rows) get one node more; offsets are the sums of the
    blocks before.  Returns [(myid, y0, ly, z0, lz), ...] by rank."""
    if nprocY < 1 or nprocZ < 1 or nprocY > ny or nprocZ > nz:
        raise ValueError(f"decomposition {nprocY} x {nprocZ} does not fit {ny} x {nz}")
    ly = [ny // nprocY + (1 if i < ny % nprocY else 0) for i in range(nprocY)]
    lz = [nz // nprocZ + (1 if k < nz % nprocZ else 0) for k in range(nprocZ)]
    y0 = np.concatenate([[0], np.cumsum(ly)[:-1]])
    z0 = np.concatenate([[0], np.cumsum(lz)[:-1]])
    out = []
    for myid in range(nprocY * nprocZ):
        iy, iz = myid % nprocY, myid // nprocY
        out.append((myid, int(y0[iy]), ly[iy], int(z0[iz]), lz[iz]))
    return out


def cntd_name(directory, istep, myid):
    """cntdflow/endrunflow2D16x8.SSSSSSS.NNN (saveload.f90:204-220)."""
    return os.path.join(directory, f"endrunflow2D16x8.{istep:07d}.{myid:03d}")


def finit_name(directory, myid):
    """initflow/finit.NNN (saveload.f90:108-114)."""
    return os.path.join(directory, f"finit.{myid:03d}")


def split_ranks(f, nprocY, nprocZ):
    """The per-rank blocks f(0:18, lx, ly_r, lz_r) of a global f (nz, ny, lx, 19)."""
    nz, ny = f.shape[0], f.shape[1]
    return [np.ascontiguousarray(f[z0:z0 + lz, y0:y0 + ly]) for _, y0, ly, z0, lz in
            rank_blocks(ny, nz, nprocY, nprocZ)]


def join_ranks(blocks, ny, nz, nprocY, nprocZ):
    """The global f (nz, ny, lx, 19) assembled from per-rank blocks."""
    lx = blocks[0].shape[2]
    f = np.empty((nz, ny, lx, 19))
    for (myid, y0, ly, z0, lz), b in zip(rank_blocks(ny, nz, nprocY, nprocZ), blocks):
        if b.shape != (lz, ly, lx, 19):
            raise ValueError(f"rank {myid}: block {b.shape}, want {(lz, ly, lx, 19)}")
        f[z0:z0 + lz, y0:y0 + ly] = b
    return f


def write_cntd_ranks(directory, f, istep, nprocY, nprocZ, istat=0, imovie=0):
    """The checkpoint set an nprocY x nprocZ run of the reference reads back
    (loadcntdflow, saveload.f90:296-332), one file per rank."""
    for myid, b in enumerate(split_ranks(f, nprocY, nprocZ)):
        write_cntd(cntd_name(directory, istep, myid), b, istep, istat, imovie)


def read_cntd_ranks(directory, istep, lx, ny, nz, nprocY, nprocZ):
    """Assemble the checkpoint set of an nprocY x nprocZ run into one lattice
    (what loadcntdflow_frmmore, saveload.f90:336-381, does for fewer ranks).
    Every rank's header must agree."""
    heads, blocks = set(), []
    for myid, _, ly, _, lz in rank_blocks(ny, nz, nprocY, nprocZ):
        head, b = read_cntd(cntd_name(directory, istep, myid), (lz, ly, lx, 19))
        heads.add(head)
        blocks.append(b)
    if len(heads) != 1:
        raise ValueError(f"checkpoint set at step {istep}: rank headers differ {sorted(heads)}")
    return heads.pop(), join_ranks(blocks, ny, nz, nprocY, nprocZ)


def write_finit_ranks(directory, f, nprocY, nprocZ, istat=0):
    for myid, b in enumerate(split_ranks(f, nprocY, nprocZ)):
        write_finit(finit_name(directory, myid), b, istat)


def read_finit_ranks(directory, lx, ny, nz, nprocY, nprocZ):
    stats, blocks = set(), []
    for myid, _, ly, _, lz in rank_blocks(ny, nz, nprocY, nprocZ):
        istat, b = read_finit(finit_name(directory, myid), (lz, ly, lx, 19))
        stats.add(istat)
        blocks.append(b)
    if len(stats) != 1:
        raise ValueError(f"finit set: rank istat values differ {sorted(stats)}")
    return stats.pop(), join_ranks(blocks, ny, nz, nprocY, nprocZ)


def read_endrunflow_zslabs(directory, istpload, lx, ly, nz, nfiles):
    """The older z-slab checkpoint set endrunflow.SSSSSS.NNN of nfiles ranks,
    each holding lz = nz/nfiles planes (rank NNN planes NNN*lz .. +lz),
    assembled into one lattice: loadcntdflow_frmmore (saveload.f90:336-381)
    with every rank's slabs gathered on one device."""
    if nz % nfiles:
        raise ValueError(f"{nz} planes do not split into {nfiles} equal slabs")
    lz = nz // nfiles
    f = np.empty((nz, ly, lx, 19))
    heads = set()
    for r in range(nfiles):
        head, b = read_cntd(os.path.join(directory, f"endrunflow.{istpload:06d}.{r:03d}"), (lz, ly, lx, 19))
        heads.add(head)
        f[r * lz:(r + 1) * lz] = b
    if len(heads) != 1:
        raise ValueError(f"z-slab set: headers differ {sorted(heads)}")
    return heads.pop(), f
