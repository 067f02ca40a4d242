"""The reference's restart file for the velocity state (Fortran unformatted, sequential).

io.f90:740-786 (checkpoint_final_v2) writes, with S_FLAG and PCon_FLAG off, one record
to restart/vel_<2000+coord>.out:

    u, v, w, RHSx, RHSy, RHSz (levels 1:nz), Cs_opt2, F_LM, F_MM, F_QN, F_NN

initial.f90:282-297 reads vel.out in the same order: with model 5 and inilag only the
first seven arrays, otherwise all eleven.  Every array is REAL(8) of shape (ld, ny, nz)
in Fortran order, which is the context's C layout [nz][ny][ld]; the record is their
concatenation.

Record framing: a 4-byte little-endian byte count before and after the data.  A record
longer than 2^31 - 1 bytes is split into subrecords; the counts of every subrecord but
the last are negated (the gfortran / flang convention).  The framing of records below
that size is pinned against amdflang-written files (tests/test_restart.py); the
subrecord form is only checked against this module's own writer.

The droplet and scalar variants (vel_pcon, vel_sc, vel_sc_pcon) also carry arrays this
implementation does not model (deposition, Real_dep, the dynamic-Schmidt F_KX..F_XX2,
theta), so they are not read or written here.
"""
import struct

import numpy as np

VEL_FIELDS = ["u", "v", "w", "RHSx", "RHSy", "RHSz", "Cs_opt2", "F_LM", "F_MM", "F_QN", "F_NN"]
VEL_FIELDS_INILAG = VEL_FIELDS[:7]  # initial.f90:290-293 (model 5 with inilag)
_MAX_SUB = 2**31 - 1


def write_record(f, arrays, max_sub=_MAX_SUB):
    """One unformatted sequential record holding the arrays' bytes back to back."""
    views = [memoryview(np.ascontiguousarray(a)).cast("B") for a in arrays]
    total = sum(len(v) for v in views)
    if total <= max_sub:
        f.write(struct.pack("<i", total))
        for v in views:
            f.write(v)
        f.write(struct.pack("<i", total))
        return
    # subrecords: counts negated on all but the last
    pieces = []  # byte counts of the subrecords
    left = total
    while left > 0:
        pieces.append(min(left, max_sub))
        left -= pieces[-1]
    it = iter(views)
    cur, pos = next(it), 0
    for n, size in enumerate(pieces):
        mark = struct.pack("<i", size if n == len(pieces) - 1 else -size)
        f.write(mark)
        need = size
        while need > 0:
            if pos == len(cur):
                cur, pos = next(it), 0
                continue
            take = min(need, len(cur) - pos)
            f.write(cur[pos:pos + take])
            pos += take
            need -= take
        f.write(mark)


def read_record(f):
    """The data bytes of the next record (subrecords joined)."""
    parts = []
    while True:
        head = f.read(4)
        if len(head) != 4:
            raise EOFError("restart file: no record marker")
        (m,) = struct.unpack("<i", head)
        data = f.read(abs(m))
        tail = f.read(4)
        if len(data) != abs(m) or len(tail) != 4 or struct.unpack("<i", tail)[0] != m:
            raise ValueError("restart file: record markers do not match")
        parts.append(data)
        if m >= 0:
            return b"".join(parts)


def write_vel(lg, path, max_sub=_MAX_SUB):
    """checkpoint_final_v2 without scalars (io.f90:772-776): the context state as REAL(8)."""
    g = lg.g
    arrays = []
    for name in VEL_FIELDS:
        a = lg.download(name).reshape(-1)[: g.ld * g.ny * g.nz]
        arrays.append(a.astype(np.float64))
    with open(path, "wb") as f:
        write_record(f, arrays, max_sub)


def read_vel(lg, path, inilag=False):
    """initial.f90:282-297 (model 5): the record's leading arrays into the context (the
    state that a restarted run starts from); returns the names that were set."""
    g = lg.g
    names = VEL_FIELDS_INILAG if inilag else VEL_FIELDS
    n = g.ld * g.ny * g.nz
    with open(path, "rb") as f:
        data = read_record(f)
    if len(data) < 8 * n * len(names):
        raise ValueError(f"restart file: {len(data)} bytes, {8 * n * len(names)} needed for {len(names)} arrays")
    vals = np.frombuffer(data, dtype="<f8", count=n * len(names))
    for q, name in enumerate(names):
        lg.upload(name, vals[q * n:(q + 1) * n].reshape(g.nz, g.ny, g.ld))
    return names
