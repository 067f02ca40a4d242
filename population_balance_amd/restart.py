"""The reference's restart files (Fortran unformatted, sequential; one record each).

io.f90:740-786 (checkpoint_final_v2) writes restart/<kind>_<2000+coord>.out with one of
four layouts, chosen by S_FLAG (temperature) and PCon_FLAG (droplets); initial_multi.f90:
48-83 reads the same four back.  Arrays in the record's order (levels 1:nz; REAL(8)):

    vel          u, v, w, RHSx, RHSy, RHSz, Cs_opt2, F_LM, F_MM, F_QN, F_NN
    vel_pcon     u, v, w, PCon, RHSx, RHSy, RHSz, RHS_PCon, Cs_opt2, F_LM, F_MM, F_QN, F_NN,
                 deposition, Real_dep, Kc_t, F_KX, F_XX, F_KX2, F_XX2
    vel_sc       u, v, w, theta, RHSx, RHSy, RHSz, RHS_T, sgs_t3(:,:,1), psi_m, Cs_opt2,
                 F_LM, F_MM, F_QN, F_NN
    vel_sc_pcon  u, v, w, theta, PCon, RHSx, RHSy, RHSz, RHS_T, sgs_t3(:,:,1), psi_m,
                 RHS_PCon, Cs_opt2, F_LM, F_MM, F_QN, F_NN, deposition, Real_dep, Kc_t,
                 F_KX, F_XX, F_KX2, F_XX2

Shapes (Fortran order; the C layout is the reverse): (ld, ny, nz) for the velocity-grid
volumes, theta, RHS_T and Kc_t; (ld, ny, nz, npcon) for PCon and RHS_PCon, which is the
context's [npcon][nz][ny][ld]; (ld, ny) for sgs_t3(:,:,1); (nx, ny) for psi_m,
deposition and Real_dep (scalars_module_crossflow_mod.f90:92-103, bottombc.f90:25);
(nx, ny, nz) for the dynamic-Schmidt averages F_KX, F_XX, F_KX2, F_XX2 (sgsmodule.f90:
37-38).

initial.f90:74-140 reads vel_sc.out (initu and initsc) and :282-297 vel.out (initu
without initsc) with a prefix of these layouts that depends on the SGS model number and
inilag (read_prefix).

The context holds u, v, w, the RHS, Cs_opt2, F_LM..F_NN, PCon and RHS_PCon, and with
the temperature on (cfg.scal.s_flag) theta, RHS_T, sgs_t3 (the record's level 1) and
psi_m.  The other arrays (droplet diagnostics: deposition, Real_dep; Kc_t; the
dynamic-Schmidt averages; the temperature's arrays for a context without it) are the
caller's: the writers take them from `extras` (zeros where absent) and the readers
return them, so a file passes through unchanged.  Kc_t is recomputed from Nu_t by every droplet step before it is
read (scalars_module_crossflow_mod.f90:3482-3549), so its stored value does not change
a restarted run.  The droplet inertia's accelerations dudt, dvdt, dwdt are not in any
layout: a restarted run (the reference's and this one's) starts them at zero.

Record framing: a 4-byte little-endian byte count before and after the data.  A record
longer than 2^31 - 1 bytes is split into subrecords; the counts of every subrecord but
the last are negated (the gfortran / flang convention).  The framing of records below
that size and the four layouts' array order are pinned against amdflang-written files
(tests/test_restart.py); the subrecord form is only checked against this module's own
writer.
"""
import struct

import numpy as np

VEL_FIELDS = ["u", "v", "w", "RHSx", "RHSy", "RHSz", "Cs_opt2", "F_LM", "F_MM", "F_QN", "F_NN"]
VEL_FIELDS_INILAG = VEL_FIELDS[:7]  # initial.f90:290-293 (model 5 with inilag)
_MAX_SUB = 2**31 - 1

_LAG = ["Cs_opt2", "F_LM", "F_MM", "F_QN", "F_NN"]
_DROP_TAIL = ["deposition", "Real_dep", "Kc_t", "F_KX", "F_XX", "F_KX2", "F_XX2"]
_SC = ["theta"]
_SC_MID = ["RHS_T", "sgs_t3", "psi_m"]
LAYOUTS = {
    "vel": ["u", "v", "w", "RHSx", "RHSy", "RHSz"] + _LAG,
    "vel_pcon": ["u", "v", "w", "PCon", "RHSx", "RHSy", "RHSz", "RHS_PCon"] + _LAG + _DROP_TAIL,
    "vel_sc": ["u", "v", "w"] + _SC + ["RHSx", "RHSy", "RHSz"] + _SC_MID + _LAG,
    "vel_sc_pcon": ["u", "v", "w"] + _SC + ["PCon", "RHSx", "RHSy", "RHSz"] + _SC_MID + ["RHS_PCon"] + _LAG
    + _DROP_TAIL,
}
CONTEXT_FIELDS = set(VEL_FIELDS) | {"PCon", "RHS_PCon"}
THETA_FIELDS = {"theta", "RHS_T", "sgs_t3", "psi_m"}  # context fields with the temperature on


def _held(lg, name):
    """whether the context holds record array `name`"""
    return name in CONTEXT_FIELDS or (name in THETA_FIELDS and lg.cfg.scal.s_flag)


def _get(lg, name, shp):
    """record array `name` (C-order shape shp) from the context"""
    a = lg.download(name).astype(np.float64)
    if name == "sgs_t3":
        return a[0]                      # sgs_t3(:,:,1)
    if name == "psi_m":
        return a[0, :, :shp[1]]          # (nx, ny) from the (ld, ny) plane
    return a.reshape(-1)[: int(np.prod(shp))].reshape(shp)


def _put(lg, name, a):
    """record array `name` into the context"""
    if name == "sgs_t3":
        v = lg.download("sgs_t3")
        v[0] = a
        a = v
    elif name == "psi_m":
        v = np.zeros(lg.download("psi_m").shape)
        v[0, :, :a.shape[1]] = a
        a = v
    lg.upload(name, a)


def kind_of(s_flag, pcon_flag):
    """io.f90:752-760: the layout checkpoint_final_v2 picks"""
    return ("vel_sc_pcon" if pcon_flag else "vel_sc") if s_flag else ("vel_pcon" if pcon_flag else "vel")


def shape_of(name, g, npcon):
    """C-order shape of a record array"""
    if name in ("PCon", "RHS_PCon"):
        return (npcon, g.nz, g.ny, g.ld)
    if name == "sgs_t3":
        return (g.ny, g.ld)
    if name in ("psi_m", "deposition", "Real_dep"):
        return (g.ny, g.nx)
    if name in ("F_KX", "F_XX", "F_KX2", "F_XX2"):
        return (g.nz, g.ny, g.nx)
    return (g.nz, g.ny, g.ld)


def read_prefix(model, inilag, initsc):
    """the arrays initial.f90 reads from the start of vel_sc.out (initsc, :89-123) or
    vel.out (:276-297) for SGS model `model` (1..5); model 4 without inilag reads two
    arrays into a scratch variable ("crap") after F_MM, returned here as F_QN, F_NN's
    slots under the names "crap1", "crap2"."""
    base = ["u", "v", "w"] + (["theta"] if initsc else []) + ["RHSx", "RHSy", "RHSz"] + \
        (["RHS_T", "sgs_t3", "psi_m"] if initsc else [])
    if model == 1:
        return base
    if model in (2, 3) or (model in (4, 5) and inilag):
        return base + ["Cs_opt2"]
    if model == 4:
        return base + (["Cs_opt2", "F_LM", "F_MM", "crap1", "crap2"] if initsc else ["Cs_opt2"])
    if model == 5:
        return base + _LAG
    raise ValueError(f"initial: invalid model number {model}")


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


def split_record(data, names, g, npcon):
    """{name: array} of a record's data bytes holding `names` in order (C-order shapes of
    shape_of; "crap1"/"crap2" are (ld, ny, nz) volumes); the record may hold more
    arrays after them"""
    shapes = [shape_of("u" if n.startswith("crap") else n, g, npcon) for n in names]
    sizes = [int(np.prod(s)) for s in shapes]
    if len(data) < 8 * sum(sizes):
        raise ValueError(f"restart file: {len(data)} bytes, {8 * sum(sizes)} needed for {len(names)} arrays")
    vals = np.frombuffer(data, dtype="<f8", count=sum(sizes))
    out, off = {}, 0
    for n, shp, sz in zip(names, shapes, sizes):
        out[n] = vals[off:off + sz].reshape(shp).copy()
        off += sz
    return out


def pack_arrays(arrays, names, g, npcon):
    """the arrays of `names` in record order, checked against their shapes"""
    out = []
    for n in names:
        a = np.ascontiguousarray(arrays[n], dtype=np.float64)
        if a.shape != shape_of(n, g, npcon):
            raise ValueError(f"restart array {n}: shape {a.shape}, expected {shape_of(n, g, npcon)}")
        out.append(a.reshape(-1))
    return out


def write_restart(lg, path, kind=None, extras=None, max_sub=_MAX_SUB):
    """checkpoint_final_v2 (io.f90:740-786): the context state (and the caller's
    `extras` for the arrays the context does not hold; zeros where absent) as one record
    of layout `kind` (default: the one io.f90:752-760 picks for the context's
    temperature and droplet flags)."""
    g = lg.g
    npc = lg.cfg.drop.npcon
    kind = kind or kind_of(lg.cfg.scal.s_flag, npc > 0)
    extras = extras or {}
    arrays = {}
    for name in LAYOUTS[kind]:
        shp = shape_of(name, g, npc)
        if _held(lg, name):
            if name in ("PCon", "RHS_PCon") and npc == 0:
                raise ValueError(f"restart layout {kind} needs droplet bins")
            arrays[name] = _get(lg, name, shp)
        else:
            arrays[name] = extras[name] if name in extras else np.zeros(shp)
    with open(path, "wb") as f:
        write_record(f, pack_arrays(arrays, LAYOUTS[kind], g, npc), max_sub)


def read_restart(lg, path, kind=None, names=None):
    """initial_multi.f90:48-83: the record's arrays (all of layout `kind`, or its first
    `names`) into the context where it holds them; returns (names set in the context,
    {name: array} of the others)."""
    g = lg.g
    npc = lg.cfg.drop.npcon
    kind = kind or kind_of(lg.cfg.scal.s_flag, npc > 0)
    names = names or LAYOUTS[kind]
    if npc == 0 and ("PCon" in names or "RHS_PCon" in names):
        raise ValueError(f"restart layout {kind} holds droplet bins; the context has none")
    with open(path, "rb") as f:
        arrays = split_record(read_record(f), names, g, npc)
    set_names, extras = [], {}
    for n in names:
        if _held(lg, n):
            _put(lg, n, arrays[n])
            set_names.append(n)
        else:
            extras[n] = arrays[n]
    return set_names, extras


def write_vel(lg, path, max_sub=_MAX_SUB):
    """checkpoint_final_v2 without scalars (io.f90:772-776): the context state as REAL(8)."""
    write_restart(lg, path, "vel", None, max_sub)


def read_vel(lg, path, inilag=False):
    """initial.f90:282-297 (model 5): the record's leading arrays into the context (the
    state that a restarted run starts from); returns the names that were set."""
    names, _ = read_restart(lg, path, "vel", VEL_FIELDS_INILAG if inilag else VEL_FIELDS)
    return names
