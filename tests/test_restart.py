"""The reference's restart records (io.f90:740-786, initial.f90:74-140 and 282-297,
initial_multi.f90:48-83) through population_balance_amd/restart.py: record framing,
array order and shapes of all four layouts against files written and read by a Fortran
compiler (amdflang), the subrecord form against the module's own writer, and on the GPU
restarts that continue a run bit for bit (with and without droplet bins)."""
import os
import shutil
import types
import subprocess
import struct

import numpy as np
import pytest

from population_balance_amd import restart

LD, NY, NZ = 6, 4, 3
HERE = os.path.dirname(os.path.abspath(__file__))


def expected():
    i = np.arange(1, LD + 1)[None, None, :]
    j = np.arange(1, NY + 1)[None, :, None]
    k = np.arange(1, NZ + 1)[:, None, None]
    return [(q * 1000 + k * 100 + j * 10 + i).astype(np.float64) for q in range(1, 12)]


def _fc():
    return shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    fc = _fc()
    if fc is None:
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("restart")
    exe = str(d / "restart_probe")
    r = subprocess.run([fc, "-O0", os.path.join(HERE, "restart_probe.f90"), "-o", exe],
                       capture_output=True, text=True, cwd=str(d))
    assert r.returncode == 0, r.stderr
    return exe, d


def test_fortran_record_read(probe):
    exe, d = probe
    f = str(d / "fortran.out")
    out = subprocess.run([exe, "write", f], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "written" in out.stdout
    with open(f, "rb") as fh:
        data = restart.read_record(fh)
        assert fh.read() == b""  # one record, nothing after it
    got = np.frombuffer(data, "<f8")
    assert got.size == 11 * LD * NY * NZ
    for q, a in enumerate(expected()):
        assert np.array_equal(got[q * a.size:(q + 1) * a.size].reshape(NZ, NY, LD), a)


def test_fortran_record_write(probe):
    exe, d = probe
    f = str(d / "python.out")
    with open(f, "wb") as fh:
        restart.write_record(fh, expected())
    out = subprocess.run([exe, "read", f], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "read ok" in out.stdout, out.stdout + out.stderr


# restart_layouts_probe.f90's grid and its layout names
PG = types.SimpleNamespace(ld=6, nx=4, ny=3, nz=2)
PNPC = 2
PKINDS = {"pcon": "vel_pcon", "sc": "vel_sc", "scpcon": "vel_sc_pcon"}


def probe_expected(kind):
    """array q (1-based, record order) holds q*1e5 + p*1e4 + k*1e3 + j*100 + i at the
    Fortran index (i, j, k, p); sgs_t3 is level 1 of a volume, surfaces have k = p = 0"""
    out = {}
    for q, n in enumerate(restart.LAYOUTS[kind], start=1):
        shp = restart.shape_of(n, PG, PNPC)
        if len(shp) == 2:
            j, i = np.meshgrid(np.arange(1, shp[0] + 1), np.arange(1, shp[1] + 1), indexing="ij")
            k = 1 if n == "sgs_t3" else 0
            out[n] = (q * 100000 + k * 1000 + j * 100 + i).astype(np.float64)
        elif len(shp) == 3:
            k, j, i = np.meshgrid(*[np.arange(1, m + 1) for m in shp], indexing="ij")
            out[n] = (q * 100000 + k * 1000 + j * 100 + i).astype(np.float64)
        else:
            p, k, j, i = np.meshgrid(*[np.arange(1, m + 1) for m in shp], indexing="ij")
            out[n] = (q * 100000 + p * 10000 + k * 1000 + j * 100 + i).astype(np.float64)
    return out


@pytest.fixture(scope="module")
def layouts_probe(tmp_path_factory):
    fc = _fc()
    if fc is None:
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("layouts")
    exe = str(d / "restart_layouts_probe")
    r = subprocess.run([fc, "-O0", os.path.join(HERE, "restart_layouts_probe.f90"), "-o", exe],
                       capture_output=True, text=True, cwd=str(d))
    assert r.returncode == 0, r.stderr
    return exe, d


@pytest.mark.parametrize("pk", list(PKINDS))
def test_layout_of_fortran_record(layouts_probe, pk):
    """a record written as checkpoint_final_v2 writes it, split by the module's layout"""
    exe, d = layouts_probe
    kind = PKINDS[pk]
    f = str(d / f"fortran_{pk}.out")
    out = subprocess.run([exe, "write", pk, f], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "written" in out.stdout, out.stdout + out.stderr
    with open(f, "rb") as fh:
        data = restart.read_record(fh)
        assert fh.read() == b""
    want = probe_expected(kind)
    assert len(data) == 8 * sum(a.size for a in want.values())  # no array left over
    got = restart.split_record(data, restart.LAYOUTS[kind], PG, PNPC)
    for n in restart.LAYOUTS[kind]:
        assert np.array_equal(got[n], want[n]), n


@pytest.mark.parametrize("pk", list(PKINDS))
def test_fortran_reads_layout(layouts_probe, pk):
    """a record packed by the module, read as initial_multi.f90 reads it"""
    exe, d = layouts_probe
    kind = PKINDS[pk]
    f = str(d / f"python_{pk}.out")
    with open(f, "wb") as fh:
        restart.write_record(fh, restart.pack_arrays(probe_expected(kind), restart.LAYOUTS[kind], PG, PNPC))
    out = subprocess.run([exe, "read", pk, f], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "read ok" in out.stdout, out.stdout + out.stderr


def test_read_prefixes():
    """initial.f90:89-123 and :276-297: what each SGS model reads from the front of the
    record"""
    vel = restart.LAYOUTS["vel"]
    sc = restart.LAYOUTS["vel_sc"]
    assert restart.read_prefix(5, False, False) == vel
    assert restart.read_prefix(5, True, False) == vel[:7]
    assert restart.read_prefix(1, False, False) == vel[:6]
    assert restart.read_prefix(3, False, False) == vel[:7]
    assert restart.read_prefix(5, False, True) == sc
    assert restart.read_prefix(5, True, True) == sc[:11]
    assert restart.read_prefix(1, False, True) == sc[:10]
    assert restart.read_prefix(4, False, True) == sc[:13] + ["crap1", "crap2"]
    with pytest.raises(ValueError):
        restart.read_prefix(6, False, False)
    assert restart.kind_of(False, True) == "vel_pcon" and restart.kind_of(True, True) == "vel_sc_pcon"


def test_subrecords_roundtrip(tmp_path):
    arrays = expected()
    f = str(tmp_path / "sub.out")
    with open(f, "wb") as fh:
        restart.write_record(fh, arrays, max_sub=100)
    raw = open(f, "rb").read()
    (m0,) = struct.unpack("<i", raw[:4])
    assert m0 == -100  # a continued subrecord
    with open(f, "rb") as fh:
        data = restart.read_record(fh)
    assert data == b"".join(a.tobytes() for a in arrays)


def test_bad_markers(tmp_path):
    f = str(tmp_path / "bad.out")
    with open(f, "wb") as fh:
        fh.write(struct.pack("<i", 16) + b"\0" * 16 + struct.pack("<i", 8))
    with open(f, "rb") as fh, pytest.raises(ValueError):
        restart.read_record(fh)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restart_continues_run(tmp_path, precision):
    """2 steps, restart file, a new context reading it (initial.f90:295-297, inilag off),
    2 more steps == 4 uninterrupted steps, bit for bit (the file holds every state the
    step reads: u, v, w, the AB2 history RHS, Cs_opt2 and the Lagrangian averages)."""
    from population_balance_amd import abi, ic
    from population_balance_amd.lesgo import Lesgo
    prec = abi.LP_FP64 if precision == "fp64" else abi.LP_MIXED
    cfg = abi.make_config(32, 32, 17, precision=prec)
    print(f"\n[case] restart continuation 32x32x17 {precision}", flush=True)
    u, v, w = ic.turbulent_ic(cfg.grid)
    a = Lesgo(cfg)
    b = Lesgo(cfg)
    for lg in (a, b):
        for k, x in (("u", u), ("v", v), ("w", w)):
            lg.upload(k, x)
    a.step(4)
    b.step(2)
    f = str(tmp_path / "vel.out")
    restart.write_vel(b, f)
    jt, jt_total = b.time()
    b.close()
    cfg_r = abi.make_config(32, 32, 17, precision=prec, initu=1, inilag=0)
    c = Lesgo(cfg_r)
    assert restart.read_vel(c, f) == restart.VEL_FIELDS
    c.set_time(jt, jt_total)
    c.step(2)
    for k in ("u", "v", "w", "RHSx", "RHSy", "RHSz", "p", "Cs_opt2", "F_LM", "F_QN"):
        assert np.array_equal(a.download(k), c.download(k)), k
    a.close()
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_droplet_restart_continues_run(tmp_path, precision):
    """with 4 droplet bins: 3 steps, a vel_pcon file (io.f90:762-766), a new context
    reading it (initial_multi.f90:66-70), 2 more steps == 5 uninterrupted steps, bit for
    bit on the velocity and droplet state; the arrays the context does not hold
    (deposition, Real_dep, Kc_t, F_KX..F_XX2) come back as they were written."""
    from population_balance_amd import abi, droplets, ic
    from population_balance_amd.lesgo import Lesgo
    prec = abi.LP_FP64 if precision == "fp64" else abi.LP_MIXED
    print(f"\n[case] droplet restart continuation 32x32x17, 4 bins, {precision}", flush=True)

    def cfg_of():
        cfg = abi.make_config(32, 32, 17, precision=prec)
        ds = droplets.DropletSet(4)
        # inertia off: the accelerations dudt..dwdt that it reads are not in the record
        # (a restarted reference run starts them at zero, as a new context does)
        ds.fill(cfg, bc=(0, 0), inertia=False)
        return cfg, ds
    cfg, ds = cfg_of()
    g = cfg.grid
    u, v, w = ic.turbulent_ic(g)
    C0 = droplets.initial_concentration(cfg, ds)
    C0[..., g.nx:] = 0.0
    a, b = Lesgo(cfg), Lesgo(cfg)
    for lg in (a, b):
        for k, x in (("u", u), ("v", v), ("w", w)):
            lg.upload(k, x)
        lg.upload("PCon", C0)
    a.step(5)
    b.step(3)
    rng = np.random.default_rng(2)
    extras = {n: rng.standard_normal(restart.shape_of(n, g, 4)) for n in ("deposition", "Real_dep", "F_KX")}
    f = str(tmp_path / "vel_pcon.out")
    restart.write_restart(b, f, extras=extras)
    jt, jt_total = b.time()
    b.close()
    cfg_r, ds_r = cfg_of()
    c = Lesgo(cfg_r)
    names, back = restart.read_restart(c, f)
    assert names == ["u", "v", "w", "PCon", "RHSx", "RHSy", "RHSz", "RHS_PCon"] + restart.VEL_FIELDS[6:]
    for n, x in extras.items():
        assert np.array_equal(back[n], x), n
    assert not np.any(back["F_XX2"]) and not np.any(back["Kc_t"])
    c.set_time(jt, jt_total)
    c.step(2)
    for k in ("u", "v", "w", "RHSx", "RHSz", "p", "Cs_opt2", "F_LM", "PCon", "RHS_PCon"):
        assert np.array_equal(a.download(k), c.download(k)), k
    a.close()
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_theta_restart_continues_run(tmp_path, precision):
    """with the temperature on (an unstable surface flux): 3 steps, a vel_sc file
    (io.f90:767-770), a new context with initsc reading it (initial_multi.f90:71-74), 2
    more steps == 5 uninterrupted steps, bit for bit (theta, RHS_T, sgs_t3(:,:,1) and
    psi_m carry the temperature's state between steps)"""
    from population_balance_amd import abi, ic
    from population_balance_amd.lesgo import Lesgo
    prec = abi.LP_FP64 if precision == "fp64" else abi.LP_MIXED
    print(f"\n[case] theta restart continuation 32x32x17 {precision}", flush=True)

    def cfg_of(initsc):
        cfg = abi.make_config(32, 32, 17, precision=prec, L_z=1.5)
        abi.set_scalar_defaults(cfg.scal, s_flag=1)
        cfg.scal.wt_s, cfg.scal.dTdz_top, cfg.scal.initsc = 0.1, 0.01, initsc
        return cfg
    cfg = cfg_of(0)
    g = cfg.grid
    u, v, w = ic.turbulent_ic(g)
    rng = np.random.default_rng(5)
    th = 0.95 + 1e-3 * rng.standard_normal((g.nz, g.ny, g.ld))
    th[..., g.nx:] = 0.0
    a, b = Lesgo(cfg), Lesgo(cfg)
    for lg in (a, b):
        for k, x in (("u", u), ("v", v), ("w", w), ("theta", th)):
            lg.upload(k, x)
    a.step(5)
    b.step(3)
    f = str(tmp_path / "vel_sc.out")
    restart.write_restart(b, f)
    jt, jt_total = b.time()
    b.close()
    c = Lesgo(cfg_of(1))
    names, extras = restart.read_restart(c, f)
    assert names == restart.LAYOUTS["vel_sc"] and not extras
    c.set_time(jt, jt_total)
    c.step(2)
    for k in ("u", "v", "w", "RHSz", "p", "theta", "RHS_T", "sgs_t3", "psi_m"):
        assert np.array_equal(a.download(k), c.download(k)), k
    a.close()
    c.close()
