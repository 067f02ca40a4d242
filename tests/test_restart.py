"""The reference's restart record (io.f90:740-786, initial.f90:282-297) through
population_balance_amd/restart.py: record framing and array order against files written
and read by a Fortran compiler (amdflang), the subrecord form against the module's own
writer, and on the GPU a restart that continues a run bit for bit."""
import os
import shutil
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
