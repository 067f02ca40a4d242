"""Shared set-up for the parity tests: one oracle context and one HIP context on the
same seeded state."""
import numpy as np

from population_balance_amd import abi, ic

STATE_FIELDS = ["u", "v", "w", "RHSx", "RHSy", "RHSz", "F_LM", "F_MM", "F_QN", "F_NN",
                "Cs_opt2", "u_lag", "v_lag", "w_lag"]


def seeded_state(cfg, seed=ic.DEFAULT_SEED, log_law=True):
    u, v, w = ic.turbulent_ic(cfg.grid, seed=seed, z0=cfg.z0, vonk=cfg.vonk, log_law=log_law)
    return {"u": u, "v": v, "w": w}


def random_fields(cfg, names, seed=1):
    rng = np.random.default_rng(seed)
    g = cfg.grid
    out = {}
    for n in names:
        a = rng.standard_normal(abi.field_shape(g, n, cfg.drop.npcon))
        a[..., g.nx:] = 0.0
        out[n] = a
    return out


def load_oracle(o, fields, jt=1):
    for k, v in fields.items():
        o.set_field(k, v)
    o.set_time(jt)


def load_gpu(lg, fields, jt=1):
    for k, v in fields.items():
        lg.upload(k, v)
    lg.set_time(jt)


def smooth_random(cfg, nlev=None, seed=3, kmax=None):
    """band-limited random planes (no Nyquist content), shape (nlev, ny, ld)."""
    g = cfg.grid
    nlev = nlev or g.nz
    rng = np.random.default_rng(seed)
    kmax = kmax or min(g.nx, g.ny) // 3
    spec = np.zeros((nlev, g.ny, g.lh), complex)
    ky = np.fft.fftfreq(g.ny, 1.0 / g.ny)
    for jy in range(g.ny):
        if abs(ky[jy]) > kmax:
            continue
        spec[:, jy, :kmax + 1] = rng.standard_normal((nlev, kmax + 1)) + 1j * rng.standard_normal((nlev, kmax + 1))
    phys = np.fft.irfft2(spec, s=(g.ny, g.nx), axes=(-2, -1)) * g.nx * g.ny / 50.0
    out = np.zeros((nlev, g.ny, g.ld))
    out[..., :g.nx] = phys
    return out
