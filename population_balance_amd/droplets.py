"""Droplet-bin parameters (host side): bin diameters, rise/settling velocities and
source strengths for the multi-bin scalar, and the lp_droplet_config they fill.

The reference reads `settling_vel` and `Q_src` from data files that it does not ship
(param.f90:668-674).  Here:
  * diameters: log-spaced (the comment at param.f90:397-398,663) or linear
    (param.f90:658-664) between d_lo and d_hi;
  * terminal velocity: Schiller-Naumann drag, w = w0 / (1 + 0.15 Re^0.687) with
    w0 = |rho_w - rho_d| g d^2 / (18 mu_w) and Re = w d / nu_w, solved by Newton's
    method on f(w) = w (1 + 0.15 Re^0.687) - w0 (monotone in w);
  * source strengths: a seeded log-normal number distribution of diameters
    (SplitMix64 + Box-Muller) histogrammed into the bins.
"""
import ctypes as C
import math

import numpy as np

from . import abi, ic


def diameters(n, d_lo=10e-6, d_hi=2e-3, spacing="log"):
    if spacing == "log":
        return d_lo * (d_hi / d_lo) ** (np.arange(n) / (n - 1))
    return d_lo + (d_hi - d_lo) * np.arange(n) / (n - 1)


def schiller_naumann(d, rho_d=850.0, rho_w=1028.3, mu_w=1.08e-3, g=9.81, tol=1e-14, maxit=100):
    """terminal velocity magnitude (m/s) of a drop of diameter d."""
    nu_w = mu_w / rho_w
    w0 = abs(rho_w - rho_d) * g * d * d / (18.0 * mu_w)
    w = w0
    for _ in range(maxit):
        re = w * d / nu_w
        f = w * (1.0 + 0.15 * re ** 0.687) - w0
        fp = 1.0 + 0.15 * 1.687 * re ** 0.687
        dw = f / fp
        w = max(w - dw, 0.5 * w)
        if abs(dw) <= tol * w0:
            break
    return w


def lognormal_sources(d, q_total=1.0, d50=300e-6, sigma_ln=0.8, seed=ic.DEFAULT_SEED, nsamp=200000):
    """Q_b: number fraction of a seeded log-normal sample falling into each bin,
    times q_total.  Bin edges are the geometric midpoints between diameters."""
    z = ic.normals(seed, nsamp, offset=7)
    samples = d50 * np.exp(sigma_ln * z)
    edges = np.concatenate([[0.0], np.sqrt(d[1:] * d[:-1]), [np.inf]])
    counts, _ = np.histogram(samples, bins=edges)
    return q_total * counts / nsamp


class DropletSet:
    """Keeps the numpy arrays alive for the pointers stored in lp_droplet_config."""

    def __init__(self, npcon=15, spacing="log", d_lo=10e-6, d_hi=2e-3, u_star=1.0, seed=ic.DEFAULT_SEED,
                 q_total=1.0):
        self.diameter = np.ascontiguousarray(diameters(npcon, d_lo, d_hi, spacing))
        self.settling_vel = np.array([schiller_naumann(x) for x in self.diameter]) / u_star
        self.Q_src = np.ascontiguousarray(lognormal_sources(self.diameter, q_total, seed=seed))

    def fill(self, cfg, src=None, pbe=True, active=True, settling=True, inertia=True, bc=(1, 1),
             sigma_cells=3.0, ngau=2, frac=0.7, jt0=1, jt1=10 ** 9, pcon_init=1, sfc_flux=0.0):
        """fill cfg.drop (reference parameters of param.f90 unless overridden)."""
        d = cfg.drop
        g = cfg.grid
        d.npcon = len(self.diameter)
        d.pbe, d.pbe_feedback = int(pbe), 0
        d.settling, d.inertia, d.active = int(settling), int(inertia), int(active)
        d.bc_x, d.bc_y = bc
        si, sj, sk = src if src is not None else (g.nx // 2, g.ny // 2, max(2, (g.nz - 1) // 4))
        d.src_i, d.src_j, d.src_k = si, sj, sk
        d.src_ngau, d.src_jt0, d.src_jt1 = ngau, jt0, jt1
        d.pcon_init, d.init_pcon = pcon_init, 0
        d.src_sigma, d.src_frac = sigma_cells, frac
        d.Sc, d.sfc_flux = 0.4, sfc_flux
        d.rho_c, d.rho_w = 850.0, 1028.3
        d.sigma_s, d.nu_c, d.nu_d, d.k_b, d.d_min = 15.5e-3, 1e-6, 5.85e-6, 0.2, 1e-7
        d.g, d.u_star, d.z_i = 9.81, 1.0, 1.0
        d.diameter = self.diameter.ctypes.data_as(C.POINTER(C.c_double))
        d.settling_vel = self.settling_vel.ctypes.data_as(C.POINTER(C.c_double))
        d.Q_src = self.Q_src.ctypes.data_as(C.POINTER(C.c_double))
        return cfg


def initial_concentration(cfg, drops, amp=1e-3):
    """C_b = amp * Q_b / max(Q) * (Gaussian footprint of the source, sigma in cells, also in z)."""
    g = cfg.grid
    d = cfg.drop
    x = np.arange(g.nx) + 1.0
    y = np.arange(g.ny) + 1.0
    z = np.arange(g.nz) + 1.0
    s = d.src_sigma
    gx = np.exp(-((x - d.src_i) ** 2) / (2 * s * s))
    gy = np.exp(-((y - d.src_j) ** 2) / (2 * s * s))
    gz = np.exp(-((z - d.src_k) ** 2) / (2 * s * s))
    foot = gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    out = np.zeros((d.npcon, g.nz, g.ny, g.ld))
    qn = drops.Q_src / max(drops.Q_src.max(), 1e-300)
    out[..., :g.nx] = amp * (0.5 + qn)[:, None, None, None] * foot[None]
    out[:, -1] = out[:, -2]
    return out
