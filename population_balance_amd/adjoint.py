"""Gradient of an objective of a run of the LES step (with its droplet bins, when the
configuration has them) with respect to its initial velocity and concentrations, by the
discrete adjoint of the frozen-coefficient step (lp_step_adjoint2) with checkpoints.

The frozen-coefficient step is the step with the eddy viscosity Nu_t and the wall u*
taken from the base run (lp_set_frozen); Kc_t = Nu_t / Sc follows Nu_t.  The velocity
part of the map it defines is quadratic and the droplet transport is linear in C on each
branch of the SMART limiter (bilinear in face velocity and face value); lp_step_adjoint2
applies the transpose of its linearisation around a base state.  The gradient is
therefore the exact gradient of the frozen-coefficient run; the sensitivities of Nu_t
(through the dynamic model's Cs^2 and |S|) and of the wall law's |U| are left out by
construction.  Droplet inertia (the acceleration in the face velocity) and the breakup
feedback are not differentiated (lp_step_adjoint2 refuses them).

Checkpointing: the forward run keeps the state every `every` steps.  The reverse sweep
takes the segments last to first: each segment is recomputed from its checkpoint in a
fresh context, keeping the segment's base states (u, v, w and PCon before each step) and
frozen coefficients, and the adjoint steps run backwards through it.  Base states of a
binary32 run are binary64 already (u, v, w and PCon are binary64 in both precisions); the
adjoint always runs in a binary64 context.

Objective (time-integrated probe): J = sum over the steps n = 1..N of the sum of one field
over a box of grid points (a velocity component, or the concentration of every bin),
    J = sum_n sum_{(k, j, i) in box} f^n[k, j, i].
"""
import torch

from . import abi
from .lesgo import Lesgo, LesgoError

VEL = ("u", "v", "w")
# the fields that carry a run from one step to the next (the step's state); the lag
# velocities only with cs_count > 1 (with cs_count = 1 they are zero at every checkpoint,
# as a new context holds them, and the model then reads u, v, w directly)
STATE = ("u", "v", "w", "RHSx", "RHSy", "RHSz", "F_LM", "F_MM", "F_QN", "F_NN", "Cs_opt2", "Beta")
LAG = ("u_lag", "v_lag", "w_lag")
DROPS = ("PCon", "RHS_PCon")


class Probe:
    """sum of field `comp` (0 u, 1 v, 2 w, or "PCon": every droplet bin) over levels
    k0..k0+nk-1, rows j0..j0+nj-1, columns i0..i0+ni-1"""

    def __init__(self, comp, k0, j0, i0, nk=4, nj=4, ni=4):
        self.comp = comp
        self.box = (slice(k0, k0 + nk), slice(j0, j0 + nj), slice(i0, i0 + ni))

    @property
    def on_droplets(self):
        return self.comp == "PCon"

    def value(self, lg):
        if self.on_droplets:
            return float(lg.get("PCon")[(slice(None),) + self.box].sum())
        return float(lg.get(VEL[self.comp])[self.box].sum())

    def seed(self, ub, cb):
        """add dJ_n/d(state) to the adjoints (ub: (3, nz, ny, ld); cb: (npcon, nz, ny, ld))"""
        if self.on_droplets:
            cb[(slice(None),) + self.box] += 1.0
        else:
            ub[self.comp][self.box] += 1.0


class CheckpointedRun:
    """a run of `nsteps` steps of configuration `cfg` (LP_FP64 or LP_MIXED) from a given
    initial velocity (and concentration), with its objective and the gradient of the
    objective by the adjoint"""

    def __init__(self, cfg, nsteps, probe, every=5):
        self.cfg, self.nsteps, self.probe, self.every = cfg, int(nsteps), probe, max(1, int(every))
        self.cfg_adj = abi.copy_config(cfg, precision=abi.LP_FP64)
        self.drops = cfg.drop.npcon > 0
        if probe.on_droplets and not self.drops:
            raise LesgoError("a concentration probe needs droplet bins (npcon > 0)")
        self.checkpoints = {}
        self.J = None
        self.t_forward = self.t_adjoint = None
        self.adjoint_steps = 0

    # ---- forward ----
    def _new(self, cfg):
        return Lesgo(cfg)

    def _save(self, lg):
        names = STATE + (LAG if self.cfg.cs_count > 1 else ()) + (DROPS if self.drops else ())
        return {n: lg.get(n) for n in names}, lg.time()

    def _restore(self, lg, ck):
        fields, t = ck
        for n, a in fields.items():
            lg.put(n, a)
        lg.set_time(*t)

    def _start(self, lg, u0, v0, w0, c0):
        for n, a in zip(VEL, (u0, v0, w0)):
            lg.put(n, a)
        if c0 is not None:
            lg.put("PCon", c0)
        lg.set_time(1, 1)

    def forward(self, u0, v0, w0, c0=None):
        """the run and its objective; keeps the checkpoints"""
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        lg = self._new(self.cfg)
        self._start(lg, u0, v0, w0, c0)
        self.checkpoints = {}
        J = 0.0
        start.record()
        for n in range(self.nsteps):
            if n % self.every == 0:
                self.checkpoints[n] = self._save(lg)
            lg.step(1)
            J += self.probe.value(lg)
        end.record()
        torch.cuda.synchronize()
        self.t_forward = start.elapsed_time(end)
        lg.close()
        self.J = J
        return J

    # ---- reverse ----
    def _segment(self, n0, n1):
        """recompute steps n0..n1-1 from the checkpoint: base u, v, w (and PCon) before
        each step and the step's Nu_t and wall u* (binary64)"""
        lg = self._new(self.cfg)
        self._restore(lg, self.checkpoints[n0])
        base = []
        for n in range(n0, n1):
            uvw = torch.stack([lg.get(m) for m in VEL])
            c = lg.get("PCon") if self.drops else None
            t = lg.time()
            lg.step(1)
            nut = lg.get("Nu_t").to(torch.float64)
            ust = lg.get("ustar").to(torch.float64)
            base.append((uvw, c, t, nut, ust))
        lg.close()
        return base

    def gradient(self):
        """dJ/d(u0, v0, w0) and, with droplets, dJ/d(PCon0) (binary64 tensors of the field
        shapes; zero on the padding)"""
        if self.J is None:
            raise LesgoError("gradient: run forward first")
        g = self.cfg.grid
        shp = abi.field_shape(g, "u")
        ub = torch.zeros((3,) + shp, dtype=torch.float64, device="cuda")
        rb = torch.zeros_like(ub)
        cb = rcb = None
        if self.drops:
            shc = abi.field_shape(g, "PCon", self.cfg.drop.npcon)
            cb = torch.zeros(shc, dtype=torch.float64, device="cuda")
            rcb = torch.zeros_like(cb)
        A = self._new(self.cfg_adj)
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        t_adj = 0.0
        self.adjoint_steps = 0
        starts = sorted(self.checkpoints)
        for si in range(len(starts) - 1, -1, -1):
            n0 = starts[si]
            n1 = starts[si + 1] if si + 1 < len(starts) else self.nsteps
            base = self._segment(n0, n1)
            for n in range(n1 - 1, n0 - 1, -1):
                uvw, c, t, nut, ust = base[n - n0]
                self.probe.seed(ub, cb)  # dJ/d(state^{n+1})
                for q, m in enumerate(VEL):
                    A.put(m, uvw[q])
                if c is not None:
                    A.put("PCon", c)
                A.set_time(*t)
                start.record()
                A.step_adjoint(nut, ust, ub, rb, cb, rcb)
                end.record()
                torch.cuda.synchronize()
                t_adj += start.elapsed_time(end)
                self.adjoint_steps += 1
        A.close()
        self.t_adjoint = t_adj
        out = (ub[0].clone(), ub[1].clone(), ub[2].clone())
        return out + ((cb.clone(),) if self.drops else ())

    def objective_at(self, u0, v0, w0, frozen=None, c0=None):
        """J of a run from (u0, v0, w0[, c0]); with frozen = list of (nut, ust) per step,
        the frozen-coefficient run (the function the adjoint differentiates)"""
        lg = self._new(self.cfg)
        self._start(lg, u0, v0, w0, c0)
        J = 0.0
        for n in range(self.nsteps):
            if frozen is not None:
                nut, ust = frozen[n]
                lg.set_frozen(nut, ust)
            lg.step(1)
            J += self.probe.value(lg)
        lg.set_frozen(None, None)
        lg.close()
        return J

    def frozen_coefficients(self, dtype):
        """Nu_t and u* of every step of the base run, in the run's field type (what
        lp_set_frozen expects)"""
        out = []
        for n0 in sorted(self.checkpoints):
            n1 = min(n0 + self.every, self.nsteps)
            for (_, _, _, nut, ust) in self._segment(n0, n1):
                out.append((nut.to(dtype).contiguous(), ust.to(dtype).contiguous()))
        return out


class RunObjective(torch.autograd.Function):
    """J(u0, v0, w0) of a CheckpointedRun as an autograd function: forward runs the LES,
    backward the checkpointed discrete adjoint.  With droplets, pass c0 as well:
    J(u0, v0, w0, c0)."""

    @staticmethod
    def forward(ctx, u0, v0, w0, run, c0=None):
        J = run.forward(u0.detach(), v0.detach(), w0.detach(), None if c0 is None else c0.detach())
        ctx.run = run
        ctx.with_c = c0 is not None
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, gJ):
        gr = ctx.run.gradient()
        gc = gJ * gr[3] if (ctx.with_c and len(gr) > 3) else None
        return gJ * gr[0], gJ * gr[1], gJ * gr[2], None, gc
