"""Adjoint sensitivities of a window of advection steps, on the GPU.

The reference has no adjoint; this module differentiates its advection step
(integration.py:1298-1611) as liblacmo_amd computes it, for objectives at the
end of a window of steps started from a given (e.g. spun-up) state:

  J_lwp  domain-mean liquid water path, 1e-18 sum xi m_w / (X dy)  [kg m^-2]
  J_nc   domain-mean cloud droplet number, mean over cells of
         1e-6 m_dry^-1 sum xi I(R)                                  [mg^-1]

with I(R) a smooth stand-in for the cloud class 0.5 um <= R < 25 um of
evaluation.py:319 (the hard class has zero derivative almost everywhere):
I(R) = s((ln R - ln 0.5)/w) s((ln 25 - ln R)/w), s logistic, w = nc_width.

Collisions are fixed events: AdjointWindow.run(record=True) records every
collision event of the window (pair, branch of all_or_nothing.py:577-632,
p_crit of a multiple collection); later runs re-apply exactly those events to
the current state instead of drawing.  The gradient is taken with respect to
the multiplicities xi and dry masses m_s of the SIPs at the window start (by
global SIP id; the other state is held), through the fixed 3-iteration
Newton with its clamps, the latent-heat and vapour feedback of the grid,
advection with the flux limiter, relaxation and the fixed events.

Memory: checkpoints of the state every `checkpoint_every` steps; the reverse
sweep re-runs one segment at a time with a tape (64 B per SIP per substep plus
per-cell Jacobians) and sweeps it backwards.
"""
import ctypes as C
import math

import numpy as np
import torch

from ._native import LM_EVENT_BYTES, LmAdjoint, check

__all__ = ["AdjointWindow", "WindowObjective", "window_objective", "membership_fingerprint", "membership_changes"]


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def collision_steps_per_adv(e):
    """Collision steps of one advection step, as lm_adv_steps places them (integration.py:1515-1568, 1095-1294)."""
    from .integration import col_steps_per_adv
    return col_steps_per_adv(e.no_col_per_adv, e.scale_dt_cond, e.act_collisions)


class AdjointWindow:
    """A window of n_steps advection steps of an Engine, its recorded collision events and checkpoints."""

    def __init__(self, eng, n_steps, checkpoint_every=60, nc_width=0.1, ev_cap=1 << 20):
        self.eng = eng
        self.n_steps = int(n_steps)
        self.K = max(1, min(int(checkpoint_every), self.n_steps))
        self.width = float(nc_width)
        e = eng.e
        dev = eng.device
        self.ncol_step = collision_steps_per_adv(e)
        self.nsub = 2 * int(e.scale_dt_cond)
        self.ncell = eng.n_runs * eng.nc
        n = max(eng.n_tot, 1)
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.ev_cap = int(ev_cap)
        self.ev = torch.zeros(self.ev_cap * LM_EVENT_BYTES // 8, **f64)
        self.ev_cnt = torch.zeros(1, **i32)
        self.ev_count = torch.zeros(self.ncell, **i32)
        self.ev_off = torch.zeros(max(self.ncol_step * self.n_steps, 1) * (self.ncell + 1), **i32)
        self.ev_order = torch.zeros(self.ev_cap, **i32)
        self.run_base = torch.as_tensor(np.asarray(eng.base, dtype=np.int64), device=dev)
        self.slot_of = torch.zeros(n, **i32)
        self.ccd = torch.zeros(self.ncell * 12, **f64)
        self.lam_mw, self.lam_ms, self.lam_xi = (torch.zeros(n, **f64) for _ in range(3))
        self.sl_mw, self.sl_ms, self.sl_xi = (torch.zeros(n, **f64) for _ in range(3))
        self.sl_of = torch.zeros(n, **i32)
        self.lam_cell = torch.zeros(8 * self.ncell, **f64)
        self.tape = None
        self.start = None
        self.recorded = False
        self.ckpt = []
        self.end = None

    # ------------------------------------------------------------ binding
    def _bind(self, mode):
        e = self.eng.e
        e.ev_mode = mode
        e.ev_cap = self.ev_cap
        e.ev = self.ev.data_ptr()
        e.ev_cnt = self.ev_cnt.data_ptr()
        e.ev_count = self.ev_count.data_ptr()
        e.ev_off = self.ev_off.data_ptr()
        e.ev_order = self.ev_order.data_ptr()
        e.ev_nstep = self.ncol_step * self.n_steps
        e.run_base = self.run_base.data_ptr()
        e.slot_of = self.slot_of.data_ptr()
        e.ccd = self.ccd.data_ptr()

    def _bind_tape(self, on):
        e = self.eng.e
        if on:
            t = self.tape
            e.tape_part, e.tape_part_i = t["part"].data_ptr(), t["part_i"].data_ptr()
            e.tape_grid, e.tape_fields = t["grid"].data_ptr(), t["fields"].data_ptr()
            e.tape_nsub, e.tape_ninc = self.K * self.nsub, 2 * self.K
        else:
            e.tape_part = e.tape_part_i = e.tape_grid = e.tape_fields = None
            e.tape_nsub = e.tape_ninc = 0
        e.tape_sub = e.tape_inc = 0

    def _unbind(self):
        e = self.eng.e
        e.ev_mode = 0
        self._bind_tape(False)

    def _alloc_tape(self):
        if self.tape is not None:
            return
        dev = self.eng.device
        n = max(self.eng.n_tot, 1)
        ns = self.K * self.nsub
        self.tape = {
            "part": torch.empty(ns * 7 * n, dtype=torch.float64, device=dev),
            "part_i": torch.empty(ns * 2 * n, dtype=torch.int32, device=dev),
            "grid": torch.empty(ns * 30 * self.ncell, dtype=torch.float64, device=dev),
            "fields": torch.empty(self.K * 2 * 2 * self.ncell, dtype=torch.float64, device=dev),
        }

    def slot_map(self):
        """slot of each global SIP id in the engine's current buffer."""
        self._bind(self.eng.e.ev_mode)
        check(self.eng.L.lm_build_slot_of(C.byref(self.eng.e), self.eng.stream))
        return self.slot_of[: self.eng.n_tot].long()

    # ---------------------------------------------------------- objectives
    def objectives(self):
        """(J_nc, J_lwp) per run of the engine's current state, as float64 arrays (n_runs,).

        The per-cell sums come compensated from the device and are added with math.fsum, so the
        objectives are the correctly rounded sums of their terms."""
        eng = self.eng
        out = torch.zeros(4 * self.ncell, dtype=torch.float64, device=eng.device)
        check(eng.L.lm_objective(C.byref(eng.e), self.width, _ptr(out), eng.stream))
        eng.check()
        o = out.cpu().numpy().reshape(4, eng.n_runs, eng.nc)
        mdinv = eng.buf["scal"][:, 8, :].cpu().numpy()
        e = eng.e
        c_lwp = 1.0E-18 / ((e.x_max - e.x_min) * (e.V_cell / (e.dx * e.dz)))
        J_lwp = np.array([c_lwp * math.fsum(np.concatenate([o[0, r], o[2, r]])) for r in range(eng.n_runs)])
        J_nc = np.array([1.0E-6 * math.fsum(np.concatenate([mdinv[r] * o[1, r], mdinv[r] * o[3, r]])) / eng.nc
                         for r in range(eng.n_runs)])
        return J_nc, J_lwp

    # -------------------------------------------------------------- forward
    def run(self, xi0=None, ms0=None, record=None, membership=None):
        """Runs the window from its start state with xi / m_s of the SIPs replaced (tensors by global id, or
        None to keep them).  The first run records the collision events (record=True); every later run
        re-applies them as fixed events.  Keeps checkpoints every K steps and the end state for gradient().
        membership: a list to which the run appends, after every advection step, the fingerprint of the SIPs'
        cell membership (membership_fingerprint); runs of perturbed inputs whose lists equal the unperturbed
        run's kept every SIP in the same cell at the end of every step.  Returns (J_nc, J_lwp) per run."""
        eng = self.eng
        if self.start is None:
            self.start = eng.snapshot()
        if record is None:
            record = not self.recorded
        if not record and not self.recorded:
            raise RuntimeError("no recorded events: run(record=True) first")
        eng.restore(self.start)
        if xi0 is not None or ms0 is not None:
            slots = self.slot_map()
            cur = int(eng.e.cur)
            if xi0 is not None:
                eng.buf[f"xi{cur}"][slots] = torch.as_tensor(xi0, dtype=torch.float64, device=eng.device)
            if ms0 is not None:
                eng.buf[f"m_s{cur}"][slots] = torch.as_tensor(ms0, dtype=torch.float64, device=eng.device)
        self._bind(1 if record else 2)
        eng.e.ev_step = 0
        if record:
            self.ev_cnt.zero_()
        self.ckpt = []
        try:
            for s0 in range(0, self.n_steps, self.K):
                self.ckpt.append(eng.snapshot())
                if membership is None:
                    eng.adv_steps(min(self.K, self.n_steps - s0))
                else:
                    for _ in range(min(self.K, self.n_steps - s0)):
                        eng.adv_steps(1)
                        membership.append(membership_fingerprint(eng))
            eng.check()
        finally:
            self._unbind()
        if record:
            self.recorded = True
            self.n_events = int(self.ev_cnt.item())
        self.end = eng.snapshot()
        return self.objectives()

    # -------------------------------------------------------------- reverse
    def gradient(self, w_nc=0.0, w_lwp=1.0, verify=False):
        """d(w_nc sum J_nc + w_lwp sum J_lwp)/d(xi, m_s at the window start) by global SIP id, at the inputs
        of the last run.  verify: check that every re-run segment reproduces the next checkpoint bit for bit.
        Leaves the engine at the window's end state."""
        if self.end is None:
            raise RuntimeError("run() the window before asking for its gradient")
        eng, e = self.eng, self.eng.e
        self._alloc_tape()
        A = LmAdjoint(self.lam_mw.data_ptr(), self.lam_ms.data_ptr(), self.lam_xi.data_ptr(),
                      self.lam_cell.data_ptr(), self.sl_mw.data_ptr(), self.sl_ms.data_ptr(), self.sl_xi.data_ptr(),
                      self.sl_of.data_ptr())
        eng.restore(self.end)
        self._bind(0)
        check(eng.L.lm_adj_seed(C.byref(e), C.byref(A), float(w_nc), float(w_lwp), self.width, eng.stream))
        try:
            for seg in reversed(range(len(self.ckpt))):
                s0 = seg * self.K
                ns = min(self.K, self.n_steps - s0)
                eng.restore(self.ckpt[seg])
                self._bind(2)
                e.ev_step = s0 * self.ncol_step
                self._bind_tape(True)
                eng.adv_steps(ns)
                if verify:
                    self._verify(self.ckpt[seg + 1] if seg + 1 < len(self.ckpt) else self.end)
                check(eng.L.lm_adj_adv_steps(C.byref(e), C.byref(A), ns, s0 * self.ncol_step, eng.stream))
                self._bind_tape(False)
            eng.check()
        finally:
            self._unbind()
        eng.restore(self.end)
        n = eng.n_tot
        return self.lam_xi[:n].clone(), self.lam_ms[:n].clone()

    def _verify(self, snap):
        eng = self.eng
        for name, t in snap["buf"].items():
            # migrant-list entries are appended by atomics in any order (readers sort them by id)
            if name in ("err", "mig_slot"):
                continue
            if not torch.equal(eng.buf[name], t):
                raise RuntimeError(f"taped re-run diverged from the recorded window in {name}")


def membership_fingerprint(eng):
    """Two int64 sums over the SIPs of all runs, sum (cell + 7) (id mod 1000 + 1) and sum (cell + 7)
    (id div 1000 + 1), with cell the engine's global cell code (run and cell; negative once inactive) and id
    the SIP's id in its run: equal for two states whose SIPs sit in the same cells (a device tensor; no host
    synchronisation).  One SIP in another cell changes the first sum."""
    cur, ccur = int(eng.e.cur), int(eng.e.ccur)
    n = eng.n_tot
    cell = eng.buf[f"cell{ccur}"][:n].long() + 7
    ids = eng.buf[f"id{cur}"][:n].long()
    return torch.stack([(cell * (ids % 1000 + 1)).sum(), (cell * (ids // 1000 + 1)).sum()])


def membership_changes(a, b):
    """Steps at whose end the fingerprints of two runs (lists from AdjointWindow.run(membership=...)) differ."""
    if len(a) != len(b):
        raise ValueError("fingerprints of windows of different lengths")
    if not a:
        return 0
    return int((torch.stack(a) != torch.stack(b)).any(dim=1).sum())


class WindowObjective(torch.autograd.Function):
    """(sum J_nc, sum J_lwp) of a window as a function of the SIPs' xi and m_s at its start (by global id).

    forward runs the window on the GPU (recording the collision events the first time); backward runs
    the HIP reverse sweep with the incoming weights of the two objectives."""

    @staticmethod
    def forward(ctx, xi0, ms0, window):
        J_nc, J_lwp = window.run(xi0.detach(), ms0.detach())
        ctx.window = window
        return torch.tensor([float(np.sum(J_nc)), float(np.sum(J_lwp))], dtype=torch.float64, device=xi0.device)

    @staticmethod
    def backward(ctx, g):
        g = g.detach().cpu().numpy()
        lam_xi, lam_ms = ctx.window.gradient(w_nc=float(g[0]), w_lwp=float(g[1]))
        return lam_xi, lam_ms, None


def window_objective(xi0, ms0, window):
    """Differentiable (J_nc, J_lwp) of `window` at start multiplicities xi0 and dry masses ms0."""
    return WindowObjective.apply(xi0, ms0, window)
