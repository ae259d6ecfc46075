"""FWI gradient and RTM image of a survey together, under an HBM budget.

A survey's gradient (exact adjoint, ``fwi.ElasticFWI``) and image (the
reference's back-propagation of the observed data with zero-lag imaging,
``batch.BatchRTM``, backward_modeling.py:539-576 and imaging.py:81-93) are
computed shot group by shot group.  One forward pass keeps checkpoints every L
steps; walking the segments last to first, each segment's forward history is
recomputed once and feeds both backward passes: the FWI adjoint reads it as
its history, and the RTM backward takes the forward wavefield of each step as
the sponge factor times the history's post-injection velocities (exactly the
forward state after the step).  Per shot the job holds the FWI state, adjoint
state, gradient accumulators, blocked-kernel workspace, checkpoints and the
L-step history, plus the RTM backward propagator; L minimises that memory
(about sqrt(nt)) and the group is the largest number of shots whose memory
fits the budget next to the shared medium.  (``share_forward=False`` runs the
two passes one after the other instead, each with its own checkpoints.)

The per-shot coefficient gradients are accumulated in shot order in float32
and the per-shot images in float64, so the result does not depend on the
grouping: any budget gives the bits of one group holding every shot.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np
import torch

from . import configs
from .batch import BatchRTM, ShotSpec, checkpoint_interval
from .field import device
from .fwi import HIST_EVERY, ElasticFWI, accumulate_in_shot_order

_F32 = 4


@dataclass
class Plan:
    group: int        # shots per group
    groups: int
    L_fwi: int        # checkpoint spacing of the FWI pass
    L_rtm: int        # of the RTM pass
    shot_bytes_fwi: int
    shot_bytes_rtm: int
    shared_bytes: int


def _fwi_shot_bytes(cells, nt, nr, L):
    # state 5 + adjoint 5 + gradients 8 + workspace 8 + cell map 2 fields,
    # checkpoints ceil(nt/L) x 5, the blocked kernels' compact history of L
    # steps (velocities 2 per step, stresses 3 per HIST_EVERY steps),
    # records / residuals / data
    fields = 28 + math.ceil(nt / L) * 5 + L * 2 + math.ceil(L / HIST_EVERY) * 3
    return fields * cells * _F32 + 12 * nr * nt * _F32


def _rtm_shot_bytes(cells, nt, nr, L, isnap):
    # forward and backward: state 5 + workspace 8 each, image 3; checkpoints
    # ceil(nt/L) x 5; snapshot buffer (L/isnap + 1) x 2; seismograms / data
    fields = 29 + math.ceil(nt / L) * 5 + (L // isnap + 1) * 2
    return fields * cells * _F32 + 12 * nr * nt * _F32


def _rtm_backward_shot_bytes(cells, nt, nr):
    # backward only (forward wavefield from the FWI history): state 5 +
    # workspace 8 + image 3; observed data / synthetic source
    return 16 * cells * _F32 + 6 * nr * nt * _F32


def plan(nx, nz, shots: Sequence[ShotSpec], budget_bytes, isnap=1, share_forward=True) -> Plan:
    """Group size and checkpoint spacings that fit ``budget_bytes``.  With
    ``share_forward`` the RTM backward pass runs beside the FWI adjoint on the
    same recomputed forward history (one forward recompute for both); else
    the two passes run one after the other, each with its own checkpoints."""
    cells = nx * nz
    nt = max(s.nt for s in shots)
    nr = max(len(np.asarray(s.rec_loc).reshape(-1, 2)) for s in shots)
    # both media (27 fields), image accumulators (6), gradient sum (8), the
    # float64 chain rule (30), margin
    shared = 80 * cells * _F32
    L_fwi = checkpoint_interval(nt, 5, 2 + 3 / HIST_EVERY)
    per_fwi = _fwi_shot_bytes(cells, nt, nr, L_fwi)
    if share_forward:
        L_rtm = L_fwi
        per_rtm = _rtm_backward_shot_bytes(cells, nt, nr)
        per_shot = per_fwi + per_rtm
    else:
        L_rtm = checkpoint_interval(nt, 5, 2, isnap, align=isnap)
        per_rtm = _rtm_shot_bytes(cells, nt, nr, L_rtm, isnap)
        per_shot = max(per_fwi, per_rtm)
    free = budget_bytes - shared
    group = min(len(shots), int(free // per_shot)) if free > 0 else 0
    if group < 1:
        raise MemoryError("one shot needs %.2f GB (FWI) / %.2f GB (RTM) next to %.2f GB shared; "
                          "budget %.2f GB" % (per_fwi / 1e9, per_rtm / 1e9, shared / 1e9,
                                              budget_bytes / 1e9))
    return Plan(group, math.ceil(len(shots) / group), L_fwi, L_rtm, per_fwi, per_rtm, shared)


class GradientAndImage:
    """FWI gradient of J = 1/2 sum |d(m) - d_obs|^2 and the stacked RTM image of
    the observed data, for shots on one (nx, nz) grid, within budget_bytes of
    device memory.  Every shot carries its observed data (u, v, w), each
    (nr, nt)."""

    def __init__(self, nx, nz, dx, dz, fs, absorbing_frame, shots: List[ShotSpec],
                 budget_bytes, isnap=1, kernel="tuned", share_forward=True):
        if any(s.observed is None for s in shots):
            raise ValueError("every shot needs its observed data")
        if share_forward and int(isnap) != 1:
            raise ValueError("sharing the forward history images every step (isnap = 1)")
        self.nx, self.nz, self.dx, self.dz, self.fs = int(nx), int(nz), dx, dz, fs
        self.fw = int(absorbing_frame)
        self.shots = list(shots)
        self.isnap = int(isnap)
        self.kernel = kernel
        self.share_forward = share_forward
        self.plan = plan(self.nx, self.nz, self.shots, budget_bytes, self.isnap, share_forward)

    def prepare(self, vp, vs, rho):
        """Model of the next groups: moduli, device copies and empty
        accumulators (misfit, per-coefficient gradient sum, float64 image)."""
        dev = device()
        vp_n, vs_n, rho_n = (np.asarray(torch.as_tensor(x).cpu(), np.float32) for x in (vp, vs, rho))
        self._mu, self._lam = configs.moduli(vp_n, vs_n, rho_n)
        self._rho = rho_n
        self._m = [torch.as_tensor(x, dtype=torch.float32, device=dev) for x in (vp_n, vs_n, rho_n)]
        self.J = 0.0
        self._gsum = None
        self._image = {c: torch.zeros((self.nx, self.nz), dtype=torch.float64, device=dev)
                       for c in ("u", "v", "w")}
        self._last = None
        # per group: (start event, end event, launches) of every adjoint call
        # when a list (set by the caller before run_group)
        self.adjoint_events = getattr(self, "adjoint_events", None)

    def group_shots(self, k):
        """Shots of group k (in survey order)."""
        g0 = k * self.plan.group
        return self.shots[g0:g0 + self.plan.group]

    def run_group(self, k):
        """Gradient and image of group k, added to the accumulators (shot
        order within and across groups is the caller's: groups in ascending
        order give the bits of one group holding every shot)."""
        dev = device()
        p = self.plan
        group = self.group_shots(k)
        fwi = ElasticFWI(self.nx, self.nz, self.dx, self.dz, self.fs, self.fw,
                         [ShotSpec(s.src_loc, s.wavelets, s.rec_loc) for s in group],
                         precision="float32", checkpoint_every=p.L_fwi,
                         kernel=self.kernel if self.kernel in ("tuned", "blocked") else "auto",
                         history_bytes=0)  # the plan holds checkpoints + one segment
        fwi.adjoint_events = self.adjoint_events
        recs = fwi.forward(*self._m)
        res = []
        for r, s in zip(recs, group):
            obs = torch.stack([torch.as_tensor(o, dtype=torch.float32, device=dev)
                               for o in s.observed])
            d = r - obs
            self.J += 0.5 * float((d.double() ** 2).sum())
            res.append(d)
        del recs
        rkernel = self.kernel if self.kernel in ("tuned", "blocked", "wave") else "blocked"
        if self.share_forward:
            # the RTM backward runs beside the adjoint on each recomputed
            # segment of the forward history
            rtm = BatchRTM(self.nx, self.nz, self.dx, self.dz, self.fs, self._mu, self._lam,
                           self._rho, self.fw, group, isnap=1, imaging="checkpoint",
                           kernel=rkernel)
            rtm.external_begin(fwi.systems, fwi.L, fwi.history_velocity_layout())
            fwi.gradient(res, on_segment=rtm.external_segment)
            codes = rtm.external_codes()
        else:
            fwi.gradient(res)
        self.shapes = dict(fwi.shapes)
        self._gsum = accumulate_in_shot_order(self._gsum, fwi.shot_coefficient_gradients)
        fwi.release()
        self._last = fwi
        if not self.share_forward:
            rtm = BatchRTM(self.nx, self.nz, self.dx, self.dz, self.fs, self._mu, self._lam,
                           self._rho, self.fw, group, isnap=self.isnap, imaging="checkpoint",
                           checkpoint_every=p.L_rtm, kernel=rkernel, snapshot_bytes=0)
            codes = rtm.run()
        if any(codes):
            raise RuntimeError("non-finite wavefield, codes %s" % codes)
        self.rtm_tuned = dict(rtm.tuned)
        for imgs in rtm.image_tensors():
            for c in ("u", "v", "w"):
                if imgs.get(c) is not None:
                    self._image[c] += imgs[c].to(torch.float64)
        del rtm, res

    def result(self):
        """(J, (g_vp, g_vs, g_rho) float64, image {"u", "v", "w"} float64) of
        the groups run since ``prepare``."""
        self.coefficient_gradient = self._gsum
        return self.J, self._last.chain(self._gsum), self._image

    def run(self, vp, vs, rho):
        """(J, (g_vp, g_vs, g_rho) float64, image {"u", "v", "w"} float64)."""
        self.prepare(vp, vs, rho)
        for k in range(self.plan.groups):
            self.run_group(k)
        return self.result()


def salt_survey(n_shots=32, nt=8000, nx=4000, nz=2000, seed=0):
    """The large budgeted workload: the salt model's observed data are made
    by the caller from ``true``; ``start`` is the layered background the
    gradient and image are computed in.  Returns (true, start, shots without
    observed data)."""
    true, start = configs.psv_salt(nx=nx, nz=nz, nt=nt, n_shots=n_shots, seed=seed)
    shots = []
    for s in true.shots:
        shots.append(ShotSpec(np.array([s.src], np.int32), true.wavelets(s),
                              np.asarray(s.rec, np.int32)))
    return true, start, shots


def observed_data(cfg, shots: List[ShotSpec], group=4, on_device=False):
    """Records of ``cfg``'s model for every shot (forward only, one
    checkpoint), (3, nr, nt) per shot: numpy, or float32 device tensors with
    ``on_device``."""
    out = []
    m = [torch.as_tensor(np.asarray(x, np.float32), device=device()) for x in (cfg.vp, cfg.vs, cfg.rho)]
    for g0 in range(0, len(shots), group):
        sub = shots[g0:g0 + group]
        p = ElasticFWI(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, sub, precision="float32",
                       history_bytes=0)
        recs = p.records(*m)
        out += recs if on_device else [r.cpu().numpy() for r in recs]
        del p
    return out
