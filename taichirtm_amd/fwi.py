"""FWI gradients through the exact discrete adjoint of the forward scheme.

The reference has no gradient: its backward propagator (backward_modeling.py)
swaps the difference directions and injects data unscaled, so it is the RTM
back-propagator, not the adjoint of its forward step.  Here:

- ``ElasticFWI.forward(vp, vs, rho)`` runs every shot with the forward scheme
  of forward_modeling.py:539-551 (same expressions) and keeps the full state
  every ``checkpoint_every`` steps; it also writes the history (the stresses
  and velocities the updates read) of as many trailing segments as fit
  ``history_bytes`` (default ``"auto"``: half the free device memory when the
  forward pass starts; 0 recomputes every segment), which then need no
  checkpoint;
- ``ElasticFWI.gradient(data_adjoint)`` walks the segments last to first,
  takes each segment's history from the forward pass or recomputes it from the
  segment's checkpoint, runs the adjoint steps over it and accumulates the gradient in
  coefficient space (lam, mu, mxz, myx, myz, inv_rho, inv_rho_u, inv_rho_w),
  then maps it through the harmonic / arithmetic averages to (lam, mu, rho) and
  to (vp, vs, rho) with mu = rho vs^2, lam = rho (vp^2 - 2 vs^2)
  (rtm.py:901-902);
- ``Seismograms.apply(vp, vs, rho, problem)`` is the torch.autograd.Function:
  forward gives the records (n_shots, 3, nr, nt); backward injects the
  incoming gradient as the data adjoint.

``precision="float64"`` runs the same kernels in double, for finite-difference
and dot-product checks of the float32 gradient.  Flat free surface only.

Step kernels (``kernel``): ``"blocked"`` (float32) advances K steps per launch
on LDS tiles with a halo, forward and adjoint, the gradient accumulated in
registers; ``"step"`` runs one step per launch (float32 or float64).  Both
give the same records and gradients bit for bit; ``"auto"`` is ``"blocked"``
with the library's shape per pass in float32 and ``"step"`` in float64;
``"tuned"`` times every compiled blocked shape on a window of this problem's
own steps, per pass (checkpointing forward, history forward, adjoint), and
keeps the fastest.  The blocked kernels keep a compact history: the
post-injection velocities of every step, but the pre-step stresses only of
every K-th step (K = the adjoint's steps per launch, 4 by default); the
adjoint recomputes the others at the start of each launch with the forward
stress update, bit for bit.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._propagator import SPONGE_A, _ptr
from .batch import ShotSpec, checkpoint_interval
from .field import as_tensor, device, free_device_bytes, stream_handle

_REAL = {"float32": (torch.float32, 4), "float64": (torch.float64, 8)}
_KERNELS = {"auto": _lib.KERNEL_AUTO, "step": _lib.KERNEL_SIMPLE, "blocked": _lib.KERNEL_BLOCKED,
            "tuned": _lib.KERNEL_BLOCKED}
_WINDOW = 64  # steps per autotuning trial
# steps between the stored stresses of the compact history of the blocked
# kernels: the blocked adjoint of K steps per launch reads stresses stored every
# min(K, HIST_EVERY) steps
HIST_EVERY = 4
GRAD_NAMES = ("g_lam", "g_mu", "g_mxz", "g_myx", "g_myz", "g_ir", "g_iru", "g_irw")


def accumulate_in_shot_order(total, per_shot):
    """total + per_shot[0] + per_shot[1] + ... per coefficient, in that order
    (total None: start from the first shot), so shots split over several calls
    sum to the same bits as one call."""
    for g in per_shot:
        if total is None:
            total = {k: g[k].clone() for k in GRAD_NAMES}
        else:
            for k in GRAD_NAMES:
                total[k] += g[k]
    return total


def blocked_launches(a, b, K, shots):
    """Kernel launches of one blocked FWI call over steps [a, b) for ``shots``
    shots (the library's schedule): ceil((b - a) / K) launches per group of up
    to 8 shots, one more when that count is odd and the first launch has more
    than one step (it is then split in two, so the state ends in the caller's
    buffers)."""
    n = -(-(b - a) // K)
    if n % 2 == 1 and min(K, b - a) > 1:
        n += 1
    return n * (-(-shots // 8))


def _plan(nx, nz, loc):
    loc = np.ascontiguousarray(loc, np.int32).reshape(-1, 2)
    n = len(loc)
    order = np.zeros(max(n, 1), np.int32)
    start = np.zeros(n + 1, np.int32)
    ng = ctypes.c_int32()
    _lib.call("trtm_plan_injection", nx, nz, n, loc.ctypes.data_as(ctypes.c_void_p),
              order.ctypes.data_as(ctypes.c_void_p), start.ctypes.data_as(ctypes.c_void_p),
              ctypes.byref(ng))
    return (as_tensor(loc, np.int32), as_tensor(order, np.int32),
            as_tensor(start[: ng.value + 1], np.int32), ng.value)


def moduli(vp, vs, rho):
    """mu = rho * vs^2 and lam = ((vp / vs)^2 - 2) * mu, in the inputs' precision,
    with the operation order of rtm.py:901-902."""
    mu = rho * (vs * vs)
    r = vp / vs
    lam = (r * r - 2) * mu
    return mu, lam


class ElasticFWI:
    """Forward records and exact gradients for shots on one (nx, nz) grid."""

    def __init__(self, nx, nz, dx, dz, fs, absorbing_frame, shots: Sequence[ShotSpec],
                 precision="float32", checkpoint_every=None, kernel="auto", tile_x=0, tile_z=0,
                 block_steps=0, history_bytes="auto"):
        if precision not in _REAL:
            raise ValueError("precision must be float32 or float64")
        if kernel not in _KERNELS:
            raise ValueError("kernel must be one of %s" % sorted(_KERNELS))
        if kernel in ("blocked", "tuned") and precision != "float32":
            raise ValueError("the blocked kernels are float32 only")
        self.kernel = kernel
        # blocked kernel shape (tile_x, tile_z, block_steps) per pass: the
        # checkpointing forward, the forward that records the history, the
        # adjoint (0 = library default)
        shape = (int(tile_x), int(tile_z), int(block_steps))
        self.shapes = {"forward": shape, "history": shape, "adjoint": shape, "scattered": shape}
        self.nx, self.nz = int(nx), int(nz)
        self.dx, self.dz, self.fs = float(dx), float(dz), float(fs)
        self.fw = int(absorbing_frame)
        if self.fw < 0 or self.fw > min(nx, nz):
            raise ValueError("absorbing_frame outside [0, min(nx, nz)]")
        self.shots = list(shots)
        self.dtype, self.rb = _REAL[precision]
        self.precision = precision
        dev = device()
        _lib.load()
        dt = 1.0 / self.fs
        self.consts = (dt, 1.0 / self.dx, 1.0 / self.dz, dt * self.dx * self.dz)
        if precision == "float32":
            tab = np.zeros(self.fw + 1, np.float32)
            _lib.call("trtm_sponge_table", self.fw, ctypes.c_float(SPONGE_A),
                      tab.ctypes.data_as(ctypes.c_void_p))
        else:
            k = np.arange(self.fw + 1, dtype=np.float64)
            tab = np.exp((-(SPONGE_A * SPONGE_A) * k) * k)
        self.sponge = torch.as_tensor(tab, device=dev)
        self.geo = []
        for s in self.shots:
            src = _plan(nx, nz, s.src_loc)
            rec = _plan(nx, nz, s.rec_loc)
            wav = [torch.as_tensor(np.asarray(w), dtype=self.dtype, device=dev).reshape(len(src[0]), -1)
                   .contiguous() for w in s.wavelets]
            self.geo.append({"src": src, "rec": rec, "wav": wav, "nt": s.nt,
                             "nr": len(rec[0]), "ns": len(src[0]),
                             "map": torch.empty(2 * nx * nz, dtype=torch.int32, device=dev)})
        self.blocked = kernel in ("blocked", "tuned") or (kernel == "auto" and
                                                          precision == "float32")
        self.tuned = {}  # pass -> (tile_x, tile_z, block_steps) chosen by kernel="tuned"
        for g in self.geo:  # second state buffer of the blocked kernels
            g["ws"] = (torch.empty(8 * nx * nz, dtype=self.dtype, device=dev) if self.blocked
                       else None)
        for g in self.geo:  # cell -> source / receiver group
            d = _lib.FwiDesc()
            d.struct_size = ctypes.sizeof(_lib.FwiDesc)
            d.nx, d.nz = self.nx, self.nz
            loc, order, start, ng = g["src"]
            d.n_src_groups, d.src_loc, d.src_order, d.src_group_start = ng, _ptr(loc), _ptr(order), _ptr(start)
            loc, order, start, ng = g["rec"]
            d.n_rec_groups, d.rec_loc, d.rec_order, d.rec_group_start = ng, _ptr(loc), _ptr(order), _ptr(start)
            d.cell_map = _ptr(g["map"])
            _lib.call("trtm_fwi_prepare", d, stream_handle())
        self.checkpoint_every = checkpoint_every
        # device memory for history written by the forward pass (the trailing
        # segments that fit skip their recompute in ``gradient``)
        if history_bytes != "auto" and int(history_bytes) < 0:
            raise ValueError("history_bytes must be 'auto' or >= 0")
        self.history_bytes = history_bytes if history_bytes == "auto" else int(history_bytes)
        self._ck = None
        self._kept = []
        self.kept_segments = 0
        # bumped by every forward(): a gradient belongs to the last forward
        self.generation = 0
        # adjoint_events: when a list, gradient() appends (start event, end
        # event, kernel launches) around every adjoint call, on the launch stream
        self.adjoint_events = None

    # -- medium ---------------------------------------------------------------

    def _medium(self, vp, vs, rho, dtype=None):
        """Moduli, averages and reciprocals of (vp, vs, rho) in ``dtype``
        (default: the problem's precision)."""
        nx, nz = self.nx, self.nz
        dtype = dtype or self.dtype
        vp = vp.to(device=device(), dtype=dtype).contiguous()
        vs = vs.to(device=device(), dtype=dtype).contiguous()
        rho = rho.to(device=device(), dtype=dtype).contiguous()
        mu, lam = moduli(vp, vs, rho)
        mu, lam = mu.contiguous(), lam.contiguous()
        c = {k: torch.empty((nx, nz), dtype=dtype, device=vp.device)
             for k in ("myx", "myz", "mxz", "rho_u", "rho_w", "inv_rho", "inv_rho_u", "inv_rho_w")}
        fn = "trtm_prepare_medium" if dtype == torch.float32 else "trtm_prepare_medium_f64"
        _lib.call(fn, nx, nz, _ptr(mu), _ptr(rho), _ptr(c["myx"]), _ptr(c["myz"]), _ptr(c["mxz"]),
                  _ptr(c["rho_u"]), _ptr(c["rho_w"]), _ptr(c["inv_rho"]), _ptr(c["inv_rho_u"]),
                  _ptr(c["inv_rho_w"]), stream_handle())
        c.update(mu=mu, lam=lam, rho=rho, vp=vp, vs=vs)
        for k in ("lam", "mu", "mxz", "inv_rho_u", "inv_rho_w"):
            if not bool(torch.isfinite(c[k][1:-1, 1:-1]).all()):
                raise ValueError("non-finite %s on the interior (in SI units the float32 "
                                 "harmonic mean of mu overflows; use g/cm^3 and kPa)" % k)
        return c

    def _desc(self, med, g):
        d = _lib.FwiDesc()
        d.struct_size = ctypes.sizeof(_lib.FwiDesc)
        d.real_bytes = self.rb
        d.nx, d.nz, d.nt, d.fw = self.nx, self.nz, g["nt"], self.fw
        d.systems = self.systems
        d.dt, d.inv_dx, d.inv_dz, d.src_factor = self.consts
        d.sponge_table = _ptr(self.sponge)
        for k in ("lam", "mu", "mxz", "myx", "myz", "inv_rho", "inv_rho_u", "inv_rho_w"):
            setattr(d, k, _ptr(med[k]))
        loc, order, start, ng = g["src"]
        d.n_src, d.n_src_groups = len(loc), ng
        d.src_loc, d.src_order, d.src_group_start = _ptr(loc), _ptr(order), _ptr(start)
        d.wav_u, d.wav_v, d.wav_w = (_ptr(w) for w in g["wav"])
        loc, order, start, ng = g["rec"]
        d.n_rec, d.n_rec_groups = len(loc), ng
        d.rec_loc, d.rec_order, d.rec_group_start = _ptr(loc), _ptr(order), _ptr(start)
        d.cell_map = _ptr(g["map"])
        d.kernel = _KERNELS[self.kernel]
        d.workspace = _ptr(g["ws"]) if g["ws"] is not None else None
        d.hist_every = getattr(self, "hist_every", 0)
        return d

    def _systems(self):
        psv = any(bool((g["wav"][0] != 0).any()) or bool((g["wav"][2] != 0).any())
                  for g in self.geo)
        sh = any(bool((g["wav"][1] != 0).any()) for g in self.geo)
        return (_lib.PSV if psv else 0) | (_lib.SH if sh else 0) or _lib.PSV

    def _fields(self):
        names = []
        if self.systems & _lib.PSV:
            names += ["u", "w", "sxx", "szz", "sxz"]
        if self.systems & _lib.SH:
            names += ["v", "syx", "syz"]
        return names

    # -- history layout -------------------------------------------------------
    # The blocked kernels keep the compact history: the post-injection
    # velocities of every step, the pre-step stresses of every hist_every-th
    # step (the blocked adjoint's steps per launch; it recomputes the others).
    # The one-step-per-launch kernels keep both of every step (hist_every 0).

    def _choose_hist_every(self):
        # the blocked adjoint of K steps per launch reads stresses stored every
        # min(K, HIST_EVERY) steps
        if not self.blocked:
            return 0
        return min(int(self.shapes["adjoint"][2]) or HIST_EVERY, HIST_EVERY)

    def _hist_counts(self):
        nv = (2 if self.systems & _lib.PSV else 0) + (1 if self.systems & _lib.SH else 0)
        ns = (3 if self.systems & _lib.PSV else 0) + (2 if self.systems & _lib.SH else 0)
        return nv, ns

    def history_fields(self, L):
        """Fields (each nx x nz) of the history of L steps."""
        nv, ns = self._hist_counts()
        if self.hist_every:
            return L * nv + -(-L // self.hist_every) * ns
        return L * (nv + ns)

    def history_fields_per_step(self):
        """Average fields per step of the history (the checkpoint spacing's
        memory model)."""
        nv, ns = self._hist_counts()
        return nv + ns / self.hist_every if self.hist_every else nv + ns

    def history_velocity_layout(self):
        """(stride, {component: offset}) of the post-injection velocities in a
        segment's history buffer: step a + j's component c is field
        j * stride + offset[c]."""
        nv, ns = self._hist_counts()
        psv = bool(self.systems & _lib.PSV)
        if self.hist_every:
            return nv, {"u": 0, "w": 1, "v": 2 if psv else 0}
        return nv + ns, {"u": 3, "w": 4, "v": 7 if psv else 2}

    # -- forward --------------------------------------------------------------

    def _tune(self, name, fn, descs, a, b, reset):
        """kernel="tuned": time every compiled shape on steps [a, b) of this
        problem (at most a 64-step window), keep the fastest for pass
        ``name``, then undo the trial steps with reset()."""
        if self.kernel != "tuned" or name in self.tuned:
            return
        from . import autotune
        if name == "adjoint":
            a = max(a, b - _WINDOW)
        else:
            b = min(b, a + _WINDOW)
        key = ("fwi", name, self.nx, self.nz, self.fw, self.systems, len(descs))

        def trial(shape):
            self._call_batch(fn, descs, a, b, shape[1:])
            reset()

        shapes = autotune.fwi_candidates(self.systems, name)
        if name == "adjoint" and self.hist_every:
            # adjoint shapes that read the history's stress interval
            shapes = [c for c in shapes if min(c[3], HIST_EVERY) == self.hist_every]
            key = key + (self.hist_every,)
        shape = autotune.tune(key, self.systems, trial, shapes=shapes)
        self.tuned[name] = shape[1:]
        self.shapes[name] = shape[1:]
        reset()

    def _call_batch(self, fn, descs, a, b, shape=(0, 0, 0)):
        for d in descs:
            d.tile_x, d.tile_z, d.block_steps = shape
        arr = (_lib.FwiDesc * len(descs))(*descs)
        _lib.call(fn, arr, len(descs), a, b, stream_handle())

    def forward(self, vp, vs, rho, reference=None):
        """Records of every shot, a list of (3, nr, nt) tensors (u, v, w);
        keeps the checkpoints for ``gradient``.  All shots advance together
        (one batched call per checkpoint segment).

        reference: (vp, vs, rho) of the model whose records are the observed
        data.  The records returned are then the residuals d(model) -
        d(reference), propagated as a scattered field beside this model's
        wavefield (trtm_fwi_forward_scattered_batch; float32 blocked kernels),
        so they carry float32 rounding relative to the residual itself, not to
        the records; the checkpoints are this model's, bit for bit, and no
        history is kept (every segment is recomputed by ``gradient``)."""
        if reference is not None:
            return self._forward_scattered(vp, vs, rho, reference)
        dev = device()
        self.generation += 1
        for g in self.geo:  # release() dropped the blocked kernels' second state buffer
            if self.blocked and g["ws"] is None:
                g["ws"] = torch.empty(8 * self.nx * self.nz, dtype=self.dtype, device=dev)
        self.systems = self._systems()
        self.hist_every = self._choose_hist_every()
        self.med = self._medium(vp, vs, rho)
        names = self._fields()
        nt_max = max(g["nt"] for g in self.geo)
        self.L = self.checkpoint_every or checkpoint_interval(nt_max, len(names),
                                                              self.history_fields_per_step())
        shape = (self.nx, self.nz)
        self.edges = list(range(0, nt_max, self.L)) + [nt_max]
        nseg = len(self.edges) - 1
        S = len(self.geo)
        n_hist = self.history_fields(self.L)
        # trailing segments whose history the forward pass keeps
        seg_bytes = n_hist * self.nx * self.nz * self.rb * S
        self._kept = []  # an earlier forward's history is released first
        self._ck = None
        budget = self.history_bytes
        if budget == "auto":
            budget = free_device_bytes() // 2
        self.kept_segments = min(nseg, int(budget) // seg_bytes)
        first_kept = nseg - self.kept_segments
        # every shot's state in one tensor, so a checkpoint is one copy
        self.state_all = torch.zeros((S, len(names)) + shape, dtype=self.dtype, device=dev)
        self.state = [{k: self.state_all[q, m] for m, k in enumerate(names)} for q in range(S)]
        self._ck = torch.empty((first_kept,) + tuple(self.state_all.shape), dtype=self.dtype,
                               device=dev)
        self._kept = [torch.empty((S, n_hist) + shape, dtype=self.dtype, device=dev)
                      for _ in range(self.kept_segments)]
        recs, descs = [], []
        for g, st in zip(self.geo, self.state):
            d = self._desc(self.med, g)
            for k in names:
                setattr(d, k, _ptr(st[k]))
            rec = {c: torch.zeros((g["nt"], g["nr"]), dtype=self.dtype, device=dev)
                   for c in ("u", "v", "w")}
            d.rec_u, d.rec_v, d.rec_w = (_ptr(rec[c]) for c in ("u", "v", "w"))
            recs.append(rec)
            descs.append(d)
        if first_kept > 0:
            self._tune("forward", "trtm_fwi_forward_batch", descs, self.edges[0], self.edges[1],
                       self.state_all.zero_)
        for k in range(nseg):
            a, b = self.edges[k], self.edges[k + 1]
            if k < first_kept:
                self._ck[k].copy_(self.state_all)
                self._call_batch("trtm_fwi_forward_batch", descs, a, b, self.shapes["forward"])
                continue
            hist = self._kept[k - first_kept]
            for q, d in enumerate(descs):
                d.hist, d.hist_base, d.hist_len = _ptr(hist[q]), a, b - a
            if "history" not in self.tuned and self.kernel == "tuned":
                start = self.state_all.clone()
                self._tune("history", "trtm_fwi_forward_batch", descs, a, b,
                           lambda: self.state_all.copy_(start))
                del start
            self._call_batch("trtm_fwi_forward_batch", descs, a, b, self.shapes["history"])
        return [torch.stack([rec[c].t() for c in ("u", "v", "w")]).contiguous() for rec in recs]

    def residual(self, vp, vs, rho, reference):
        """d(model) - d(reference) per shot, (3, nr, nt) each, and the
        checkpoints for ``gradient`` (``forward`` with ``reference``)."""
        return self.forward(vp, vs, rho, reference=reference)

    def _forward_scattered(self, vp, vs, rho, reference):
        if not (self.blocked and self.precision == "float32"):
            raise ValueError("the scattered-field residual runs the float32 blocked kernels "
                             "(precision='float32', kernel 'auto', 'blocked' or 'tuned')")
        dev = device()
        self.generation += 1
        for g in self.geo:
            if g["ws"] is None:
                g["ws"] = torch.empty(8 * self.nx * self.nz, dtype=self.dtype, device=dev)
        self.systems = self._systems()
        self.med = self._medium(vp, vs, rho)
        # the models' coefficient differences, formed in double and rounded
        # once: they are the whole source of the scattered field, so rounding
        # each model's coefficients to float32 first would add noise wherever
        # the two differ by less than a float32 ulp
        m64 = self._medium(vp, vs, rho, torch.float64)
        r64 = self._medium(*(torch.as_tensor(x) for x in reference), dtype=torch.float64)
        dif = {}
        for k in ("lam", "mu", "mxz", "myx", "myz", "inv_rho", "inv_rho_u", "inv_rho_w"):
            a, b = m64[k], r64[k]
            dif[k] = torch.where(a == b, torch.zeros_like(a), a - b).to(self.dtype).contiguous()
        del m64, r64
        self.hist_every = self._choose_hist_every()
        names = self._fields()
        nt_max = max(g["nt"] for g in self.geo)
        self.L = self.checkpoint_every or checkpoint_interval(nt_max, len(names),
                                                              self.history_fields_per_step())
        shape = (self.nx, self.nz)
        self.edges = list(range(0, nt_max, self.L)) + [nt_max]
        nseg = len(self.edges) - 1
        S = len(self.geo)
        self._kept = []
        self._ck = None
        self.kept_segments = 0
        self.state_all = torch.zeros((S, len(names)) + shape, dtype=self.dtype, device=dev)
        self.state = [{k: self.state_all[q, m] for m, k in enumerate(names)} for q in range(S)]
        dstate = torch.zeros_like(self.state_all)
        dws = [torch.empty(8 * self.nx * self.nz, dtype=self.dtype, device=dev) for _ in range(S)]
        self._ck = torch.empty((nseg,) + tuple(self.state_all.shape), dtype=self.dtype, device=dev)
        recs, descs = [], []
        for q, (g, st) in enumerate(zip(self.geo, self.state)):
            d = self._desc(self.med, g)
            for m, k in enumerate(names):
                setattr(d, k, _ptr(st[k]))
                setattr(d, "d" + k, _ptr(dstate[q, m]))
            for k in ("lam", "mu", "mxz", "myx", "myz", "inv_rho", "inv_rho_u", "inv_rho_w"):
                setattr(d, "dif_" + k, _ptr(dif[k]))
            d.scat_workspace = _ptr(dws[q])
            rec = {c: torch.zeros((g["nt"], g["nr"]), dtype=self.dtype, device=dev)
                   for c in ("u", "v", "w")}
            d.rec_u, d.rec_v, d.rec_w = (_ptr(rec[c]) for c in ("u", "v", "w"))
            recs.append(rec)
            descs.append(d)

        def reset():
            self.state_all.zero_()
            dstate.zero_()
        self._tune("scattered", "trtm_fwi_forward_scattered_batch", descs, self.edges[0],
                   self.edges[1], reset)
        for k in range(nseg):
            a, b = self.edges[k], self.edges[k + 1]
            self._ck[k].copy_(self.state_all)
            self._call_batch("trtm_fwi_forward_scattered_batch", descs, a, b,
                             self.shapes["scattered"])
        out = [torch.stack([rec[c].t() for c in ("u", "v", "w")]).contiguous() for rec in recs]
        del dstate, dws, dif
        return out

    # -- gradient ---------------------------------------------------------------

    def gradient(self, data_adjoint: List[torch.Tensor], on_segment=None):
        """Gradient of J with respect to (vp, vs, rho), float64 (nx, nz) each,
        for dJ/d(records) given per shot as (3, nr, nt).  Segments last to
        first: every shot's history is taken from the forward pass or
        recomputed from its checkpoint, and the adjoint steps run over it, all
        shots in one batched call each; each
        shot accumulates its own coefficient-space gradient, summed in shot
        order at the end.  ``on_segment(a, b, hists)``, when given, is called
        after the adjoint of steps [a, b) with each shot's history buffer of
        those steps, (fields, nx, nz): the post-injection velocities there
        (at the fields ``history_velocity_layout`` gives) times the sponge are
        the forward wavefield after each step."""
        if self._ck is None:
            raise RuntimeError("call forward first")
        if self.kept_segments and not self._kept:
            raise RuntimeError("the history kept by the forward pass was used by an earlier "
                               "gradient; run forward again")
        dev = device()
        names = self._fields()
        shape = (self.nx, self.nz)
        gnames = GRAD_NAMES
        n_hist = self.history_fields(self.L)
        keep, descs = [], []
        S = len(self.geo)
        nseg = len(self.edges) - 1
        first_kept = nseg - self.kept_segments
        adj_all = torch.zeros((S, len(names)) + shape, dtype=self.dtype, device=dev)
        grad_all = torch.zeros((S, len(gnames)) + shape, dtype=self.dtype, device=dev)
        # recompute buffer, for the segments the forward pass did not keep
        rebuf = (torch.empty((S, n_hist) + shape, dtype=self.dtype, device=dev)
                 if first_kept > 0 else None)
        for q, (g, st, da) in enumerate(zip(self.geo, self.state, data_adjoint)):
            adj = {"p" + k: adj_all[q, m] for m, k in enumerate(names)}
            grads = {k: grad_all[q, m] for m, k in enumerate(gnames)}
            hist = rebuf[q] if rebuf is not None else None
            da = torch.as_tensor(da, device=dev).to(self.dtype)
            res = [da[c].t().contiguous() for c in range(3)]  # (nt, nr) each
            d = self._desc(self.med, g)
            for k in names:
                setattr(d, k, _ptr(st[k]))
                setattr(d, "p" + k, _ptr(adj["p" + k]))
            for k in gnames:
                setattr(d, k, _ptr(grads[k]))
            d.res_u, d.res_v, d.res_w = (_ptr(r) for r in res)
            keep.append((adj, grads, res))
            descs.append(d)
        for k in range(nseg - 1, -1, -1):
            a, b = self.edges[k], self.edges[k + 1]
            if k >= first_kept:
                hists = list(self._kept[k - first_kept])
            else:
                hists = list(rebuf)
            for d, h in zip(descs, hists):
                d.hist, d.hist_base, d.hist_len = _ptr(h), a, b - a
            if k < first_kept:
                # a shot whose record ends before a skips these steps; its state is unused
                self.state_all.copy_(self._ck[k])
                ck = self._ck[k]
                self._tune("history", "trtm_fwi_forward_batch", descs, a, b,
                           lambda: self.state_all.copy_(ck))
                self._call_batch("trtm_fwi_forward_batch", descs, a, b, self.shapes["history"])

            def reset_adjoint():
                adj_all.zero_()
                grad_all.zero_()
            if k == nseg - 1:
                self._tune("adjoint", "trtm_fwi_adjoint_batch", descs, a, b, reset_adjoint)
            timed = self.adjoint_events is not None and self.blocked and self.shapes["adjoint"][2]
            if timed:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            self._call_batch("trtm_fwi_adjoint_batch", descs, a, b, self.shapes["adjoint"])
            if timed:
                e1.record()
                self.adjoint_events.append(
                    (e0, e1, blocked_launches(a, b, self.shapes["adjoint"][2], S)))
            if on_segment is not None:  # the segment's history is still in place
                on_segment(a, b, hists)
            if k >= first_kept:
                self._kept[k - first_kept] = None  # used: free it
        self._kept = []
        self.shot_coefficient_gradients = [gq for _, gq, _ in keep]
        grads = accumulate_in_shot_order(None, self.shot_coefficient_gradients)
        self.coefficient_gradient = grads
        return self.chain(grads)

    def records(self, vp, vs, rho):
        """Records of every shot, as ``forward`` gives them, without keeping
        checkpoints or history for a gradient."""
        saved = (self.checkpoint_every, self.history_bytes)
        self.checkpoint_every = max(g["nt"] for g in self.geo)
        self.history_bytes = 0
        try:
            return self.forward(vp, vs, rho)
        finally:
            self.checkpoint_every, self.history_bytes = saved
            self.release_forward()

    def release_forward(self):
        """Drop what forward() keeps for a gradient (checkpoints, kept
        history, state); the next gradient() needs a new forward()."""
        self._ck = None
        self._kept = []
        self.state_all = self.state = None

    def release(self):
        """Drop the checkpoints, state, workspaces and per-shot gradients;
        the medium stays for ``chain``."""
        self._ck = self.state_all = self.state = None
        self._kept = []
        self.shot_coefficient_gradients = None
        for g in self.geo:
            g["ws"] = None

    def chain(self, grads):
        """Coefficient-space gradient (dict over lam, mu, mxz, myx, myz, ir,
        iru, irw) to (vp, vs, rho) in float64 for the medium of the last
        forward."""
        shape = (self.nx, self.nz)
        m = self.med
        dbl = lambda t: t.to(torch.float64).contiguous()
        out = [torch.empty(shape, dtype=torch.float64, device=device()) for _ in range(3)]
        args = [dbl(m["mu"]), dbl(m["rho"]), dbl(m["vp"]), dbl(m["vs"])] + \
               [dbl(grads[k]) for k in GRAD_NAMES]
        _lib.call("trtm_fwi_chain_gradient", self.nx, self.nz, *[_ptr(t) for t in args],
                  *[_ptr(t) for t in out], stream_handle())
        return tuple(out)


class Seismograms(torch.autograd.Function):
    """(vp, vs, rho) -> records (n_shots, 3, nr, nt); all shots need the same
    nr and nt.  Backward: the exact adjoint gradient.

    records_from (optional): a float64 ElasticFWI of the same shots.  The
    records returned are then its float64 ones, so a loss formed from them
    (a residual) carries no float32 rounding of the forward runs; ``problem``
    still runs its forward pass for the adjoint's checkpoints and history,
    and backward runs ``problem``'s adjoint on the incoming gradient rounded
    to its precision.

    reference (optional): (vp, vs, rho) of the model the observed data come
    from.  The output is then the residual d(vp, vs, rho) - d(reference),
    propagated as a scattered field (``ElasticFWI.residual``), so a loss
    1/2 |output|^2 is the L2 misfit against those data."""

    @staticmethod
    def forward(ctx, vp, vs, rho, problem: ElasticFWI, records_from=None, reference=None):
        if reference is not None:
            if records_from is not None:
                raise ValueError("give records_from or reference, not both")
            recs = problem.forward(vp.detach(), vs.detach(), rho.detach(),
                                   reference=[torch.as_tensor(x).detach() for x in reference])
        else:
            recs = problem.forward(vp.detach(), vs.detach(), rho.detach())
        out_dtype = vp.dtype
        if records_from is not None:
            if records_from.dtype != torch.float64:
                raise ValueError("records_from must be a float64 problem")
            recs = records_from.records(*(x.detach().to(torch.float64) for x in (vp, vs, rho)))
            out_dtype = torch.float64
        if len({tuple(r.shape) for r in recs}) != 1:
            raise ValueError("autograd records need equal receiver counts and record lengths")
        ctx.problem = problem
        ctx.generation = problem.generation
        ctx.dtypes = (vp.dtype, vs.dtype, rho.dtype)
        return torch.stack(recs).to(out_dtype)

    @staticmethod
    def backward(ctx, grad):
        if ctx.problem.generation != ctx.generation:
            raise RuntimeError("the problem ran forward again after these records were made "
                               "(its checkpoints now belong to another model); apply "
                               "Seismograms again before calling backward")
        g_vp, g_vs, g_rho = ctx.problem.gradient([grad[q] for q in range(grad.shape[0])])
        return (g_vp.to(ctx.dtypes[0]), g_vs.to(ctx.dtypes[1]), g_rho.to(ctx.dtypes[2]), None,
                None, None)


def l2_misfit_and_gradient(problem: ElasticFWI, vp, vs, rho, observed=None,
                           records_from: Optional[ElasticFWI] = None, reference=None):
    """J = 1/2 sum (d - d_obs)^2 over every shot, component, receiver and
    sample, and its gradient (vp, vs, rho) in float64.

    records_from: a float64 ElasticFWI of the same shots.  Its forward pass
    gives the records, so the residual is formed in double and rounded to
    ``problem``'s precision once; ``problem`` still runs its own forward pass
    for the adjoint's checkpoints and history, and the adjoint.  A float32
    problem then keeps its adjoint's accuracy where the misfit is so small
    that float32 rounding of the records would dominate the residual.

    reference: (vp, vs, rho) of the model whose records are the observed data
    (a synthetic inversion).  The residual d(model) - d(reference) is then
    propagated as a scattered field in float32 (``ElasticFWI.residual``):
    no float64 run, and no cancellation between two float32 record sets;
    ``observed`` is not used."""
    if reference is not None:
        res = problem.residual(vp, vs, rho, reference)
        J = 0.5 * sum(float((x.double() ** 2).sum()) for x in res)
        return J, problem.gradient(res)
    if observed is None:
        raise ValueError("observed data or a reference model is needed")
    if records_from is None:
        recs = problem.forward(vp, vs, rho)
        res = [r - torch.as_tensor(o, device=r.device).to(r.dtype) for r, o in zip(recs, observed)]
        J = 0.5 * sum(float((x.double() ** 2).sum()) for x in res)
        return J, problem.gradient(res)
    if records_from.dtype != torch.float64:
        raise ValueError("records_from must be a float64 problem")
    recs = records_from.records(*(torch.as_tensor(x).to(torch.float64) for x in (vp, vs, rho)))
    res64 = [r - torch.as_tensor(o, device=r.device).to(torch.float64)
             for r, o in zip(recs, observed)]
    problem.forward(vp, vs, rho)
    J = 0.5 * sum(float((x ** 2).sum()) for x in res64)
    return J, problem.gradient([x.to(problem.dtype) for x in res64])
