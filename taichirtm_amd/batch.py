"""Many-shot RTM on one medium, batched into one launch per time block.

The reference migrates one shot per ForwardModeling / BackwardModeling pair
(rtm.py:935-1000) and stacks the per-shot images afterwards
(imaging.py:57-129).  ``BatchRTM`` runs every shot of a survey through the
blocked step kernel together: the grid's second dimension is the shot, so the
medium is shared and one launch advances all shots by a block of time steps.
Shots may have different record lengths and receiver spreads.

Two imaging modes give the same image bit for bit:

- ``"snapshot"``: the reference's scheme, forward wavefields saved every
  ``isnap`` steps (forward_modeling.py:568-572) and correlated backward
  (backward_modeling.py:556-564).  Memory: nt/isnap wavefields per shot.
- ``"checkpoint"``: the forward pass keeps the full state every L steps;
  the backward pass walks the segments last to first, recomputes the forward
  wavefields of one segment from its checkpoint into a segment-sized buffer,
  then runs the backward steps of that segment with the image accumulated in
  the step kernel.  Memory: nt/L states + L/isnap wavefields per shot, with L
  from a memory budget (about sqrt(nt) by default); cost: one extra forward
  sweep.  Recomputation is deterministic, so the wavefields, and the image,
  are the snapshot mode's exactly.  The forward pass also writes the segment
  buffers of as many trailing segments as fit ``snapshot_bytes`` (default
  ``"auto"``: half the free device memory when the forward pass starts; only
  shots still running in a segment count); those segments need neither
  checkpoint nor recompute.  Either budget is capped at the free memory less
  what the checkpoints and the backward pass's recompute buffers still need.
"""
from __future__ import annotations

import ctypes
import math
import warnings
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._propagator import Medium, PSV_STATE, SH_STATE, _ptr
from .backward_modeling import BackwardModeling
from .field import (Field, SnapshotField, ZeroField, as_host, device, free_device_bytes,
                    stream_handle)
from .forward_modeling import ForwardModeling


@dataclass
class ShotSpec:
    """One shot: source cells (ns, 2), source wavelets (u, v, w) each (ns, nt),
    receiver cells (nr, 2), and optionally observed data (u, v, w) each
    (nr, nt); without observed data the shot migrates its own forward
    seismogram."""
    src_loc: np.ndarray
    wavelets: Sequence[np.ndarray]
    rec_loc: np.ndarray
    observed: Optional[Sequence[np.ndarray]] = None

    @property
    def nt(self):
        return int(np.asarray(self.wavelets[0]).shape[-1])


def _desc_array(descs):
    arr = (_lib.StepDesc * len(descs))()
    for q, d in enumerate(descs):
        ctypes.memmove(ctypes.byref(arr[q]), ctypes.byref(d), ctypes.sizeof(_lib.StepDesc))
    return arr


def observed_arrays(observed):
    """A shot's observed data as float32 arrays, or as the device tensors
    given (data already resident on the device is not copied)."""
    return [o if isinstance(o, torch.Tensor) else np.asarray(o, np.float32) for o in observed]


def checkpoint_interval(nt, n_state, n_snap, isnap=1, align=1, budget_bytes=None, cell_bytes=4,
                        cells=1):
    """Segment length L (a multiple of align) minimising checkpoint storage
    ceil(nt/L)*n_state plus segment buffer (L/isnap + 1)*n_snap fields, or the
    smallest L whose total fits budget_bytes when one is given."""
    def fields(L):
        return math.ceil(nt / L) * n_state + (L // isnap + 1) * n_snap
    best = max(align, int(math.sqrt(nt * n_state * isnap / max(n_snap, 1))))
    best = ((best + align - 1) // align) * align
    cands = sorted({best} | {((L + align - 1) // align) * align for L in
                             (best // 2, best * 2, nt)} - {0})
    L = min(cands, key=fields)
    if budget_bytes is not None and fields(L) * cell_bytes * cells > budget_bytes:
        raise MemoryError("checkpointing needs %.2f GB per shot, budget %.2f GB"
                          % (fields(L) * cell_bytes * cells / 1e9, budget_bytes / 1e9))
    return L


class BatchRTM:
    """Reverse time migration of many shots sharing one model.

    Parameters: grid and sampling (nx, nz, dx, dz, fs), the model (mu_field,
    lam_field, rho_field, optional surface_matrix_field), absorbing_frame, the
    shots, the snapshot interval isnap, imaging ("snapshot" or "checkpoint"),
    checkpoint_every (segment length; default from the memory model),
    stability_check_interval, the temporally blocked kernel ("blocked": LDS
    tiles, "wave": one wavefront per tile with the state in registers, or
    "tuned": family and shape chosen by the autotuner on this batch) and its
    tile_x / tile_z / block_steps (0 = library default), snapshot_bytes
    (checkpoint mode: device memory for segment wavefields kept from the
    forward pass), and graph (every batched step range recorded into a HIP
    graph and launched as one; same results).
    """

    def __init__(self, nx, nz, dx, dz, fs, mu_field, lam_field, rho_field, absorbing_frame,
                 shots: List[ShotSpec], isnap=1, imaging="snapshot", checkpoint_every=None,
                 stability_check_interval=100, surface_matrix_field=None, tile_x=0, tile_z=0,
                 block_steps=0, kernel="blocked", snapshot_bytes="auto", graph=False):
        if imaging not in ("snapshot", "checkpoint"):
            raise ValueError("imaging must be 'snapshot' or 'checkpoint'")
        if kernel not in ("blocked", "wave", "tuned"):
            raise ValueError("kernel must be 'blocked', 'wave' or 'tuned'")
        if not shots:
            raise ValueError("no shots")
        self.medium = Medium(nx, nz, dx, dz, fs, absorbing_frame, mu_field, lam_field, rho_field,
                             surface_matrix_field)
        self.nx, self.nz = int(nx), int(nz)
        self.shots = shots
        self.isnap = int(isnap)
        self.imaging = imaging
        self.checkpoint_every = checkpoint_every
        self.sci = int(stability_check_interval)
        self.block = dict(tile_x=tile_x, tile_z=tile_z, block_steps=block_steps, kernel=kernel)
        # checkpoint mode: device memory for segment snapshots written by the
        # forward pass (the trailing segments that fit need no recompute)
        if snapshot_bytes != "auto" and int(snapshot_bytes) < 0:
            raise ValueError("snapshot_bytes must be 'auto' or >= 0")
        self.snapshot_bytes = snapshot_bytes if snapshot_bytes == "auto" else int(snapshot_bytes)
        self.kept = {}
        self.first_kept = None
        # every batched step range recorded into a HIP graph and launched as
        # one (taichirtm_amd.plan.StepPlan), released at the end of run()
        self.graph = bool(graph)
        self._plans = []
        self.tuned = {}
        self.codes = []
        self.fw: List[ForwardModeling] = []
        self.bw: List[BackwardModeling] = []

    # -- helpers ------------------------------------------------------------

    def _common(self, shot):
        src = np.asarray(shot.src_loc, np.int32).reshape(-1, 2)
        rec = np.asarray(shot.rec_loc, np.int32).reshape(-1, 2)
        return dict(nx=self.nx, nz=self.nz, dx=self.medium.dx, dz=self.medium.dz, nt=shot.nt,
                    fs=self.medium.fs, absorbing_frame=self.medium.fw, medium=self.medium,
                    src_loc_field=src, recv_loc_field=rec, num_sources=len(src),
                    num_receivers=len(rec), **self.block)

    def _tune(self, name, fn, descs, it_begin, it_end, extra, reset):
        """kernel="tuned": time every compiled shape on a window of this
        batch's own steps, keep the fastest for every descriptor, then undo
        the trial steps with reset()."""
        if self.block.get("kernel") != "tuned":
            return
        from . import autotune
        from ._propagator import set_shape
        key = ("batch", name, self.nx, self.nz, self.medium.fw, self.systems, len(descs),
               self.imaging, self.isnap)
        hi = min(it_end, it_begin + 64)

        def trial(shape):
            for d in descs:
                set_shape(d, shape)
            self._batch(fn, descs, it_begin, hi, *extra)

        shape = autotune.tune(key, self.systems, trial)
        for d in descs:
            set_shape(d, shape)
        self.tuned[name] = shape
        reset()

    def _batch(self, fn, descs, it_begin, it_end, *extra):
        if self.graph:
            from .plan import StepPlan
            fwd = fn == "trtm_forward_steps_batch"
            plan = StepPlan(descs, it_begin, it_end, forward=fwd, save=fwd and bool(extra[0]))
            plan.launch()
            self._plans.append(plan)
            return plan._bws
        arr = _desc_array(descs)
        n = len(descs)
        nbytes = int(_lib.load().trtm_batch_workspace_bytes(n))
        bws = torch.empty(nbytes, dtype=torch.uint8, device=device())
        _lib.call(fn, arr, n, _ptr(bws), nbytes, it_begin, it_end, *extra, stream_handle())
        return bws

    # -- forward ------------------------------------------------------------

    def _forward(self):
        self.fw = [ForwardModeling(wavelet_u_field=s.wavelets[0], wavelet_v_field=s.wavelets[1],
                                   wavelet_w_field=s.wavelets[2], isnap=self.isnap,
                                   **self._common(s)) for s in self.shots]
        systems = {f.systems for f in self.fw}
        if len(systems) != 1:
            raise ValueError("all shots of a batch must excite the same wave systems")
        self.systems = systems.pop()
        dev = device()
        descs, keep = [], []
        self.flags_f = []
        nt_max = max(s.nt for s in self.shots)
        snapshot = self.imaging == "snapshot"
        names = [k for k in PSV_STATE if self.systems & _lib.PSV] + \
                [k for k in SH_STATE if self.systems & _lib.SH]
        self._state_names = names
        # every shot's state in one (shots, fields, nx, nz) tensor, so saving
        # or restoring a checkpoint is one copy (or gather) for the batch
        self._fstate = torch.zeros((len(self.fw), len(names), self.nx, self.nz),
                                   dtype=torch.float32, device=dev)
        for q, f in enumerate(self.fw):
            for m, name in enumerate(names):
                setattr(f, name, Field(tensor=self._fstate[q, m]))
        if not snapshot:
            n_snap = (2 if self.systems & _lib.PSV else 0) + (1 if self.systems & _lib.SH else 0)
            L = self.checkpoint_every or checkpoint_interval(nt_max, len(names), n_snap,
                                                             self.isnap, align=self.isnap)
            if L % self.isnap:
                raise ValueError("checkpoint_every must be a multiple of isnap")
            self.L = L
        for f in self.fw:
            d = f._desc()
            fl = torch.zeros(f.nt // self.sci + 1, dtype=torch.int32, device=dev)
            d.check_interval, d.check_flags = self.sci, _ptr(fl)
            d.isnap = self.isnap
            if snapshot:
                f.num_snaps = f.nt // self.isnap
                for comp, sysbit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV)):
                    fld = (SnapshotField(tensor=torch.zeros((f.num_snaps, self.nx, self.nz),
                                                            dtype=torch.float32, device=dev))
                           if f.systems & sysbit else ZeroField((self.nx, self.nz, f.num_snaps)))
                    setattr(f, comp + "_save_field", fld)
                d.num_snaps = f.num_snaps
                d.snap_u, d.snap_v, d.snap_w = (_ptr(getattr(f, c + "_save_field"))
                                                for c in ("u", "v", "w"))
            d.rec_u, d.rec_v, d.rec_w = (_ptr(f._seis.get(c)) for c in ("u", "v", "w"))
            if not snapshot:  # size the workspace for the recompute's segment buffer too
                d.num_snaps = self.L // self.isnap + 1
            keep.append(f._workspace(d))
            d.num_snaps = f.num_snaps if snapshot else 0
            descs.append(d)
            self.flags_f.append(fl)
        self._fw_descs, self._fw_ws = descs, keep

        def reset_fw():
            for f, fl in zip(self.fw, self.flags_f):
                f.reset_state()
                fl.zero_()

        self._tune("forward", "trtm_forward_steps_batch", descs, 0, nt_max,
                   (1 if snapshot else 0,), reset_fw)
        if snapshot:
            self._batch("trtm_forward_steps_batch", descs, 0, nt_max, 1)
            return
        # checkpoint mode: keep the state at every multiple of L, or, for the
        # trailing segments that fit snapshot_bytes, the segment's snapshots
        L = self.L
        self.edges = list(range(0, nt_max, L)) + [nt_max]
        nseg = len(self.edges) - 1
        nslot = L // self.isnap + 1
        comps = [(c, bit) for c, bit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV))
                 if self.systems & bit]
        seg_shot_bytes = nslot * len(comps) * self.nx * self.nz * 4
        # the budget never takes memory that the rest of the job still needs:
        # the checkpoints of the segments below the first kept one, a
        # recompute segment per shot when any segment is recomputed, and slack
        self.kept, self.ckpt = {}, []  # an earlier run's buffers are released first
        free = free_device_bytes()
        state_shot = len(names) * self.nx * self.nz * 4
        budget = free // 2 if self.snapshot_bytes == "auto" else int(self.snapshot_bytes)
        first_kept, kept_bytes, capped = nseg, 0, False
        for k in range(nseg - 1, -1, -1):
            need = seg_shot_bytes * sum(1 for f in self.fw if self.edges[k] < f.nt)
            if kept_bytes + need > budget:
                break
            ckpt = state_shot * sum(min((f.nt + L - 1) // L, k) for f in self.fw)
            recompute = seg_shot_bytes * len(self.fw) if k > 0 else 0
            if kept_bytes + need + ckpt + recompute + (4 << 30) > free:
                capped = True
                break
            kept_bytes += need
            first_kept = k
        self.first_kept = first_kept
        # bytes of segment wavefields this run keeps; capped: an explicit
        # snapshot_bytes was cut to what the device leaves free
        self.snapshot_budget_used = kept_bytes
        self.snapshot_budget_capped = capped and self.snapshot_bytes != "auto"
        if self.snapshot_budget_capped:
            warnings.warn("snapshot_bytes %d capped at %d bytes (%d of %d segments kept): the "
                          "device has %d bytes free for the job" % (
                              int(self.snapshot_bytes), kept_bytes, nseg - first_kept, nseg,
                              free), RuntimeWarning)
        self.kept = {}
        # checkpoint k holds the shots still running at its first step
        # (ckpt_idx[k]; None: every shot)
        self.ckpt, self.ckpt_idx = [], []
        for k in range(first_kept):
            act = [q for q, f in enumerate(self.fw) if self.edges[k] < f.nt]
            self.ckpt.append(torch.empty((len(act), len(names), self.nx, self.nz),
                                         dtype=torch.float32, device=dev))
            self.ckpt_idx.append(None if len(act) == len(self.fw) else
                                 torch.tensor(act, dtype=torch.int64, device=dev))
        spare = None  # snapshot target of shots whose record ended before the segment
        for k in range(nseg):
            a, b = self.edges[k], self.edges[k + 1]
            if k < first_kept:
                if self.ckpt_idx[k] is None:
                    self.ckpt[k].copy_(self._fstate)
                else:
                    torch.index_select(self._fstate, 0, self.ckpt_idx[k], out=self.ckpt[k])
                self._batch("trtm_forward_steps_batch", descs, a, b, 0)
                continue
            bufs = []
            for f, d in zip(self.fw, descs):
                if a < f.nt:
                    buf = {c: torch.empty((nslot, self.nx, self.nz), dtype=torch.float32,
                                          device=dev) for c, _ in comps}
                else:
                    if spare is None:
                        spare = {c: torch.empty((nslot, self.nx, self.nz), dtype=torch.float32,
                                                device=dev) for c, _ in comps}
                    buf = spare
                bufs.append(buf)
                d.num_snaps, d.snap_base = nslot, a // self.isnap - 1
                d.snap_u, d.snap_v, d.snap_w = (_ptr(buf.get(c)) for c in ("u", "v", "w"))
            self.kept[k] = bufs
            self._batch("trtm_forward_steps_batch", descs, a, b, 1)
        for d in descs:
            d.num_snaps, d.snap_base = 0, 0
            d.snap_u = d.snap_v = d.snap_w = None

    # -- backward -----------------------------------------------------------

    def _backward_propagators(self):
        """The backward propagators (their set-up reads the data on the host
        side, which waits for the device): built before the forward pass is
        queued when every shot carries its observed data, so the device does
        not idle between the two passes."""
        dev = device()
        self.bw = []
        for q, s in enumerate(self.shots):
            if s.observed is not None:
                obs = observed_arrays(s.observed)
            else:
                f = self.fw[q]
                seis = f.seismogram_tensors()
                obs = []
                for c in ("u", "v", "w"):
                    t = seis.get(c)
                    if t is None:
                        obs.append(np.zeros((f.num_receivers, f.nt), np.float32))
                    else:
                        o = torch.empty((f.num_receivers, f.nt), dtype=torch.float32, device=dev)
                        _lib.call("trtm_transpose", f.nt, f.num_receivers, _ptr(t), _ptr(o),
                                  stream_handle())
                        obs.append(o)
            self.bw.append(BackwardModeling(obsdata_u_field=obs[0], obsdata_v_field=obs[1],
                                            obsdata_w_field=obs[2], **self._common(s)))
        if len({b.systems for b in self.bw}) != 1:
            raise ValueError("the observed data of every shot of a batch must excite the same "
                             "wave systems (a shot without data migrates to a zero image on "
                             "its own)")

    def _backward(self):
        dev = device()
        if self.bw is None:
            self._backward_propagators()
        self.flags_b = []
        descs, keep, self._isn = [], [], []
        nt_max = max(s.nt for s in self.shots)
        snapshot = self.imaging == "snapshot"
        for q, (f, b) in enumerate(zip(self.fw, self.bw)):
            d = b._desc()
            fl = torch.zeros(b.nt // self.sci + 1, dtype=torch.int32, device=dev)
            d.check_interval, d.check_flags = self.sci, _ptr(fl)
            d.isnap = self.isnap
            d.image_u, d.image_v, d.image_w = (_ptr(b._img.get(c)) for c in ("u", "v", "w"))
            if snapshot:
                isn = np.array([(k + 1) * self.isnap for k in range(f.num_snaps)], np.int32)
                isn[isn >= f.nt] = 0  # a slot the forward never wrote
                self._isn.append(isn)
                d.num_snaps = f.num_snaps
                d.snap_u, d.snap_v, d.snap_w = (_ptr(getattr(f, c + "_save_field"))
                                                for c in ("u", "v", "w"))
                d.isnaps_host = isn.ctypes.data_as(ctypes.c_void_p)
            else:  # size the workspace for the segment buffer's isnaps
                d.num_snaps = self.L // self.isnap + 1
            keep.append(b._workspace(d))
            descs.append(d)
            self.flags_b.append(fl)
        self._bw_ws = keep

        def reset_bw():
            for b, fl in zip(self.bw, self.flags_b):
                b.reset_state()
                fl.zero_()

        if snapshot:
            self._tune("backward", "trtm_backward_steps_batch", descs, 0, nt_max, (), reset_bw)
            self._batch("trtm_backward_steps_batch", descs, 0, nt_max)
            return
        # checkpoint mode: segments last to first
        L = self.L
        nslot = L // self.isnap + 1
        bufs = []
        if self.first_kept > 0:  # recompute buffers
            for f in self.fw:
                bufs.append({c: (torch.zeros((nslot, self.nx, self.nz), dtype=torch.float32,
                                             device=dev) if f.systems & bit else None)
                             for c, bit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV))})
        else:
            bufs = [{} for _ in self.fw]
        fdescs = []
        for f, d0, buf in zip(self.fw, self._fw_descs, bufs):
            d = _lib.StepDesc.from_buffer_copy(d0)
            d.num_snaps = nslot
            d.snap_u, d.snap_v, d.snap_w = (_ptr(buf.get(c)) for c in ("u", "v", "w"))
            d.rec_u = d.rec_v = d.rec_w = None
            fdescs.append(d)
        for k in range(len(self.edges) - 2, -1, -1):
            a, b = self.edges[k], self.edges[k + 1]
            base = a // self.isnap - 1
            if k >= self.first_kept:
                seg = self.kept.pop(k)  # written by the forward pass
            else:
                seg = bufs
                if self.ckpt_idx[k] is None:
                    self._fstate.copy_(self.ckpt[k])
                else:
                    self._fstate.index_copy_(0, self.ckpt_idx[k], self.ckpt[k])
                self.ckpt[k] = None  # used once: free it
                for q in range(len(self.fw)):
                    fdescs[q].snap_base = base
                self._batch("trtm_forward_steps_batch", fdescs, a, b, 1)
            isns = []
            for q, f in enumerate(self.fw):
                isn = np.full(nslot, -1, np.int32)
                for j in range(nslot):
                    it = (j + base + 1) * self.isnap
                    if a <= it < min(b, f.nt) and it >= 1:
                        isn[j] = it
                isns.append(isn)
                descs[q].snap_base = base
                descs[q].num_snaps = nslot
                descs[q].snap_u, descs[q].snap_v, descs[q].snap_w = (
                    _ptr(seg[q].get(c)) for c in ("u", "v", "w"))
                descs[q].isnaps_host = isn.ctypes.data_as(ctypes.c_void_p)
                descs[q].step_offset = f.nt
            if k == len(self.edges) - 2:
                self._tune("backward", "trtm_backward_steps_batch", descs, -b, -a, (), reset_bw)
            self._batch("trtm_backward_steps_batch", descs, -b, -a)

    # -- backward fed with forward wavefields from elsewhere ------------------

    def external_begin(self, fwd_systems, L, layout=None):
        """Prepare the backward passes for ``external_segment``: the forward
        wavefield comes segment by segment from an FWI history (the FWI
        forward is the reference's forward step, forward_modeling.py:539-551,
        so A times its post-injection velocity is the forward state after each
        step).  ``layout`` = (stride, {component: offset}) of those velocities
        in a history buffer (``ElasticFWI.history_velocity_layout``; default:
        the full layout, 5 P-SV + 3 SH fields per step).  No forward
        propagators, checkpoints or snapshot buffers are held here; imaging
        every step (isnap 1), blocked kernel."""
        if self.isnap != 1:
            raise ValueError("external forward wavefields need isnap = 1")
        if any(s.observed is None for s in self.shots):
            raise ValueError("external forward wavefields need every shot's observed data")
        self.block["kernel"] = "blocked" if self.block["kernel"] == "wave" else self.block["kernel"]
        if layout is None:
            layout = ((5 if fwd_systems & _lib.PSV else 0) + (3 if fwd_systems & _lib.SH else 0),
                      {"u": 3, "w": 4, "v": 7 if fwd_systems & _lib.PSV else 2})
        self._ext_nhist, self._ext_fields = int(layout[0]), dict(layout[1])
        self.L = int(L)
        dev = device()
        self.bw, self.flags_b, descs, keep = [], [], [], []
        for s in self.shots:
            obs = observed_arrays(s.observed)
            common = self._common(s)
            if common["kernel"] == "tuned":
                common["kernel"] = "blocked"
            b = BackwardModeling(obsdata_u_field=obs[0], obsdata_v_field=obs[1],
                                 obsdata_w_field=obs[2], **common)
            if b.systems & ~fwd_systems:
                raise ValueError("observed data drive a wave system the forward history lacks")
            d = b._desc()
            d.kernel = _lib.KERNEL_BLOCKED
            fl = torch.zeros(b.nt // self.sci + 1, dtype=torch.int32, device=dev)
            d.check_interval, d.check_flags = self.sci, _ptr(fl)
            d.isnap = 1
            d.image_u, d.image_v, d.image_w = (_ptr(b._img.get(c)) for c in ("u", "v", "w"))
            d.num_snaps = self.L  # workspace sizing
            keep.append(b._workspace(d))
            d.snap_sponge = 1
            self.bw.append(b)
            self.flags_b.append(fl)
            descs.append(d)
        self.systems = {b.systems for b in self.bw}.pop() if len({b.systems for b in self.bw}) == 1 \
            else None
        if self.systems is None:
            raise ValueError("all shots of a batch must excite the same wave systems")
        self._ext_descs, self._bw_ws, self._ext_isn = descs, keep, []
        self._ext_tuned = False

    def external_segment(self, a, b, hists):
        """Backward steps of forward times b-1 down to a for every shot, the
        image taking the forward wavefield F(t) = sponge x the post-injection
        velocities of step t in the shot's history buffer (at the layout given
        to ``external_begin``)."""
        n = self.nx * self.nz
        nslot = b - a
        if nslot > self.L:
            raise ValueError("segment longer than the history")
        self._ext_isn = []
        descs = self._ext_descs
        for q, (s, d, h) in enumerate(zip(self.shots, descs, hists)):
            isn = np.full(self.L, -1, np.int32)
            for j in range(nslot):
                t = a + j
                if 1 <= t < s.nt:
                    isn[j] = t
            self._ext_isn.append(isn)
            base = h.data_ptr()
            d.snap_base = a - 1
            d.num_snaps = self.L
            d.snap_stride = self._ext_nhist * n
            d.snap_u, d.snap_v, d.snap_w = (
                base + 4 * self._ext_fields[c] * n if self.systems & bit else None
                for c, bit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV)))
            d.isnaps_host = isn.ctypes.data_as(ctypes.c_void_p)
            d.step_offset = s.nt
        if not self._ext_tuned and self.block.get("kernel") == "tuned":
            from . import autotune
            from ._propagator import set_shape
            key = ("batch-external", self.nx, self.nz, self.medium.fw, self.systems, len(descs))
            hi_ = min(b, a + 64)

            def reset():  # the first segment taken starts from rest
                for bw, fl in zip(self.bw, self.flags_b):
                    bw.reset_state()
                    fl.zero_()

            def trial(shape):
                for d in descs:
                    set_shape(d, shape)
                self._batch("trtm_backward_steps_batch", descs, -hi_, -a)
                reset()

            shape = autotune.tune(key, self.systems, trial,
                                  shapes=[c for c in autotune.candidates(self.systems)
                                          if c[0] == "blocked"])
            for d in descs:
                set_shape(d, shape)
            self.tuned["backward"] = shape
            self._ext_tuned = True
        self._batch("trtm_backward_steps_batch", descs, -b, -a)

    def external_codes(self) -> List[int]:
        """Per-shot backward codes (0, or 4-6 with the first failing check step)."""
        self._close_plans()
        out = []
        for fl in self.flags_b:
            h = fl.cpu().numpy()
            nz_ = np.nonzero(h)[0]
            out.append(int(h[nz_[0]]) if len(nz_) else 0)
        return out

    def _close_plans(self):
        for plan in self._plans:
            plan.close()
        self._plans = []

    # -- driver -------------------------------------------------------------

    def run(self) -> List[int]:
        """Forward then backward for every shot; returns per-shot codes
        (0, 1-3 forward, 4-6 backward; the first failing check step)."""
        self.bw = None
        if all(s.observed is not None for s in self.shots):
            self._backward_propagators()
        self._forward()
        self._backward()
        self._close_plans()
        self.codes = []
        for q, s in enumerate(self.shots):
            code = 0
            for fl in (self.flags_f[q], self.flags_b[q]):
                h = fl.cpu().numpy()
                nz_ = np.nonzero(h)[0]
                if len(nz_):
                    code = int(h[nz_[0]])
                    break
            self.codes.append(code)
        return self.codes

    def images(self):
        """Per-shot images [(u, v, w)] as (nx, nz) numpy arrays."""
        return [b.get_results() for b in self.bw]

    def image_tensors(self):
        return [b.image_tensors() for b in self.bw]

    def stacked_image(self, dtype=torch.float64):
        """Sum of the per-shot images on the device, per component, in shot
        order and in ``dtype`` (the reference stacks in float32 on the host,
        imaging.py:86-93)."""
        out = {}
        for c in ("u", "v", "w"):
            acc = None
            for b in self.bw:
                t = b._img.get(c)
                if t is None:
                    continue
                acc = t.to(dtype).clone() if acc is None else acc.add_(t.to(dtype))
            out[c] = acc if acc is not None else torch.zeros((self.nx, self.nz), dtype=dtype,
                                                             device=device())
        return out

    def seismograms(self):
        return [f.get_seismogram() for f in self.fw]
