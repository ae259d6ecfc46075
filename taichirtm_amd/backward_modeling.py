"""Backward (time-reversed) propagation with zero-lag imaging on the MI355X.

Drop-in counterpart of the reference's BackwardModeling
(src/backward_modeling.py:32-645).  The imaging condition accumulates in
float32 inside the step kernel, in the reference's t-descending order, so no
per-step host synchronisation (the reference syncs on _check_isnap_match every
imaging step, backward_modeling.py:562) is needed: snapshot validity is
checked on the host once from isnaps.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib
from ._propagator import Propagator, _ptr
from .field import Field, SnapshotField, ZeroField, as_host, as_tensor, device, stream_handle


class BackwardModeling(Propagator):
    """Backward modeling with cross-correlation imaging.

    Keyword arguments (backward_modeling.py:125-151): nx, nz, dx, dz, nt, fs,
    mu_field, lam_field, rho_field, absorbing_frame (60), src_loc_field,
    obsdata_u_field / obsdata_v_field / obsdata_w_field (num_receivers, nt),
    recv_loc_field, surface_matrix_field (optional), num_sources,
    num_receivers.  Extra keywords: kernel = "auto" | "fused" | "simple";
    record_synthetic_source (False): also fill synsrc_u/v/w, which the
    reference writes every step but never reads (backward_modeling.py:462-470).
    """

    _direction = "backward"

    @staticmethod
    def estimate_memory_bytes(nx: int, nz: int, nt: int, num_sources: int,
                              num_receivers: int) -> int:
        """The reference's memory accounting (backward_modeling.py:81-123):
        20 float32 grid fields plus three (num_sources, nt) synthetic sources."""
        return 20 * nx * nz * 4 + 3 * num_sources * nt * 4

    def __init__(self, **kwargs):
        self._setup(kwargs, ("obsdata_u_field", "obsdata_v_field", "obsdata_w_field"))
        self.obsdata_u_field = kwargs["obsdata_u_field"]
        self.obsdata_v_field = kwargs["obsdata_v_field"]
        self.obsdata_w_field = kwargs["obsdata_w_field"]
        self._alloc_state()
        dev = device()
        nx, nz = self.nx, self.nz
        self._img = {}
        for comp, sysbit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV)):
            if self.systems & sysbit:
                self._img[comp] = torch.zeros((nx, nz), dtype=torch.float32, device=dev)
                setattr(self, "result_" + comp, Field(tensor=self._img[comp]))
            else:
                setattr(self, "result_" + comp, ZeroField((nx, nz)))
        self._syn = {}
        if kwargs.get("record_synthetic_source", False) and self.num_sources:
            for comp, sysbit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV)):
                if self.systems & sysbit:
                    self._syn[comp] = torch.zeros((self.nt, self.num_sources),
                                                  dtype=torch.float32, device=dev)

    def _syn_field(self, comp):
        t = self._syn.get(comp)
        if t is None:
            return ZeroField((self.num_sources, self.nt))
        return Field(tensor=torch.as_tensor(
            self._transpose_to_host(t, self.nt, self.num_sources), device=t.device))

    synsrc_u = property(lambda self: self._syn_field("u"))
    synsrc_v = property(lambda self: self._syn_field("v"))
    synsrc_w = property(lambda self: self._syn_field("w"))

    def _snapshot_tensor(self, f, num_snaps):
        """Snapshot-major (num_snaps, nx, nz) device tensor for a snapshot
        argument in the reference's (nx, nz, num_snaps) layout."""
        if f is None or isinstance(f, ZeroField):
            return None
        if isinstance(f, SnapshotField):
            t = f.tensor
        elif isinstance(f, Field):
            t = f.tensor.permute(2, 0, 1).contiguous()
        else:
            t = as_tensor(f).permute(2, 0, 1).contiguous()
        if tuple(t.shape) != (num_snaps, self.nx, self.nz):
            raise ValueError("snapshots have shape %s, expected (nx, nz, num_snaps) = %s"
                             % ((t.shape[1], t.shape[2], t.shape[0]), (self.nx, self.nz, num_snaps)))
        return t

    def run_calc(self, import_fwdata_u, import_fwdata_v, import_fwdata_w, isnaps_field,
                 num_snaps: int, isnap_interval: int, method: str = "cross_correlation",
                 display_callback: Optional[Callable] = None,
                 stability_check_interval: int = 100) -> int:
        """Run backward modeling with correlation (backward_modeling.py:498-579).

        ``method`` is accepted and ignored, as in the reference (the body never
        reads it).  Returns 0, or 4 / 5 / 6 when u / v / w became non-finite.
        """
        d, flags, ws, sci, sysmask, imaging, snaps, isnaps = self._prepare_calc(  # noqa: F841
            import_fwdata_u, import_fwdata_v, import_fwdata_w, isnaps_field, num_snaps,
            isnap_interval, stability_check_interval)
        nx, nz, nt = self.nx, self.nz, self.nt
        num_snaps, isnap = int(num_snaps), int(isnap_interval)
        if imaging and sysmask != self.systems:
            # a propagating system without forward data: its image stays 0,
            # run it without imaging in a second descriptor
            return self._run_split(d, sysmask, flags, sci, display_callback, ws)

        image_steps = [nt - 1 - t for t in range(nt - 1, 0, -1)
                       if imaging and t % isnap == 0 and 0 <= t // isnap - 1 < num_snaps
                       and isnaps[t // isnap - 1] == t]
        stops = image_steps if display_callback is not None else ()
        s = stream_handle()
        code = 0
        for a, b in self._chunks(nt, sci, stops):
            _lib.call("trtm_backward_steps", d, a, b, s)
            host_flags = flags.cpu().numpy()
            code, _ = self._first_failure(host_flags, sci, a, b - 1)
            if code:
                break
            if display_callback is not None and (b - 1) in stops:
                display_callback(self.u.to_numpy(), self.v.to_numpy(), self.w.to_numpy(),
                                 nt - b, nx, nz, self.dx, self.dz)
            code, _ = self._first_failure(host_flags, sci, b - 1, b)
            if code:
                break
        if code:
            return code
        print("Backward modeling completed")
        return 0

    def capture_calc(self, import_fwdata_u, import_fwdata_v, import_fwdata_w, isnaps_field,
                     num_snaps: int, isnap_interval: int, stability_check_interval: int = 100):
        """run_calc's whole run as a ``CapturedCalc``: recorded once into a HIP
        graph (blocked or wave kernels) reading the given snapshots, replayed
        by ``launch()``; no display callback and no early stop (``result()``
        reads the check flags afterwards, codes 4-6).  Each launch continues
        from the current state and adds to the current image: call
        ``reset_state()`` first for a fresh run.  Snapshots passed as a
        forward propagator's own snapshot fields are read where they live
        (a replayed forward capture refreshes them); other snapshot arrays and
        the isnaps are copied when the run is captured."""
        from .plan import StepPlan
        d, flags, ws, sci, sysmask, imaging, snaps, isnaps = self._prepare_calc(
            import_fwdata_u, import_fwdata_v, import_fwdata_w, isnaps_field, num_snaps,
            isnap_interval, stability_check_interval)
        if imaging and sysmask != self.systems:
            raise NotImplementedError("capture with partially imaged wave systems")
        plan = StepPlan([d], 0, self.nt, forward=False,
                        keep=[ws, flags] + [t for t in snaps.values() if t is not None])
        return CapturedCalc(self, plan, flags, sci)

    def _prepare_calc(self, import_fwdata_u, import_fwdata_v, import_fwdata_w, isnaps_field,
                      num_snaps, isnap_interval, stability_check_interval):
        """Snapshot tensors, descriptor, check flags and workspace of a
        backward run; the autotuner's choice when kernel="tuned"."""
        sci = int(stability_check_interval)
        if sci < 1:
            raise ValueError("stability_check_interval must be >= 1")
        nx, nz, nt = self.nx, self.nz, self.nt
        num_snaps = int(num_snaps)
        isnap = int(isnap_interval)
        if isnap < 1:
            raise ValueError("isnap_interval must be >= 1")
        dev = device()
        snaps = {}
        if num_snaps > 0:
            for comp, f in (("u", import_fwdata_u), ("v", import_fwdata_v), ("w", import_fwdata_w)):
                snaps[comp] = self._snapshot_tensor(f, num_snaps)
            isnaps = np.ascontiguousarray(as_host(isnaps_field, np.int32).reshape(-1)[:num_snaps])
            if len(isnaps) < num_snaps:
                raise ValueError("isnaps_field has fewer than num_snaps entries")
        else:
            isnaps = np.zeros(1, np.int32)
        # a component imaged here whose forward snapshot is identically zero
        # correlates to zero; give the kernel explicit zeros only if the other
        # component of its system has data
        for c1, c2 in (("u", "w"), ("w", "u")):
            if snaps.get(c1) is None and snaps.get(c2) is not None and c1 in self._img:
                snaps[c1] = torch.zeros((num_snaps, nx, nz), dtype=torch.float32, device=dev)
        imaging = num_snaps > 0 and any(
            snaps.get(c) is not None and c in self._img for c in ("u", "v", "w"))

        d = self._desc()
        flags = torch.zeros(nt // sci + 1, dtype=torch.int32, device=dev)
        d.check_interval = sci
        d.check_flags = _ptr(flags)
        d.isnap = isnap
        if imaging:
            # image only the systems whose forward snapshots exist
            sysmask = 0
            if snaps.get("u") is not None and "u" in self._img:
                sysmask |= _lib.PSV
            if snaps.get("v") is not None and "v" in self._img:
                sysmask |= _lib.SH
            d.num_snaps = num_snaps
            d.snap_u, d.snap_v, d.snap_w = (_ptr(snaps.get(c)) for c in ("u", "v", "w"))
            d.isnaps_host = isnaps.ctypes.data_as(ctypes.c_void_p)
            d.image_u, d.image_v, d.image_w = (_ptr(self._img.get(c)) for c in ("u", "v", "w"))
        d.rec_u, d.rec_v, d.rec_w = (_ptr(self._syn.get(c)) for c in ("u", "v", "w"))
        ws = self._workspace(d)
        self._autotune(d, "backward",
                       lambda dd, n: _lib.call("trtm_backward_steps", dd, 0, n, stream_handle()))
        flags.zero_()
        sysmask = sysmask if imaging else 0
        return d, flags, ws, sci, sysmask, imaging, snaps, isnaps

    def _run_split(self, d, sysmask, flags, sci, display_callback, ws):
        """Imaging for the systems in sysmask, plain propagation for the rest."""
        d_img = _lib.StepDesc.from_buffer_copy(d)
        d_img.systems = sysmask
        d_rest = _lib.StepDesc.from_buffer_copy(d)
        d_rest.systems = self.systems & ~sysmask
        d_rest.num_snaps = 0
        if display_callback is not None:
            raise NotImplementedError("display_callback with partially imaged systems")
        s = stream_handle()
        d_img.rec_u = d_img.rec_v = d_img.rec_w = None
        for dd in (d_img, d_rest):
            if dd.systems:
                _lib.call("trtm_backward_steps", dd, 0, self.nt, s)
        code, _ = self._first_failure(flags.cpu().numpy(), sci, 0, self.nt)
        if code:
            return code
        print("Backward modeling completed")
        return 0

    def get_results(self) -> tuple:
        """Images (u, v, w), each (nx, nz) numpy float32 (backward_modeling.py:581-587)."""
        return tuple(getattr(self, "result_" + c).to_numpy() for c in ("u", "v", "w"))

    def image_tensors(self) -> dict:
        return dict(self._img)

    def cleanup(self):
        """Drop every device array this instance holds (backward_modeling.py:589-645)."""
        for name in ("sxx", "sxz", "szz", "syx", "syz", "u", "v", "w", "mxz", "myx", "myz",
                     "rho_u", "rho_w", "inv_rho_u", "inv_rho_w", "inv_rho", "absorb_coeff",
                     "result_u", "result_v", "result_w", "src_loc_field", "recv_loc_field",
                     "obsdata_u_field", "obsdata_v_field", "obsdata_w_field", "surface_matrix",
                     "mu", "lam", "rho_field"):
            if hasattr(self, name):
                setattr(self, name, None)
        self._img = {}
        self._syn = {}
        self._coef = {}
        self._inj = []


class CapturedCalc:
    """A backward run captured by ``BackwardModeling.capture_calc``."""

    def __init__(self, prop, plan, flags, sci):
        self.prop, self.plan, self.flags, self.sci = prop, plan, flags, sci

    def launch(self):
        """Clear the check flags and replay the steps on the current stream."""
        self.flags.zero_()
        self.plan.launch()

    def result(self) -> int:
        """Wait for the last launch; 0, or the code of the first failing check."""
        code, _ = self.prop._first_failure(self.flags.cpu().numpy(), self.sci, 0, self.prop.nt)
        if code:
            return code
        print("Backward modeling completed")
        return 0

    def close(self):
        self.plan.close()
