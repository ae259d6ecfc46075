"""Forward elastic P-SV / SH modeling on the MI355X.

Drop-in counterpart of the reference's ForwardModeling
(src/forward_modeling.py:32-646): same constructor kwargs, ``run`` signature,
return codes, snapshot attributes and ``get_seismogram``.  Each ``run`` hands
whole ranges of time steps to the HIP library (one fused launch per step)
instead of issuing seven Taichi kernels per step from a Python loop.
"""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np
import torch

from . import _lib
from ._propagator import Propagator, _ptr
from .field import Field, SnapshotField, ZeroField, device, stream_handle


class ForwardModeling(Propagator):
    """Forward modeling for seismic wave propagation.

    Keyword arguments (forward_modeling.py:126-153): nx, nz, dx, dz, nt, fs,
    mu_field, lam_field, rho_field, absorbing_frame (60), src_loc_field
    (num_sources, 2), wavelet_u_field / wavelet_v_field / wavelet_w_field
    (num_sources, nt), recv_loc_field (num_receivers, 2), isnap (10),
    surface_matrix_field (optional), num_sources, num_receivers.  Fields may
    be ``Field`` objects, torch tensors or numpy arrays.  Extra keyword:
    kernel = "auto" | "fused" | "simple" (step kernel selection).
    """

    _direction = "forward"

    @staticmethod
    def estimate_memory_bytes(nx: int, nz: int, nt: int, num_sources: int,
                              num_receivers: int) -> int:
        """The reference's memory accounting (forward_modeling.py:81-124):
        17 float32 grid fields plus three (num_receivers, nt) seismograms."""
        return 17 * nx * nz * 4 + 3 * num_receivers * nt * 4

    def __init__(self, **kwargs):
        self.isnap = int(kwargs.get("isnap", 10))
        self._setup(kwargs, ("wavelet_u_field", "wavelet_v_field", "wavelet_w_field"))
        self.wavelet_u_field = kwargs["wavelet_u_field"]
        self.wavelet_v_field = kwargs["wavelet_v_field"]
        self.wavelet_w_field = kwargs["wavelet_w_field"]
        self._alloc_state()
        dev = device()
        nr = self.num_receivers
        # time-major recording buffers (nt, nr); get_seismogram transposes once
        self._seis = {}
        for comp, sysbit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV)):
            if self.systems & sysbit and nr:
                self._seis[comp] = torch.zeros((self.nt, nr), dtype=torch.float32, device=dev)

    # seismogram fields with the reference's (num_receivers, nt) shape
    def _seis_field(self, comp):
        t = self._seis.get(comp)
        if t is None:
            return ZeroField((self.num_receivers, self.nt))
        return Field(tensor=torch.as_tensor(
            self._transpose_to_host(t, self.nt, self.num_receivers), device=t.device))

    seismogram_u = property(lambda self: self._seis_field("u"))
    seismogram_v = property(lambda self: self._seis_field("v"))
    seismogram_w = property(lambda self: self._seis_field("w"))

    def run(self, save: bool = False, display_callback: Optional[Callable] = None,
            stability_check_interval: int = 100) -> int:
        """Run forward modeling (forward_modeling.py:504-575).

        Returns 0 on success or 1 / 2 / 3 when u / v / w became non-finite
        (|x| > 1e30 or NaN) at a check step it % stability_check_interval == 0.
        """
        d, flags, ws, sci = self._prepare_run(save, stability_check_interval)  # noqa: F841
        nx, nz, nt = self.nx, self.nz, self.nt
        s = stream_handle()
        stops = range(0, nt, self.isnap) if display_callback is not None else ()
        code, it_fail = 0, None
        for a, b in self._chunks(nt, sci, stops):
            _lib.call("trtm_forward_steps", d, a, b, int(bool(save)), s)
            host_flags = flags.cpu().numpy()
            code, it_fail = self._first_failure(host_flags, sci, a, b - 1)
            if code:
                break
            # order inside a step: record -> callback -> finite check (:551-566)
            if display_callback is not None and (b - 1) % self.isnap == 0:
                display_callback(self.u.to_numpy(), self.v.to_numpy(), self.w.to_numpy(),
                                 b - 1, nx, nz, self.dx, self.dz)
            code, it_fail = self._first_failure(host_flags, sci, b - 1, b)
            if code:
                break
        return self._finish_run(save, code, it_fail)

    def _prepare_run(self, save, stability_check_interval):
        """Snapshot buffers, descriptor, check flags and workspace of a run;
        the autotuner's choice when kernel="tuned"."""
        sci = int(stability_check_interval)
        if sci < 1:
            raise ValueError("stability_check_interval must be >= 1")
        if self.isnap < 1:
            raise ValueError("isnap must be >= 1")
        nx, nz, nt = self.nx, self.nz, self.nt
        dev = device()
        if save:
            num_snaps = nt // self.isnap
            self.num_snaps = num_snaps
            written = (nt - 1) // self.isnap  # slots 0..written-1 are written by this run
            for comp, sysbit in (("u", _lib.PSV), ("v", _lib.SH), ("w", _lib.PSV)):
                old = getattr(self, comp + "_save_field", None)
                if not self.systems & sysbit:
                    f = ZeroField((nx, nz, num_snaps))
                elif isinstance(old, SnapshotField) and old.shape == (nx, nz, num_snaps):
                    # a repeated run reuses the buffers; every slot it reads is rewritten
                    f = old
                    f.tensor[written:].zero_()
                else:
                    f = SnapshotField(tensor=torch.zeros((num_snaps, nx, nz), dtype=torch.float32,
                                                         device=dev))
                setattr(self, comp + "_save_field", f)
            if not isinstance(getattr(self, "isnaps_field", None), Field) or \
                    self.isnaps_field.shape != (num_snaps,):
                self.isnaps_field = Field(np.int32, (num_snaps,))
        d = self._desc()
        flags = torch.zeros(nt // sci + 1, dtype=torch.int32, device=dev)
        d.check_interval = sci
        d.check_flags = _ptr(flags)
        d.isnap = self.isnap
        if save:
            d.num_snaps = self.num_snaps
            d.snap_u, d.snap_v, d.snap_w = (_ptr(getattr(self, c + "_save_field"))
                                            for c in ("u", "v", "w"))
        d.rec_u, d.rec_v, d.rec_w = (_ptr(self._seis.get(c)) for c in ("u", "v", "w"))
        ws = self._workspace(d)  # kept alive for the queued kernels
        s = stream_handle()
        self._autotune(d, "forward",
                       lambda dd, n: _lib.call("trtm_forward_steps", dd, 0, n, int(bool(save)), s))
        flags.zero_()
        return d, flags, ws, sci

    def _finish_run(self, save, code, it_fail):
        """isnaps of the written snapshots; on a blow-up, outputs cut at the
        failing step as the reference's early return leaves them."""
        nt = self.nt
        last = nt if it_fail is None else it_fail  # steps whose snapshot the reference kept
        if save:
            isn = np.zeros(self.num_snaps, np.int32)
            for sl in range(self.num_snaps):
                it = (sl + 1) * self.isnap
                if it < last:
                    isn[sl] = it
            self.isnaps_field.from_numpy(isn)
        if code:
            for t in self._seis.values():  # the reference stops recording at the failing step
                t[it_fail + 1:].zero_()
            if save:  # and keeps no snapshot from the failing step on
                for comp in ("u", "v", "w"):
                    f = getattr(self, comp + "_save_field")
                    if isinstance(f, SnapshotField):
                        f.tensor[max(0, -(-it_fail // self.isnap) - 1):].zero_()
            return code
        print("Forward modeling completed")
        return 0

    def capture(self, save: bool = False, stability_check_interval: int = 100):
        """The whole run [0, nt) as a ``CapturedRun``: recorded once into a
        HIP graph (blocked or wave kernels), replayed by ``launch()``; no
        display callback and no early stop (``result()`` reads the check
        flags afterwards, with run()'s codes).  Every launch continues from
        the current state: call ``reset_state()`` first for a fresh run."""
        from .plan import StepPlan
        d, flags, ws, sci = self._prepare_run(save, stability_check_interval)
        snaps = [getattr(self, c + "_save_field", None) for c in ("u", "v", "w")] if save else []
        plan = StepPlan([d], 0, self.nt, forward=True, save=save,
                        keep=[ws, flags] + [f for f in snaps if f is not None])
        return CapturedRun(self, plan, flags, sci, save)


    def get_seismogram(self) -> tuple:
        """Seismograms (u, v, w), each (num_receivers, nt) numpy float32
        (forward_modeling.py:577-583)."""
        out = []
        for comp in ("u", "v", "w"):
            t = self._seis.get(comp)
            if t is None:
                out.append(np.zeros((self.num_receivers, self.nt), np.float32))
            else:
                out.append(self._transpose_to_host(t, self.nt, self.num_receivers))
        return tuple(out)

    def seismogram_tensors(self) -> dict:
        """Time-major (nt, num_receivers) device tensors of the active components."""
        return dict(self._seis)

    def cleanup(self):
        """Drop every device array this instance holds (forward_modeling.py:585-646)."""
        for name in ("sxx", "sxz", "szz", "syx", "syz", "u", "v", "w", "mxz", "myx", "myz",
                     "rho_u", "rho_w", "inv_rho_u", "inv_rho_w", "inv_rho", "absorb_coeff",
                     "u_save_field", "v_save_field", "w_save_field", "isnaps_field",
                     "src_loc_field", "recv_loc_field", "wavelet_u_field", "wavelet_v_field",
                     "wavelet_w_field", "surface_matrix", "mu", "lam", "rho_field"):
            if hasattr(self, name):
                setattr(self, name, None)
        self._seis = {}
        self._coef = {}
        self._inj = []


class CapturedRun:
    """A propagator's run captured by ``capture``."""

    def __init__(self, prop, plan, flags, sci, save):
        self.prop, self.plan, self.flags, self.sci, self.save = prop, plan, flags, sci, save

    def launch(self):
        """Clear the check flags and replay the steps on the current stream."""
        self.flags.zero_()
        self.plan.launch()

    def result(self) -> int:
        """Wait for the last launch; 0, or the code of the first failing check."""
        code, it_fail = self.prop._first_failure(self.flags.cpu().numpy(), self.sci, 0,
                                                 self.prop.nt)
        return self.prop._finish_run(self.save, code, it_fail)

    def close(self):
        self.plan.close()
