"""Reverse Time Migration orchestrator (src/rtm.py:43-1097) on the MI355X path.

Same public surface as the reference: ``init_taichi`` / ``reset_taichi``, and
``ReverseTimeMigration`` with its fluent setters, deprecated kwargs
constructor, context manager, ``run``, ``save_result``, ``get_results`` and
``get_axes_extent``.  Host-side steps (velocity scan, grid set-up, topography
mask, snapshot budget, CFL retry) follow the reference; the propagators are
``ForwardModeling`` / ``BackwardModeling`` of this package.

One deliberate difference: the reference's CFL-halving retry loop has no bound
(rtm.py:882), so non-finite *input* data halves dx forever; here it stops after
``max_cfl_retries`` halvings with a RuntimeError.
"""
from __future__ import annotations

import os
import warnings
from typing import Callable, Optional

import numpy as np
import torch

from .backward_modeling import BackwardModeling
from .field import device
from .forward_modeling import ForwardModeling


def reset_taichi():
    """Release cached device memory (the reference calls ti.reset(), rtm.py:43-53)."""
    if torch.cuda.is_available():
        torch.cuda.synchronize()
        torch.cuda.empty_cache()


def init_taichi(backend: str = "cpu", extreme_optimization: bool = False, **kwargs):
    """Initialise the device (rtm.py:56-133).

    ``backend`` and the Taichi keyword arguments are accepted for signature
    compatibility; the only backend is the HIP device, and the kernels are
    always the IEEE float32 ones (no fast-math).  Raises when no HIP device or
    no built library is available, since there is no CPU fallback."""
    device()
    from . import _lib
    _lib.load()
    return {"backend": "hip", "device": torch.cuda.get_device_name(0),
            "requested_backend": backend, "extreme_optimization": extreme_optimization}


class ReverseTimeMigration:
    """Reverse Time Migration for seismic exploration (rtm.py:136-1097)."""

    def __init__(self, **kwargs):
        self.observed_u = self.observed_v = self.observed_w = None
        self.source_u = self.source_v = self.source_w = None
        self.receiver_loc = None
        self.receiver_num = None
        self.source_loc = None
        self.fs = None
        self.nt = None
        self.rho = 1500
        self.poisson = 0.33
        self.isnap = 216
        self.absorbing_frame = 50
        self.vmin = 10.0
        self.vmax = 500.0
        self.vstep = 10
        self.v_fix = None
        self.total_allocate_memory_gb = 8.0
        self.receivers_height = None
        self.debug = False
        self.max_cfl_retries = 8
        self.kernel = "auto"
        self._fw_instance = None
        self._bw_instance = None
        self._run_fields = []
        if kwargs:
            warnings.warn(
                "The kwargs-based constructor is deprecated. "
                "Please use the fluent API (set_observed_data, set_source, etc.) instead.",
                DeprecationWarning, stacklevel=2)
            self._init_from_kwargs(kwargs)

    def _init_from_kwargs(self, kwargs):
        self.observed_u = np.asarray(kwargs["observed_u"], dtype=np.float32)
        self.observed_v = np.asarray(kwargs["observed_v"], dtype=np.float32)
        self.observed_w = np.asarray(kwargs["observed_w"], dtype=np.float32)
        self.source_u = np.asarray(kwargs["source_u"], dtype=np.float32)
        self.source_v = np.asarray(kwargs["source_v"], dtype=np.float32)
        self.source_w = np.asarray(kwargs["source_w"], dtype=np.float32)
        self.receiver_loc = np.asarray(kwargs["receiver_loc"], dtype=np.float32)
        self.receiver_num = len(self.receiver_loc)
        self.source_loc = float(kwargs["source_loc"])
        self.fs = float(kwargs["fs"])
        self.rho = kwargs.get("rho", 1500)
        self.poisson = kwargs.get("poisson_ratio", 0.33)
        self.isnap = kwargs.get("isnap", 216)
        self.nt = self.observed_u.shape[1]
        self.absorbing_frame = kwargs.get("absorbing_frame", 50)
        self.vmin = kwargs.get("vmin", 10.0)
        self.vmax = kwargs.get("vmax", 500.0)
        self.vstep = kwargs.get("vstep", 10)
        self.v_fix = kwargs.get("v_fix", None)
        self.total_allocate_memory_gb = kwargs.get("total_allocate_memory_gb", 8.0)
        self.receivers_height = kwargs.get("receivers_height", None)
        if self.receivers_height is not None:
            self.receivers_height = np.asarray(self.receivers_height, dtype=np.float32)
            self.receivers_height = self.receivers_height - np.max(self.receivers_height)
        self._check_parameters()

    # context manager ------------------------------------------------------
    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self.cleanup()
        return False

    def cleanup(self, reset_taichi_runtime: bool = True):
        if self._fw_instance is not None:
            self._fw_instance.cleanup()
            self._fw_instance = None
        if self._bw_instance is not None:
            self._bw_instance.cleanup()
            self._bw_instance = None
        self._run_fields.clear()
        if reset_taichi_runtime:
            reset_taichi()

    # fluent API (rtm.py:332-561) -------------------------------------------
    def set_observed_data(self, observed_u, observed_v, observed_w):
        self.observed_u = np.asarray(observed_u, dtype=np.float32)
        self.observed_v = np.asarray(observed_v, dtype=np.float32)
        self.observed_w = np.asarray(observed_w, dtype=np.float32)
        self.nt = self.observed_u.shape[1]
        return self

    def set_source(self, source_u, source_v, source_w, source_loc):
        self.source_u = np.asarray(source_u, dtype=np.float32)
        self.source_v = np.asarray(source_v, dtype=np.float32)
        self.source_w = np.asarray(source_w, dtype=np.float32)
        self.source_loc = float(source_loc)
        return self

    def set_receivers(self, receiver_loc):
        self.receiver_loc = np.asarray(receiver_loc, dtype=np.float32)
        self.receiver_num = len(self.receiver_loc)
        return self

    def set_frequency(self, fs):
        self.fs = float(fs)
        return self

    def set_velocity_range(self, vmin, vmax, vstep):
        self.vmin, self.vmax, self.vstep = float(vmin), float(vmax), int(vstep)
        return self

    def fix_velocity(self, velocity):
        self.v_fix = float(velocity)
        return self

    def set_boundary(self, absorbing_frame):
        self.absorbing_frame = int(absorbing_frame)
        return self

    def set_memory(self, total_allocate_memory_gb):
        self.total_allocate_memory_gb = float(total_allocate_memory_gb)
        return self

    def set_density(self, rho):
        self.rho = float(rho)
        return self

    def set_poisson_ratio(self, poisson_ratio):
        self.poisson = float(poisson_ratio)
        return self

    def set_topography(self, receivers_height):
        self.receivers_height = np.asarray(receivers_height, dtype=np.float32)
        self.receivers_height = self.receivers_height - np.max(self.receivers_height)
        return self

    def set_debug(self, debug: bool = True):
        self.debug = debug
        return self

    def set_kernel(self, kernel: str):
        """Step kernel of the propagators: "auto", "fused" or "simple"."""
        self.kernel = kernel
        return self

    # checks and host preprocessing ------------------------------------------
    def _check_parameters(self):
        """rtm.py:563-601 (same messages)."""
        if self.observed_u is None or self.observed_v is None or self.observed_w is None:
            raise ValueError("Observed data must be provided. "
                             "Use set_observed_data(observed_u, observed_v, observed_w) to set it.")
        if self.source_u is None or self.source_v is None or self.source_w is None:
            raise ValueError("Source functions must be provided. "
                             "Use set_source(source_u, source_v, source_w, source_loc) to set them.")
        if self.receiver_loc is None:
            raise ValueError("Receiver locations must be provided. "
                             "Use set_receivers(receiver_loc) to set them.")
        if self.source_loc is None:
            raise ValueError("Source location must be provided. "
                             "Use set_source(source_u, source_v, source_w, source_loc) to set it.")
        if self.fs is None:
            raise ValueError("Sampling frequency must be provided. Use set_frequency(fs) to set it.")
        if self.observed_u.shape != self.observed_v.shape or self.observed_u.shape != self.observed_w.shape:
            raise ValueError("Observed data shapes must match")
        if self.source_u.shape != self.source_v.shape or self.source_u.shape != self.source_w.shape:
            raise ValueError("Source function shapes must match")
        if self.fs <= 0:
            raise ValueError("Sampling frequency must be positive")
        if self.receivers_height is not None and len(self.receivers_height) != self.receiver_num:
            raise ValueError("Receiver height array size must match receiver count")

    def estimate_velocity(self, array: np.ndarray, vmin: float, vmax: float, vstep: int = 10) -> float:
        """Slant-stack velocity scan (rtm.py:603-652): for each of vstep
        geometrically spaced trial velocities, align traces by the moveout
        |x_r - x_s| / v and take |sum_t prod_r trace_r(t + shift_r)|; return the
        velocity with the largest value."""
        print("Estimating velocity...")
        _, L = array.shape
        r = (vmax / vmin) ** (1 / (vstep - 1))
        best, v_est = 0, vmin
        for v in vmin * r ** np.arange(vstep):
            shifts = np.round(self.fs * np.abs(self.receiver_loc - self.source_loc) / v).astype(np.int32)
            shifts = shifts - np.min(shifts)
            span = L - np.max(shifts)
            if span <= 0:
                continue
            idx = shifts[:, None] + np.arange(span)
            aligned = array[np.arange(array.shape[0])[:, None], idx]
            score = np.abs(np.sum(np.prod(aligned, axis=0)))
            if score > best:
                best, v_est = score, v
        print(f"Estimated velocity: {v_est:.2f} m/s")
        return float(v_est)

    def _setup_modeling_conditions(self, estimated_v: float, CFL: float, absorbing_frame: int):
        """rtm.py:654-666; Vp from Vs and Poisson's ratio as the reference writes it."""
        dx = estimated_v / self.fs / CFL
        dz = dx
        nx = int((np.max(self.receiver_loc) - np.min(self.receiver_loc)) / dx)
        nx += 4 * absorbing_frame
        nz = nx
        rho = self.rho * np.ones((nx, nz), dtype=np.float32)
        vs = estimated_v * np.ones((nx, nz), dtype=np.float32)
        vp = (np.sqrt((2 * self.poisson + 1) / (1 - 2 * self.poisson)) * vs).astype(np.float32)
        return dx, dz, nx, nz, rho, vs, vp

    def _setup_surface_and_locations(self, nx, nz, dx, dz, absorbing_frame):
        """Grid indices of receivers and source, and the topography mask (rtm.py:668-717)."""
        offset = 2 * absorbing_frame
        receiver_loc_step = [[int(loc / dx + offset), 1] for loc in self.receiver_loc]
        src_loc_step = [[int(self.source_loc / dx + offset), 1]]
        surface_matrix = np.ones((nx, nz), dtype=np.float32)
        height_array = np.zeros(nx, dtype=np.float32)
        if self.receivers_height is not None:
            hstep = -1 * self.receivers_height / dz
            for i, locstep in enumerate(receiver_loc_step):
                h = int(hstep[i])
                receiver_loc_step[i][1] = h
                if i == 0:
                    surface_matrix[:locstep[0], :h] = 0
                    height_array[:locstep[0]] = h
                if i > 0:
                    x1, x0 = receiver_loc_step[i][0], receiver_loc_step[i - 1][0]
                    h1, h0 = hstep[i], hstep[i - 1]
                    for ix in range(x0, x1):
                        lean = (h1 - h0) / (x1 - x0)
                        hix = int(lean * (ix - x0) + h0)
                        surface_matrix[ix, :hix] = 0
                        height_array[ix] = hix
            height_array[receiver_loc_step[-1][0]:] = hstep[-1]
            surface_matrix[receiver_loc_step[-1][0]:, :int(hstep[-1])] = 0
            src_loc_step[0][1] = int(height_array[src_loc_step[0][0]]) + 1
            self.height_array = height_array
            self.surface_matrix = surface_matrix
        else:
            self.surface_matrix = None
        return receiver_loc_step, src_loc_step, surface_matrix if self.surface_matrix is not None else None

    def _compute_isnap(self, nx: int, nz: int, nt: int, num_sources: int, num_receivers: int) -> int:
        """Snapshot interval from the memory budget, with the reference's
        accounting (rtm.py:719-844)."""
        total = int(self.total_allocate_memory_gb * 1024 * 1024 * 1024)
        inputs = (3 * nx * nz * 4 + (num_sources * 2 + num_receivers * 2) * 4
                  + 3 * num_sources * nt * 4 + nx * nz * 4)
        fw_mem = ForwardModeling.estimate_memory_bytes(nx, nz, nt, num_sources, num_receivers)
        obs_mem = 3 * num_receivers * nt * 4
        bw_mem = BackwardModeling.estimate_memory_bytes(nx, nz, nt, num_sources, num_receivers)
        per_snap = nx * nz * 4 * 3 * 2 + 4
        fixed = 300 * 1024 * 1024 + inputs + fw_mem + obs_mem + bw_mem
        avail = total - fixed - int(total * 0.05)
        if avail <= 0:
            print(f"Warning: Very limited memory available ({self.total_allocate_memory_gb:.1f} GB), "
                  "using minimum snapshots")
            return nt
        max_snaps = avail // per_snap
        if max_snaps <= 0:
            return nt
        if max_snaps >= nt:
            return 1
        isnap = max(1, int(np.ceil(nt / max_snaps)))
        if self.debug:
            print(f"  Snapshot interval: isnap={isnap} ({nt // isnap} snapshots)")
        return isnap

    def run(self, method: str = "cross_correlation", display_callback: Optional[Callable] = None):
        """Forward with snapshots, then backward with imaging; halve CFL and
        retry on a non-finite code (rtm.py:846-1029)."""
        self._check_parameters()
        if self.v_fix is not None:
            print(f"Using fixed velocity: {self.v_fix} m/s")
            estimated_v = self.v_fix
        else:
            estimated_v = self.estimate_velocity(self.observed_v, self.vmin, self.vmax, self.vstep)
        CFL = 0.8
        fw_frame = self.absorbing_frame
        flag = 99
        retries = 0
        print("\nStarting forward modeling...")
        while flag:
            if retries > self.max_cfl_retries:
                raise RuntimeError("modeling still non-finite after %d CFL halvings"
                                   % self.max_cfl_retries)
            retries += 1
            CFL *= 0.5
            flag = 0
            dx, dz, nx, nz, rho, vs, vp = self._setup_modeling_conditions(estimated_v, CFL, fw_frame)
            rec_step, src_step, surface_matrix = self._setup_surface_and_locations(
                nx, nz, dx, dz, fw_frame)
            wu = self.source_u.reshape(1, -1) if self.source_u.ndim == 1 else self.source_u
            wv = self.source_v.reshape(1, -1) if self.source_v.ndim == 1 else self.source_v
            ww = self.source_w.reshape(1, -1) if self.source_w.ndim == 1 else self.source_w
            ns, nr = len(src_step), len(rec_step)
            isnap = self._compute_isnap(nx, nz, self.nt, ns, nr)
            mu_np = (rho * vs ** 2).astype(np.float32)
            lam_np = (((vp / vs) ** 2 - 2) * mu_np).astype(np.float32)
            common = dict(nx=nx, nz=nz, dx=dx, dz=dz, nt=self.nt, fs=self.fs,
                          mu_field=mu_np, lam_field=lam_np, rho_field=rho,
                          absorbing_frame=fw_frame,
                          src_loc_field=np.array(src_step, dtype=np.int32),
                          recv_loc_field=np.array(rec_step, dtype=np.int32),
                          surface_matrix_field=surface_matrix, num_sources=ns,
                          num_receivers=nr, kernel=self.kernel)
            fw = ForwardModeling(wavelet_u_field=np.asarray(wu, np.float32),
                                 wavelet_v_field=np.asarray(wv, np.float32),
                                 wavelet_w_field=np.asarray(ww, np.float32), isnap=isnap, **common)
            self._fw_instance = fw
            flag = fw.run(save=True, display_callback=display_callback)
            if flag:
                print(f"Forward modeling failed with flag {flag}, reducing CFL...")
                fw.cleanup()
                self._fw_instance = None
                continue
            print("Forward modeling completed successfully.")
            bw = BackwardModeling(obsdata_u_field=self.observed_u, obsdata_v_field=self.observed_v,
                                  obsdata_w_field=self.observed_w, **common)
            self._bw_instance = bw
            flag = bw.run_calc(import_fwdata_u=fw.u_save_field, import_fwdata_v=fw.v_save_field,
                               import_fwdata_w=fw.w_save_field, isnaps_field=fw.isnaps_field,
                               num_snaps=fw.num_snaps, isnap_interval=isnap, method=method,
                               display_callback=display_callback)
            if flag:
                print(f"Backward modeling failed with flag {flag}, reducing CFL...")
                bw.cleanup()
                fw.cleanup()
                self._bw_instance = self._fw_instance = None
                continue
            print("Backward modeling completed successfully.")
        print("\nSimulation completed.")
        print(f"Estimated velocity: {estimated_v} m/s")
        print(f"CFL: {CFL}")
        print(f"Grid: nx={nx}, nz={nz}, dx={dx:.4f}")
        self.image_u, self.image_v, self.image_w = bw.get_results()
        self.dx, self.dz, self.nx, self.nz = dx, dz, nx, nz
        self.offset = 2 * fw_frame * dx
        self.CFL = CFL
        self.src_loc_step = src_step
        # not kept by the reference: the receiver grid cells and snapshot interval used
        self.receiver_loc_step = rec_step
        self.isnap_used = isnap
        return self

    def save_result(self, directory: str, savename: str):
        """npz with the reference's keys and values (rtm.py:1031-1070); a
        missing surface matrix is stored as None, as the reference stores it
        (numpy keeps it as a pickled object entry, which load_rtm_results
        returns as None without unpickling)."""
        xmin = float(-self.offset)
        xmax = float(self.dx * self.nx - self.offset)
        sm = getattr(self, "surface_matrix", None)
        if not os.path.exists(directory):
            os.makedirs(directory)
        np.savez_compressed(
            os.path.join(directory, savename + ".npz"),
            u=self.image_u, v=self.image_v, w=self.image_w, receiver_loc=self.receiver_loc,
            src_loc_step=np.asarray(self.src_loc_step), offset=self.offset, xmin=xmin, xmax=xmax,
            zmin=0.0, zmax=float(self.dz * self.nz), dx=self.dx, dz=self.dz, nx=self.nx,
            nz=self.nz, surface_matrix=sm)
        print(f"Results saved to {os.path.join(directory, savename)}.npz")

    def get_results(self) -> tuple:
        return self.image_u, self.image_v, self.image_w

    def get_axes_extent(self) -> dict:
        return {"xmin": -self.offset, "xmax": self.dx * self.nx - self.offset,
                "zmin": 0.0, "zmax": self.dz * self.nz}
