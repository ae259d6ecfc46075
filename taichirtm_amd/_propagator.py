"""Set-up shared by ForwardModeling and BackwardModeling.

The reference repeats the same preparation in both classes
(forward_modeling.py:126-307, backward_modeling.py:125-310): kwargs, averaged
moduli and densities, reciprocal densities and the sponge.  Here it runs once
per propagator through the C ABI, and the per-step work is a single call into
the library for a whole range of steps.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .field import Field, ZeroField, as_host, as_tensor, device, stream_handle

SPONGE_A = 0.0053  # forward_modeling.py:240
STATE = ("u", "v", "w", "sxx", "szz", "sxz", "syx", "syz")
PSV_STATE = ("u", "w", "sxx", "szz", "sxz")
SH_STATE = ("v", "syx", "syz")
KERNELS = {"auto": _lib.KERNEL_AUTO, "simple": _lib.KERNEL_SIMPLE, "fused": _lib.KERNEL_FUSED,
           "blocked": _lib.KERNEL_BLOCKED, "wave": _lib.KERNEL_WAVE,
           # shape and family chosen by the autotuner on the first run
           "tuned": _lib.KERNEL_AUTO}


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, Field):
        p = t.data_ptr()
        return ctypes.c_void_p(p) if p else None
    return ctypes.c_void_p(t.data_ptr())


def _all_finite(t) -> bool:
    return bool(torch.isfinite(t).all().item())


class Medium:
    """Model arrays shared by every propagator on one grid: the caller's mu,
    lam, rho (and optional surface matrix), the averaged moduli and densities
    and their reciprocals (forward_modeling.py:243-307), the sponge
    (forward_modeling.py:203-241) and the float32 step constants."""

    def __init__(self, nx, nz, dx, dz, fs, absorbing_frame, mu, lam, rho, surface_matrix=None):
        self.nx, self.nz = int(nx), int(nz)
        self.dx, self.dz, self.fs = float(dx), float(dz), float(fs)
        self.fw = fw = int(absorbing_frame)
        if fw < 0 or fw > min(nx, nz):
            raise ValueError("absorbing_frame %d must lie in [0, min(nx, nz)] = [0, %d]"
                             % (fw, min(nx, nz)))
        device()
        _lib.load()
        self.mu = as_tensor(mu, shape=(nx, nz), name="mu_field")
        self.lam = as_tensor(lam, shape=(nx, nz), name="lam_field")
        self.rho_field = as_tensor(rho, shape=(nx, nz), name="rho_field")
        self.surface_matrix = (None if surface_matrix is None else
                               as_tensor(surface_matrix, shape=(nx, nz), name="surface_matrix_field"))
        dev = device()
        self.coef = {k: torch.empty((nx, nz), dtype=torch.float32, device=dev)
                     for k in ("myx", "myz", "mxz", "rho_u", "rho_w", "inv_rho", "inv_rho_u",
                               "inv_rho_w")}
        c = self.coef
        s = stream_handle()
        _lib.call("trtm_prepare_medium", nx, nz, _ptr(self.mu), _ptr(self.rho_field),
                  _ptr(c["myx"]), _ptr(c["myz"]), _ptr(c["mxz"]), _ptr(c["rho_u"]),
                  _ptr(c["rho_w"]), _ptr(c["inv_rho"]), _ptr(c["inv_rho_u"]),
                  _ptr(c["inv_rho_w"]), s)
        tab = np.zeros(fw + 1, np.float32)
        _lib.call("trtm_sponge_table", fw, ctypes.c_float(SPONGE_A),
                  tab.ctypes.data_as(ctypes.c_void_p))
        self.sponge_table = as_tensor(tab)
        self.absorb_coeff = Field(np.float32, (nx, nz))
        _lib.call("trtm_sponge_fill", nx, nz, fw, _ptr(self.sponge_table),
                  _ptr(self.absorb_coeff.tensor), s)
        # float32 constants as the reference narrows them (forward_modeling.py:320-322,430)
        dt = 1.0 / self.fs
        self.dt32 = np.float32(dt)
        self.idx32 = np.float32(1.0 / self.dx)
        self.idz32 = np.float32(1.0 / self.dz)
        self.factor32 = np.float32(dt * self.dx * self.dz)


def set_shape(d, shape):
    """Kernel family and block shape (family, tile_x, tile_z, block_steps) into a
    step descriptor."""
    family, tx, tz, k = shape
    d.kernel = KERNELS[family]
    d.tile_x, d.tile_z, d.block_steps = tx, tz, k


class Propagator:
    """Medium, sponge and geometry of one forward or backward propagator."""

    _direction = "forward"

    def _setup(self, kwargs, inj_keys):
        self.nx = int(kwargs["nx"])
        self.nz = int(kwargs["nz"])
        self.dx = float(kwargs["dx"])
        self.dz = float(kwargs["dz"])
        self.nt = int(kwargs["nt"])
        self.fs = float(kwargs["fs"])
        self.absorbing_frame = int(kwargs.get("absorbing_frame", 60))
        self.dt = 1.0 / self.fs
        self.num_sources = int(kwargs["num_sources"])
        self.num_receivers = int(kwargs["num_receivers"])
        kernel = kwargs.get("kernel", "auto")
        if kernel not in KERNELS:
            raise ValueError("kernel must be one of %s" % sorted(KERNELS))
        self.kernel = kernel
        self.tile_x = int(kwargs.get("tile_x", 0))
        self.block_steps = int(kwargs.get("block_steps", 0))
        self.tile_z = int(kwargs.get("tile_z", 0))
        self._tuned = None  # (family, tile_x, tile_z, block_steps) once tuned
        nx, nz, nt = self.nx, self.nz, self.nt
        if nx < 2 or nz < 2 or nt < 1:
            raise ValueError("grid must be at least 2 x 2 with nt >= 1")
        fw = self.absorbing_frame
        device()  # fail loudly without a HIP device
        _lib.load()

        self.src_loc_field = kwargs["src_loc_field"]
        self.recv_loc_field = kwargs["recv_loc_field"]
        self._src_host = as_host(self.src_loc_field, np.int32).reshape(-1, 2)
        self._rec_host = as_host(self.recv_loc_field, np.int32).reshape(-1, 2)
        if len(self._src_host) != self.num_sources:
            raise ValueError("src_loc_field has %d rows, num_sources is %d"
                             % (len(self._src_host), self.num_sources))
        if len(self._rec_host) != self.num_receivers:
            raise ValueError("recv_loc_field has %d rows, num_receivers is %d"
                             % (len(self._rec_host), self.num_receivers))
        self._src_dev = as_tensor(self._src_host, np.int32)
        self._rec_dev = as_tensor(self._rec_host, np.int32)

        # medium (kept by reference, as forward_modeling.py:194-197); shots on
        # the same model can share one Medium
        med = kwargs.get("medium")
        if med is None:
            sm = kwargs.get("surface_matrix_field", None)
            med = Medium(nx, nz, self.dx, self.dz, self.fs, fw, kwargs["mu_field"],
                         kwargs["lam_field"], kwargs["rho_field"], sm)
        elif (med.nx, med.nz, med.fw) != (nx, nz, fw) or (med.dx, med.dz, med.fs) != (
                self.dx, self.dz, self.fs):
            raise ValueError("medium does not match nx, nz, dx, dz, fs, absorbing_frame")
        self.medium = med
        self.mu, self.lam, self.rho_field = med.mu, med.lam, med.rho_field
        self.surface_matrix = med.surface_matrix
        self._coef = med.coef
        for k, t in med.coef.items():
            setattr(self, k, Field(tensor=t))
        self._sponge_table = med.sponge_table
        self.absorb_coeff = med.absorb_coeff
        self._dt32, self._idx32, self._idz32, self._factor32 = (
            med.dt32, med.idx32, med.idz32, med.factor32)

        # injection data and its deterministic per-cell plan
        self._inj = [as_tensor(kwargs[k], shape=(len(self._inj_host_loc()), nt), name=k)
                     for k in inj_keys]
        self._plan_injection()
        self.systems = self._active_systems()

    # -- geometry -----------------------------------------------------------

    def _inj_host_loc(self):
        return self._src_host if self._direction == "forward" else self._rec_host

    def _inj_dev_loc(self):
        return self._src_dev if self._direction == "forward" else self._rec_dev

    def _rec_host_loc(self):
        return self._rec_host if self._direction == "forward" else self._src_host

    def _rec_dev_loc(self):
        return self._rec_dev if self._direction == "forward" else self._src_dev

    def _plan_injection(self):
        loc = np.ascontiguousarray(self._inj_host_loc(), dtype=np.int32)
        n = len(loc)
        order = np.zeros(max(n, 1), np.int32)
        start = np.zeros(n + 1, np.int32)
        ng = ctypes.c_int32(0)
        _lib.call("trtm_plan_injection", self.nx, self.nz, n, loc.ctypes.data_as(ctypes.c_void_p),
                  order.ctypes.data_as(ctypes.c_void_p), start.ctypes.data_as(ctypes.c_void_p),
                  ctypes.byref(ng))
        rec = np.ascontiguousarray(self._rec_host_loc())
        bad = (rec[:, 0] < 0) | (rec[:, 0] >= self.nx) | (rec[:, 1] < 0) | (rec[:, 1] >= self.nz)
        if np.any(bad):
            raise ValueError("location (%d, %d) outside the %d x %d grid"
                             % (*rec[np.argmax(bad)], self.nx, self.nz))
        self._inj_order = as_tensor(order, np.int32)
        self._inj_start = as_tensor(start[: ng.value + 1], np.int32)
        self._n_groups = ng.value

    def _active_systems(self):
        """Systems that can leave exact zero: a system whose input is all zero
        stays zero only if every coefficient its update multiplies by is
        finite (0 * inf = NaN in the reference too, e.g. mxz overflow)."""
        I = (slice(1, self.nx - 1), slice(1, self.nz - 1))
        c = self._coef
        loc = self._inj_host_loc()
        cells = (torch.as_tensor(loc[:, 0], device=self.mu.device).long(),
                 torch.as_tensor(loc[:, 1], device=self.mu.device).long())
        fwd = self._direction == "forward"
        sm_ok = self.surface_matrix is None or _all_finite(self.surface_matrix)
        u, v, w = self._inj

        def stays_zero(data, coefs, inj_coefs):
            if any(bool((d != 0).any().item()) for d in data):
                return False
            if not sm_ok or not all(_all_finite(t[I]) for t in coefs):
                return False
            if fwd and len(loc) and not all(_all_finite(t[cells]) for t in inj_coefs):
                return False
            return True

        systems = 0
        if not stays_zero((u, w), (self.lam, self.mu, c["mxz"], c["inv_rho_u"], c["inv_rho_w"]),
                          (c["inv_rho_u"], c["inv_rho_w"])):
            systems |= _lib.PSV
        if not stays_zero((v,), (c["myx"], c["myz"], c["inv_rho"]), (c["inv_rho"],)):
            systems |= _lib.SH
        return systems

    def _alloc_state(self):
        dev = device()
        for name in STATE:
            active = ((name in PSV_STATE and self.systems & _lib.PSV)
                      or (name in SH_STATE and self.systems & _lib.SH))
            if active:
                setattr(self, name, Field(tensor=torch.zeros((self.nx, self.nz),
                                                             dtype=torch.float32, device=dev)))
            else:
                setattr(self, name, ZeroField((self.nx, self.nz)))

    def reset_state(self):
        """Zero the wavefield (and any recorded output) so the next run starts
        from rest, as a freshly constructed propagator does."""
        for name in STATE:
            t = getattr(self, name).tensor
            if t is not None:
                t.zero_()
        for buf in (getattr(self, "_seis", {}), getattr(self, "_img", {}), getattr(self, "_syn", {})):
            for t in buf.values():
                t.zero_()

    def _desc(self):
        d = _lib.StepDesc()
        d.struct_size = ctypes.sizeof(_lib.StepDesc)
        d.nx, d.nz, d.nt = self.nx, self.nz, self.nt
        d.dt, d.inv_dx, d.inv_dz = self._dt32, self._idx32, self._idz32
        d.src_factor = self._factor32
        d.fw = self.absorbing_frame
        d.systems = self.systems
        d.kernel = KERNELS[self.kernel]
        d.tile_x, d.block_steps, d.tile_z = self.tile_x, self.block_steps, self.tile_z
        if self._tuned is not None:
            set_shape(d, self._tuned)
        d.sponge_table = _ptr(self._sponge_table)
        c = self._coef
        d.lam, d.mu = _ptr(self.lam), _ptr(self.mu)
        for k in ("mxz", "myx", "myz", "inv_rho", "inv_rho_u", "inv_rho_w"):
            setattr(d, k, _ptr(c[k]))
        d.surface_matrix = _ptr(self.surface_matrix)
        for k in STATE:
            setattr(d, k, _ptr(getattr(self, k)))
        loc = self._inj_host_loc()
        d.n_inj, d.n_inj_groups = len(loc), self._n_groups
        d.inj_loc = _ptr(self._inj_dev_loc()) if len(loc) else None
        d.inj_order = _ptr(self._inj_order)
        d.inj_group_start = _ptr(self._inj_start)
        d.inj_u, d.inj_v, d.inj_w = (_ptr(t) if t.numel() else None for t in self._inj)
        rl = self._rec_host_loc()
        d.n_rec = len(rl)
        d.rec_loc = _ptr(self._rec_dev_loc()) if len(rl) else None
        return d

    def _autotune(self, d, direction, launch):
        """kernel="tuned": choose the blocked kernel family and shape by timing
        every compiled shape on this propagator's own buffers (``launch(d,
        steps)`` runs steps [0, steps)), then zero the state and every output
        the trials wrote.  The caller clears its check flags."""
        if self.kernel != "tuned" or self._tuned is not None:
            return
        from . import autotune
        key = (direction, self.nx, self.nz, self.absorbing_frame, self.systems,
               int(d.num_snaps > 0), self.surface_matrix is not None, self.num_receivers > 0)
        window = min(self.nt, 64)

        def trial(shape):
            set_shape(d, shape)
            launch(d, window)

        self._tuned = autotune.tune(key, self.systems, trial)
        set_shape(d, self._tuned)
        self.reset_state()

    def block_config(self):
        """(tile rows, tile columns, threads, steps per launch) of the blocked
        kernel for this propagator's wave systems and requests."""
        d = self._desc()
        out = (ctypes.c_int32 * 4)()
        _lib.call("trtm_block_config", d, out)
        return tuple(out)

    def _workspace(self, d):
        if self.kernel == "simple":
            return None
        nbytes = int(_lib.load().trtm_workspace_bytes(ctypes.byref(d)))
        if nbytes < 0:
            raise _lib.HIPError("trtm_workspace_bytes rejected the descriptor")
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device())
        d.workspace = _ptr(ws)
        d.workspace_bytes = nbytes
        return ws

    def _first_failure(self, flags_host, sci, it_lo, it_hi):
        """First check step in [it_lo, it_hi) with a nonzero code, as the
        reference returns at the first failing check."""
        first = (it_lo + sci - 1) // sci
        for k in range(first, (it_hi - 1) // sci + 1):
            if flags_host[k]:
                return int(flags_host[k]), k * sci
        return 0, None

    @staticmethod
    def _chunks(nt, sci, stops):
        """Step ranges ending after each step in ``stops`` (callback points)
        and at least every ~2048 steps so a blow-up is seen early."""
        span = sci * max(1, 2048 // sci)
        edges = set(range(span, nt, span)) | {s + 1 for s in stops if s + 1 < nt} | {nt}
        out, a = [], 0
        for b in sorted(edges):
            out.append((a, b))
            a = b
        return out

    def _transpose_to_host(self, t, rows, cols):
        """(rows, cols) device tensor -> numpy (cols, rows)."""
        out = torch.empty((cols, rows), dtype=torch.float32, device=t.device)
        _lib.call("trtm_transpose", rows, cols, _ptr(t), _ptr(out), stream_handle())
        return out.cpu().numpy()
