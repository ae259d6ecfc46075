"""CPU oracle for the TaichiRTM elastic FD / RTM hot path — TEST INFRASTRUCTURE.

Only the test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only to check results or to time the
reference-equivalent CPU path.  The product (``taichirtm_amd``) never imports it.

It is a ctypes front end over ``fd_oracle.c``, a plain-C restatement of the
reference's per-step kernel sequence (forward_modeling.py:203-575,
backward_modeling.py:206-579) in float32 and float64.  Arrays follow the
reference layouts: grid fields (nx, nz) C-order with z contiguous, seismograms
and wavelets (n, nt), snapshots (nx, nz, num_snaps).

Parity status of each oracle function and the check that pins it:
  sponge    closed form d = min(i, nx-1-i, nz-1-j) plus hand-derived overwrite
            order for nx < 2*FW (tests/test_oracle_closed_forms.py)
  prepare   harmonic / arithmetic means recomputed in float64 from their
            textbook definitions (tests/test_oracle_closed_forms.py)
  forward   2-D SH and P-SV analytic Green's functions (Hankel functions) in a
            homogeneous medium; linearity and time-shift invariants
            (tests/test_oracle_physics.py)
  backward  reflection identity: one backward step equals the forward step on
            the grid reflected in x and z (homogeneous medium, no sponge);
            raw receiver injection checked against the closed form
            (tests/test_oracle_physics.py)
  imaging   image equals the explicit sum over stored forward/backward
            wavefields with the reference's gating (tests/test_oracle_physics.py)
  topography  a surface matrix zero on row 0 and one elsewhere equals the flat
            free surface value for value; a stepped mask equals the same steps
            with syz, sxz, szz masked in numpy before each step
            (tests/test_oracle_physics.py)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ST_NAMES = ("u", "v", "w", "sxx", "szz", "sxz", "syx", "syz")


class _Problem(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("nt", ctypes.c_int),
        ("dx", ctypes.c_double), ("dz", ctypes.c_double), ("fs", ctypes.c_double),
        ("ns", ctypes.c_int), ("nr", ctypes.c_int),
        ("src_loc", ctypes.c_void_p), ("rec_loc", ctypes.c_void_p),
        ("isnap", ctypes.c_int), ("sci", ctypes.c_int),
        ("surface_matrix", ctypes.c_void_p),
    ]


class _Medium(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("mu", "lam", "myx", "myz", "mxz", "inv_rho", "inv_rho_u", "inv_rho_w", "absorb")]


_LIBS: dict = {}


def build(force: bool = False, plants: bool = False) -> None:
    """Compile the oracle variants with the committed Makefile."""
    args = ["make", "-s", "-C", HERE, "all"] + (["plants"] if plants else [])
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)


def _lib(variant="strict"):
    """variant: "strict" (no FMA contraction), "fma" (contraction allowed) or
    "plantN" (the strict build with deliberate mistake N, for pin tests)."""
    if variant is True:
        variant = "fma"
    elif variant is False or variant is None:
        variant = "strict"
    if variant not in _LIBS:
        name = {"strict": "liboracle.so", "fma": "liboracle_fma.so"}.get(
            variant, "liboracle_%s.so" % variant)
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build(plants=variant.startswith("plant"))
        lib = ctypes.CDLL(path)
        for sfx in ("_f32", "_f64"):
            getattr(lib, "oracle_forward" + sfx).restype = ctypes.c_int
            getattr(lib, "oracle_backward" + sfx).restype = ctypes.c_int
        _LIBS[variant] = lib
    return _LIBS[variant]


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def sponge(nx, nz, fw, a=0.0053, dtype=np.float32, variant="strict"):
    """forward_modeling.py:203-241."""
    out = np.empty((nx, nz), dtype=dtype)
    getattr(_lib(variant), "oracle_sponge" + _sfx(dtype))(
        ctypes.c_int(nx), ctypes.c_int(nz), ctypes.c_int(fw), ctypes.c_double(a), _p(out))
    return out


def prepare(mu, rho, dtype=np.float32, variant="strict"):
    """forward_modeling.py:243-307: averaged moduli, densities and reciprocals."""
    mu = np.ascontiguousarray(mu, dtype=dtype)
    rho = np.ascontiguousarray(rho, dtype=dtype)
    nx, nz = mu.shape
    names = ("myx", "myz", "mxz", "rho_u", "rho_w", "inv_rho", "inv_rho_u", "inv_rho_w")
    out = {n: np.empty((nx, nz), dtype=dtype) for n in names}
    with np.errstate(all="ignore"):
        getattr(_lib(variant), "oracle_prepare" + _sfx(dtype))(
            ctypes.c_int(nx), ctypes.c_int(nz), _p(mu), _p(rho), *[_p(out[n]) for n in names])
    return out


class Model:
    """One propagation problem: medium, geometry, and the state of one run."""

    def __init__(self, mu, lam, rho, dx, dz, fs, nt, src_loc, rec_loc, fw,
                 isnap=10, sci=100, surface_matrix=None, dtype=np.float32, variant="strict"):
        self.dtype = np.dtype(dtype)
        self.variant = variant
        self.mu = np.ascontiguousarray(mu, dtype=dtype)
        self.lam = np.ascontiguousarray(lam, dtype=dtype)
        self.rho = np.ascontiguousarray(rho, dtype=dtype)
        self.nx, self.nz = self.mu.shape
        self.dx, self.dz, self.fs, self.nt = float(dx), float(dz), float(fs), int(nt)
        self.src_loc = np.ascontiguousarray(np.asarray(src_loc, dtype=np.int32).reshape(-1, 2))
        self.rec_loc = np.ascontiguousarray(np.asarray(rec_loc, dtype=np.int32).reshape(-1, 2))
        self.fw = int(fw)
        self.isnap, self.sci = int(isnap), int(sci)
        self.surface_matrix = (None if surface_matrix is None
                               else np.ascontiguousarray(surface_matrix, dtype=dtype))
        self.coef = prepare(self.mu, self.rho, dtype, variant)
        self.coef["absorb"] = sponge(self.nx, self.nz, self.fw, 0.0053, dtype, variant)
        self._prob = _Problem(self.nx, self.nz, self.nt, self.dx, self.dz, self.fs,
                              len(self.src_loc), len(self.rec_loc),
                              _p(self.src_loc), _p(self.rec_loc), self.isnap, self.sci,
                              _p(self.surface_matrix))
        c = self.coef
        self._med = _Medium(_p(self.mu), _p(self.lam), _p(c["myx"]), _p(c["myz"]), _p(c["mxz"]),
                            _p(c["inv_rho"]), _p(c["inv_rho_u"]), _p(c["inv_rho_w"]), _p(c["absorb"]))

    def new_state(self):
        return {n: np.zeros((self.nx, self.nz), dtype=self.dtype) for n in ST_NAMES}

    def _st(self, state):
        for n in ST_NAMES:
            a = state[n]
            assert a.dtype == self.dtype and a.flags.c_contiguous and a.shape == (self.nx, self.nz)
        return (ctypes.c_void_p * 8)(*[_p(state[n]) for n in ST_NAMES])

    def forward(self, wavelets, state=None, seis=None, save=False, snaps=None,
                it0=0, it1=None):
        """forward_modeling.py:504-575. Returns dict with code, it_stop, state,
        seis (3 x (nr, nt)), snaps (3 x (nx, nz, num_snaps)), isnaps."""
        nt = self.nt
        it1 = nt if it1 is None else it1
        state = self.new_state() if state is None else state
        wav = [np.ascontiguousarray(np.asarray(w, dtype=self.dtype).reshape(len(self.src_loc), nt))
               for w in wavelets]
        if seis is None:
            seis = [np.zeros((len(self.rec_loc), nt), dtype=self.dtype) for _ in range(3)]
        num_snaps = nt // self.isnap
        if save and snaps is None:
            snaps = ([np.zeros((self.nx, self.nz, num_snaps), dtype=self.dtype) for _ in range(3)]
                     + [np.zeros(num_snaps, dtype=np.int32)])
        sn = snaps if save else [None, None, None, None]
        stop = ctypes.c_int(0)
        with np.errstate(all="ignore"):
            code = getattr(_lib(self.variant), "oracle_forward" + _sfx(self.dtype))(
                ctypes.byref(self._prob), self._st(state), ctypes.byref(self._med),
                *[_p(w) for w in wav], *[_p(s) for s in seis], ctypes.c_int(int(save)),
                *[_p(s) for s in sn], ctypes.c_int(it0), ctypes.c_int(it1), ctypes.byref(stop))
        return {"code": code, "it_stop": stop.value, "state": state, "seis": seis,
                "snaps": snaps, "num_snaps": num_snaps}

    def backward(self, obs, snaps, isnaps, num_snaps, isnap_interval, state=None,
                 result=None, synsrc=None, it0=0, it1=None):
        """backward_modeling.py:498-579. snaps: 3 x (nx, nz, num_snaps)."""
        nt = self.nt
        it1 = nt if it1 is None else it1
        state = self.new_state() if state is None else state
        ob = [np.ascontiguousarray(np.asarray(o, dtype=self.dtype).reshape(len(self.rec_loc), nt))
              for o in obs]
        if result is None:
            result = [np.zeros((self.nx, self.nz), dtype=self.dtype) for _ in range(3)]
        if synsrc is None:
            synsrc = [np.zeros((len(self.src_loc), nt), dtype=self.dtype) for _ in range(3)]
        if snaps is None or num_snaps == 0:
            snaps = [np.zeros((self.nx, self.nz, 1), dtype=self.dtype) for _ in range(3)]
            isnaps = np.zeros(1, dtype=np.int32)
            num_snaps = 0
        snaps = [np.ascontiguousarray(s, dtype=self.dtype) for s in snaps]
        isnaps = np.ascontiguousarray(isnaps, dtype=np.int32)
        stop = ctypes.c_int(0)
        with np.errstate(all="ignore"):
            code = getattr(_lib(self.variant), "oracle_backward" + _sfx(self.dtype))(
                ctypes.byref(self._prob), self._st(state), ctypes.byref(self._med),
                *[_p(o) for o in ob], *[_p(s) for s in synsrc], *[_p(s) for s in snaps],
                _p(isnaps), ctypes.c_int(num_snaps), ctypes.c_int(isnap_interval),
                *[_p(r) for r in result], ctypes.c_int(it0), ctypes.c_int(it1), ctypes.byref(stop))
        return {"code": code, "it_stop": stop.value, "state": state, "result": result,
                "synsrc": synsrc}


def rtm(model: Model, wavelets, obs=None, isnap=None):
    """Reference RTM sequencing (rtm.py:935-1000): forward with snapshots, then
    backward from ``obs`` (default: the forward's own seismograms) with
    imaging. Returns dict(fw=..., bw=...)."""
    isnap = model.isnap if isnap is None else isnap
    model.isnap = isnap
    model._prob.isnap = isnap
    fw = model.forward(wavelets, save=True)
    s = fw["snaps"]
    bw = model.backward(fw["seis"] if obs is None else obs, s[:3], s[3], fw["num_snaps"], isnap)
    return {"fw": fw, "bw": bw}


def relative_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
