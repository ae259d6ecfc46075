"""Closed-form 2-D elastodynamic Green's functions for pinning the oracle.

Homogeneous isotropic full space, line force f(t) (N/m) applied as a body
force f(t) delta(x) delta(z).  Frequency domain with numpy's e^{+i omega t}
convention, where the outgoing scalar Green's function of (lap + k^2) g = -delta
is g = (-i/4) H0^(2)(k r).

  SH:    U_y = F g_s / mu
  P-SV:  U_ij = [k_s^2 delta_ij g_s + d_i d_j (g_s - g_p)] / (rho omega^2) * F_j

Velocities are i omega U, sampled at (it + 1/2) dt: the reference records the
velocity after the step that takes it from t_{it-1/2} to t_{it+1/2}, and adds
the wavelet sample it during that step (forward_modeling.py:547-551).

The reference injects  v += wavelet * dt*dx*dz / rho  (forward_modeling.py:430-443)
where a body force density f/(dx dz) on one cell gives  v += f dt / (rho dx dz);
so the equivalent line force is  f = wavelet * (dx*dz)**2.
"""
import numpy as np
from scipy.special import hankel2


def ricker(f0, nt, dt, t0=None):
    """r(t) = (1 - 2 (pi f0 (t-t0))^2) exp(-(pi f0 (t-t0))^2), t = it*dt, t0 = 1/f0."""
    t0 = 1.0 / f0 if t0 is None else t0
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def _g(k, r):
    return -0.25j * hankel2(0, k * r)


def _didj_g(k, X, Z, i, j):
    r = np.hypot(X, Z)
    gam = (X / r, Z / r)
    h0 = hankel2(0, k * r)
    h1 = hankel2(1, k * r)
    d1 = -k * h1
    d2 = -k ** 2 * (h0 - h1 / (k * r))
    return -0.25j * (d2 * gam[i] * gam[j] + d1 * (float(i == j) - gam[i] * gam[j]) / r)


class Green:
    def __init__(self, wavelet, dt, dx, dz, vp, vs, rho, nfft=None):
        self.nt = len(wavelet)
        self.dt = dt
        self.N = nfft or int(2 ** np.ceil(np.log2(8 * self.nt)))
        self.W = np.fft.rfft(np.asarray(wavelet, np.float64), self.N) * (dx * dz) ** 2
        om = 2 * np.pi * np.fft.rfftfreq(self.N, dt)
        om[0] = 1.0
        self.om = om
        self.vp, self.vs, self.rho = vp, vs, rho
        self.shift = np.exp(1j * om * dt / 2)

    def _to_time(self, U):
        V = 1j * self.om * U * self.shift
        V[0] = 0.0
        return np.fft.irfft(V, self.N)[: self.nt]

    def sh_velocity(self, X, Z):
        r = np.hypot(X, Z)
        mu = self.rho * self.vs ** 2
        return self._to_time(self.W * _g(self.om / self.vs, r) / mu)

    def psv_velocity(self, X, Z, i, j):
        """Velocity component i (0 = x, 1 = z) from a force along j."""
        ks, kp = self.om / self.vs, self.om / self.vp
        r = np.hypot(X, Z)
        G = (ks ** 2 * float(i == j) * _g(ks, r) + _didj_g(ks, X, Z, i, j) - _didj_g(kp, X, Z, i, j))
        G = G / (self.rho * self.om ** 2)
        return self._to_time(self.W * G)
