"""Size of FMA contraction at full length, on the CPU: the oracle built with
-ffp-contract=fast (liboracle_fma.so, the contraction Taichi's default
fast_math allows, rtm.py:104-128) against the operand-by-operand build
(liboracle.so) for the whole single-shot RTM.  The 200 x 200 x 1500 P-SV
shot runs by default (a few seconds); the 1200 x 600 x 4000 SH shot (about
two minutes on 8 cores) runs with TRTM_FULL_LENGTH_CPU=1.  Measured: P-SV
seismogram w 5.0e-7, image w 2.1e-7; SH seismogram v 9.8e-7, image v 6.7e-7
(isnap 10)."""
import os

import pytest

import oracle
from taichirtm_amd import configs


def _rtm(cfg, isnap, variant):
    shot = cfg.shots[0]
    mu, lam = cfg.moduli()
    m = oracle.Model(mu, lam, cfg.rho, cfg.dx, cfg.dz, cfg.fs, shot.nt, [shot.src], shot.rec,
                     cfg.fw, isnap=isnap, sci=100, variant=variant)
    r = oracle.rtm(m, list(cfg.wavelets(shot)), isnap=isnap)
    assert r["fw"]["code"] == 0 and r["bw"]["code"] == 0
    return r


def _spread(cfg, isnap, comp):
    a = _rtm(cfg, isnap, "strict")
    b = _rtm(cfg, isnap, "fma")
    e_seis = oracle.relative_l2(b["fw"]["seis"][comp], a["fw"]["seis"][comp])
    e_img = oracle.relative_l2(b["bw"]["result"][comp], a["bw"]["result"][comp])
    print("[fma] %s seismogram rel-L2 %.3e image rel-L2 %.3e" % (cfg.name, e_seis, e_img))
    return e_seis, e_img


def test_fma_contraction_psv_anomaly_full_length():
    e_seis, e_img = _spread(configs.psv_anomaly(), 1, 2)
    assert 0 < e_seis < 2e-6 and 0 < e_img < 2e-6


@pytest.mark.slow
@pytest.mark.skipif(os.environ.get("TRTM_FULL_LENGTH_CPU") != "1",
                    reason="two minutes of CPU; set TRTM_FULL_LENGTH_CPU=1")
def test_fma_contraction_sh_layered_full_length():
    e_seis, e_img = _spread(configs.sh_layered(), 10, 1)
    assert 0 < e_seis < 2e-6 and 0 < e_img < 2e-6
