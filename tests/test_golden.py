"""Committed golden fixture (tests/golden/tiny_rtm.npz, written by
tests/golden/make_golden.py from the oracle): the oracle still reproduces it
bit for bit, and the HIP path matches it bit for bit on the device."""
import numpy as np
import pytest

import oracle
from tests.golden import make_golden

FIX = np.load(make_golden.PATH, allow_pickle=False)


def test_oracle_reproduces_golden_fixture():
    got = make_golden.generate()
    for k in FIX.files:
        assert np.array_equal(FIX[k], got[k], equal_nan=True), k


def test_golden_fixture_is_nontrivial():
    assert np.abs(FIX["f32_image"]).max() > 0
    assert np.abs(FIX["f32_seis"]).max() > 0
    # float32 vs float64 run of the same scheme
    assert oracle.relative_l2(FIX["f32_seis"], FIX["f64_seis"]) < 1e-5
    assert oracle.relative_l2(FIX["f32_image"], FIX["f64_image"]) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["fused", "simple"])
def test_hip_matches_golden_fixture(kernel):
    print("\n[case] golden tiny_rtm kernel=%s" % kernel, flush=True)
    from tests.parity import run_hip
    kw = dict(nx=int(FIX["nx"]), nz=int(FIX["nz"]), dx=float(FIX["dx"]), dz=float(FIX["dz"]),
              nt=int(FIX["nt"]), fs=float(FIX["fs"]), mu_field=FIX["mu"], lam_field=FIX["lam"],
              rho_field=FIX["rho"], absorbing_frame=int(FIX["fw"]), src_loc_field=FIX["src"],
              recv_loc_field=FIX["rec"], num_sources=len(FIX["src"]),
              num_receivers=len(FIX["rec"]))
    h = run_hip(kw, tuple(FIX["wavelets"]), isnap=int(FIX["isnap"]), kernel=kernel, sci=1)
    assert h["code_f"] == 0 and h["code_b"] == 0
    assert np.array_equal(np.stack(h["seis"]), FIX["f32_seis"])
    assert np.array_equal(np.stack(h["image"]), FIX["f32_image"])
    assert np.array_equal(np.stack([h["state_f"][k] for k in oracle.ST_NAMES]), FIX["f32_state_f"])
    assert np.array_equal(np.stack([h["state_b"][k] for k in oracle.ST_NAMES]), FIX["f32_state_b"])
