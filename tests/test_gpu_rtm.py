"""The ReverseTimeMigration orchestrator end to end on the device: the image
of the fluent-API run equals the oracle's forward + backward + image on the
grid, locations, medium and snapshot interval the orchestrator derives
(rtm.py:846-1029), bit for bit; with topography too."""
import numpy as np
import pytest

from taichirtm_amd import configs
from taichirtm_amd.rtm import ReverseTimeMigration
from tests.parity import run_oracle, same

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case]", name, flush=True)


def _inputs(nt=200, fs=500.0, seed=0):
    rng = np.random.default_rng(seed)
    rec = np.arange(0.0, 20.5, 0.5)
    obs = rng.standard_normal((3, len(rec), nt)).astype(np.float32)
    k = np.exp(-0.5 * (np.arange(-6, 7) / 2.0) ** 2)
    obs = np.apply_along_axis(lambda x: np.convolve(x, k / k.sum(), "same"), 2, obs).astype(np.float32)
    w = configs.ricker(20.0, nt, 1.0 / fs).astype(np.float32)
    return rec, obs, w, fs


def _oracle_image(r, rec, obs, w):
    fw = r.absorbing_frame
    rho = np.full((r.nx, r.nz), 1500.0, np.float32)
    vs = np.full((r.nx, r.nz), r.v_fix, np.float32)
    vp = (np.sqrt((2 * r.poisson + 1) / (1 - 2 * r.poisson)) * vs).astype(np.float32)
    mu = (rho * vs ** 2).astype(np.float32)
    lam = (((vp / vs) ** 2 - 2) * mu).astype(np.float32)
    recs = np.array(r.receiver_loc_step, np.int32)
    if r.receivers_height is None:
        assert [p[0] for p in r.receiver_loc_step] == [int(x / r.dx + 2 * fw) for x in rec]
    kw = dict(nx=r.nx, nz=r.nz, dx=r.dx, dz=r.dz, nt=obs.shape[2], fs=r.fs, mu_field=mu,
              lam_field=lam, rho_field=rho, absorbing_frame=fw,
              src_loc_field=np.array(r.src_loc_step, np.int32), recv_loc_field=recs)
    wav = (w[None], 0.5 * w[None], w[None])
    return run_oracle(kw, wav, isnap=r.isnap_used, obs=tuple(obs),
                      surface=getattr(r, "surface_matrix", None))


@pytest.mark.parametrize("topo", [False, True])
def test_fluent_rtm_equals_oracle(topo):
    _case("ReverseTimeMigration fluent run, topography=%s" % topo)
    rec, obs, w, fs = _inputs()
    with ReverseTimeMigration() as r:
        r.set_observed_data(*obs).set_source(w, 0.5 * w, w, 7.3).set_receivers(rec)
        r.set_frequency(fs).fix_velocity(100.0).set_boundary(10).set_memory(1.0)
        if topo:
            r.set_topography(-0.08 * np.abs(rec - 10.0))
        assert r.run() is r
        # the first attempt (CFL 0.4: dx 0.5 m, vp dt / dx = 1.05) blows up and
        # is retried at CFL 0.2 (dx 1 m)
        assert r.CFL == 0.2 and r.dx == 1.0 and r.nx == r.nz == 60
        o = _oracle_image(r, rec, obs, w)
        assert o["code_f"] == 0 and o["code_b"] == 0
        for c, img in enumerate(r.get_results()):
            assert same(img, o["image"][c])
        assert np.abs(r.image_w).max() > 0
        ext = r.get_axes_extent()
        assert ext["xmin"] == -20.0 and ext["zmax"] == 60.0


def test_cfl_retry_follows_the_oracle_code():
    """The oracle's forward on the CFL 0.4 grid returns the same non-finite
    code the orchestrator retried on."""
    _case("CFL retry code")
    rec, obs, w, fs = _inputs()
    r = ReverseTimeMigration().set_observed_data(*obs).set_source(w, 0.5 * w, w, 7.3)
    r.set_receivers(rec).set_frequency(fs).fix_velocity(100.0).set_boundary(10)
    dx, dz, nx, nz, rho, vs, vp = r._setup_modeling_conditions(100.0, 0.4, 10)
    rec_step, src_step, _ = r._setup_surface_and_locations(nx, nz, dx, dz, 10)
    mu = (rho * vs ** 2).astype(np.float32)
    lam = (((vp / vs) ** 2 - 2) * mu).astype(np.float32)
    kw = dict(nx=nx, nz=nz, dx=dx, dz=dz, nt=obs.shape[2], fs=fs, mu_field=mu, lam_field=lam,
              rho_field=rho, absorbing_frame=10, src_loc_field=np.array(src_step, np.int32),
              recv_loc_field=np.array(rec_step, np.int32))
    o = run_oracle(kw, (w[None], 0.5 * w[None], w[None]), isnap=1, backward=False)
    assert o["code_f"] in (1, 2, 3)
