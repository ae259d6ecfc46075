"""Parity of the HIP path against the oracle: forward seismograms, final
wavefields (all 8 fields), images and return codes, bit for bit, for the fused
and the simple step kernels.  Bar: bitwise equality (both sides evaluate the
reference's float32 expressions operand by operand, without contraction)."""
import numpy as np
import pytest

import oracle
from taichirtm_amd import configs
from tests.parity import assert_bitwise, run_hip, run_oracle, same

pytestmark = pytest.mark.gpu
KERNELS = ["blocked", "wave", "tuned", "fused", "simple"]


def _case(name):
    print("\n[case]", name, flush=True)


def small_psv(nt=300):
    return configs.psv_anomaly(nx=60, nz=50, nt=nt, fw=10, radius=8, src=(30, 5),
                               rec_range=(2, 58), f0=40.0)


def small_sh(nt=250):
    return configs.sh_layered(nx=96, nz=70, nt=nt, fw=12, f0=40.0, seed=3)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("isnap", [1, 7])
def test_psv_rtm_bitwise(kernel, isnap):
    _case("psv small kernel=%s isnap=%d" % (kernel, isnap))
    cfg = small_psv()
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    h = run_hip(kw, wav, isnap=isnap, kernel=kernel)
    o = run_oracle(kw, wav, isnap=isnap)
    assert_bitwise(h, o)
    assert np.abs(h["image"][2]).max() > 0
    assert same(h["fw"].isnaps_field.to_numpy(), o["isnaps"])


@pytest.mark.parametrize("kernel", KERNELS)
def test_sh_rtm_bitwise_border_receivers(kernel):
    """SH with a receiver on every surface cell, including i = 0 and nx-1,
    which are never stencil-updated but receive injected data backward."""
    _case("sh small kernel=%s" % kernel)
    cfg = small_sh()
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    h = run_hip(kw, wav, isnap=3, kernel=kernel)
    o = run_oracle(kw, wav, isnap=3)
    assert_bitwise(h, o)
    assert np.abs(h["image"][1]).max() > 0


@pytest.mark.parametrize("kernel", KERNELS)
def test_both_systems_duplicates_and_random_obs(kernel):
    """All three force components at once, two sources on one cell, duplicate
    receivers, and backward injection of unrelated data."""
    _case("both systems kernel=%s" % kernel)
    cfg = small_psv(nt=180)
    kw = cfg.kwargs(cfg.shots[0])
    kw["src_loc_field"] = np.array([[30, 5], [30, 5], [12, 40]], np.int32)
    kw["num_sources"] = 3
    rec = np.array([[5, 1], [5, 1], [59, 1], [0, 1], [33, 33], [20, 49]], np.int32)
    kw["recv_loc_field"], kw["num_receivers"] = rec, len(rec)
    rng = np.random.default_rng(7)
    w = configs.ricker(40.0, cfg.nt, cfg.dt)
    wav = (np.stack([w, -0.5 * w, 0.25 * w]), np.stack([0.3 * w, w, w]),
           np.stack([w, w, -w]))
    obs = tuple(rng.standard_normal((len(rec), cfg.nt)).astype(np.float32) * 1e-3
                for _ in range(3))
    h = run_hip(kw, wav, isnap=2, kernel=kernel, obs=obs)
    o = run_oracle(kw, wav, isnap=2, obs=obs)
    assert_bitwise(h, o)
    # border cells are never stencil-updated but carry injected data backward
    assert h["state_b"]["u"][0, 1] != 0 and h["state_b"]["v"][59, 1] != 0


@pytest.mark.parametrize("kernel", KERNELS)
def test_synthetic_source_record_matches_oracle(kernel):
    """record_synthetic_source=True: the backward propagator's record at the
    source cells (backward_modeling.py:462-470, row t of step it = nt-1-t),
    both systems, two sources on one cell and one elsewhere, against the
    oracle's synsrc bit for bit."""
    _case("synthetic source kernel=%s" % kernel)
    cfg = small_psv(nt=150)
    kw = cfg.kwargs(cfg.shots[0])
    kw["src_loc_field"] = np.array([[30, 5], [30, 5], [12, 40]], np.int32)
    kw["num_sources"] = 3
    w = configs.ricker(40.0, cfg.nt, cfg.dt)
    wav = (np.stack([w, -0.5 * w, 0.25 * w]), np.stack([0.3 * w, w, w]),
           np.stack([w, w, -w]))
    h = run_hip(kw, wav, isnap=3, kernel=kernel, bw_opts=dict(record_synthetic_source=True))
    o = run_oracle(kw, wav, isnap=3)
    assert_bitwise(h, o)
    for c, comp in enumerate("uvw"):
        got = getattr(h["bw"], "synsrc_" + comp).to_numpy()
        assert got.shape == (3, cfg.nt)
        assert same(got, o["synsrc"][c]), comp
        assert np.abs(got).max() > 0


@pytest.mark.parametrize("kernel", KERNELS)
def test_topography_surface_matrix(kernel):
    """surface_matrix multiplies syz, sxz, szz every step (forward_modeling.py:412-419)."""
    _case("topography kernel=%s" % kernel)
    cfg = small_psv(nt=200)
    kw = cfg.kwargs(cfg.shots[0])
    sm = np.ones((cfg.nx, cfg.nz), np.float32)
    for i in range(cfg.nx):
        sm[i, : 2 + (i // 10) % 3] = 0.0
    w = configs.ricker(40.0, cfg.nt, cfg.dt)
    wav = (np.zeros_like(w)[None], w[None], w[None])
    h = run_hip(kw, wav, isnap=4, kernel=kernel, surface=sm)
    o = run_oracle(kw, wav, isnap=4, surface=sm)
    assert_bitwise(h, o)


@pytest.mark.parametrize("kernel", KERNELS)
def test_nonfinite_codes(kernel):
    _case("finite-check codes kernel=%s" % kernel)
    cfg = small_psv(nt=300)
    kw = cfg.kwargs(cfg.shots[0])
    w = configs.ricker(40.0, cfg.nt, cfg.dt)
    bad = w.copy()
    bad[5] = np.nan
    z = np.zeros_like(w)
    for wav, sci in (((bad[None], z[None], w[None]), 100), ((bad[None], z[None], w[None]), 1)):
        h = run_hip(kw, wav, sci=sci, kernel=kernel, backward=False)
        o = run_oracle(kw, wav, sci=sci, backward=False)
        assert h["code_f"] == o["code_f"] and h["code_f"] != 0
        assert same(h["seis"][2][:, : o["it_f"] + 1], o["seis"][2][:, : o["it_f"] + 1])
    obs = [np.zeros((len(kw["recv_loc_field"]), cfg.nt), np.float32) for _ in range(3)]
    obs[2][0, cfg.nt - 1 - 7] = np.inf
    h = run_hip(kw, (z[None], z[None], w[None]), sci=1, kernel=kernel, obs=obs)
    o = run_oracle(kw, (z[None], z[None], w[None]), sci=1, obs=obs)
    assert h["code_b"] == o["code_b"] == 6


@pytest.mark.parametrize("kernel", KERNELS)
def test_zero_receivers_and_border_source(kernel):
    """No receivers at all (empty seismograms, nothing injected backward: the
    image stays zero), and a source on the grid border, where 1 / rho_w is
    +inf (forward_modeling.py:291-307), so the reference's w blows up and the
    forward returns code 3 at the first check: both as the oracle."""
    _case("zero receivers, border source kernel=%s" % kernel)
    cfg = small_psv(nt=60)
    kw = cfg.kwargs(cfg.shots[0])
    kw["recv_loc_field"], kw["num_receivers"] = np.zeros((0, 2), np.int32), 0
    wav = cfg.wavelets(cfg.shots[0])
    h = run_hip(kw, wav, isnap=2, kernel=kernel)
    o = run_oracle(kw, wav, isnap=2)
    assert_bitwise(h, o)
    assert all(np.asarray(x).shape == (0, cfg.nt) for x in h["seis"])
    assert not np.any(h["image"][2])
    kw["src_loc_field"] = np.array([[0, 1]], np.int32)
    h = run_hip(kw, wav, isnap=2, kernel=kernel, backward=False)
    o = run_oracle(kw, wav, isnap=2, backward=False)
    assert h["code_f"] == o["code_f"] == 3


@pytest.mark.parametrize("kernel", KERNELS)
def test_short_records(kernel):
    """Records of 1, 2 and 9 steps (fewer steps than one temporally blocked
    launch, and one more than K = 8), snapshots every step: as the oracle."""
    _case("short records kernel=%s" % kernel)
    for nt in (1, 2, 9):
        cfg = small_sh(nt=nt)
        kw = cfg.kwargs(cfg.shots[0])
        wav = cfg.wavelets(cfg.shots[0])
        h = run_hip(kw, wav, isnap=1, kernel=kernel)
        o = run_oracle(kw, wav, isnap=1)
        assert_bitwise(h, o)


@pytest.mark.parametrize("kernel", KERNELS)
def test_sponge_edge_cases(kernel):
    """The sponge the step kernels evaluate from the cell index, where the
    reference's strips meet unusually (forward_modeling.py:203-241): no
    sponge (absorbing_frame 0), a frame wider than half the grid (the right
    strip overwrites the left, nx < 2 FW), and a frame equal to the depth."""
    _case("sponge edge cases kernel=%s" % kernel)
    for nx, nz, fw in ((60, 50, 0), (30, 50, 20), (40, 20, 20)):
        cfg = configs.psv_anomaly(nx=nx, nz=nz, nt=70, fw=fw, radius=5, src=(nx // 2, 5),
                                  rec_range=(2, nx - 2), f0=40.0)
        kw = cfg.kwargs(cfg.shots[0])
        wav = cfg.wavelets(cfg.shots[0])
        h = run_hip(kw, wav, isnap=2, kernel=kernel)
        o = run_oracle(kw, wav, isnap=2)
        assert_bitwise(h, o)


def test_mxz_overflow_reproduced():
    """SI rock moduli overflow mu^4 in float32 (forward_modeling.py:271-276):
    the reference's shear stress turns NaN; the HIP path reports the same code."""
    _case("SI-unit overflow")
    cfg = small_psv(nt=220)
    mu, lam = configs.moduli(cfg.vp, cfg.vs, cfg.rho * 1000.0)
    kw = cfg.kwargs(cfg.shots[0])
    kw.update(mu_field=mu, lam_field=lam, rho_field=(cfg.rho * 1000.0).astype(np.float32))
    wav = cfg.wavelets(cfg.shots[0])
    h = run_hip(kw, wav, backward=False)
    o = run_oracle(kw, wav, backward=False)
    assert h["code_f"] == o["code_f"]


def test_prepare_and_sponge_bitwise():
    _case("medium preparation and sponge")
    import taichirtm_amd as trtm
    rng = np.random.default_rng(1)
    for nx, nz, fw in ((70, 66, 20), (7, 20, 5), (130, 65, 0)):
        mu = rng.uniform(1e6, 9e9, (nx, nz)).astype(np.float32)
        rho = rng.uniform(1.5, 2.6, (nx, nz)).astype(np.float32)
        lam = mu.copy()
        fm = trtm.ForwardModeling(nx=nx, nz=nz, dx=5.0, dz=5.0, nt=4, fs=2500.0, mu_field=mu,
                                  lam_field=lam, rho_field=rho, absorbing_frame=fw,
                                  src_loc_field=[[1, 1]], recv_loc_field=[[1, 1]],
                                  wavelet_u_field=np.zeros((1, 4)), wavelet_v_field=np.zeros((1, 4)),
                                  wavelet_w_field=np.zeros((1, 4)), num_sources=1, num_receivers=1)
        ref = oracle.prepare(mu, rho)
        for k, v in ref.items():
            assert same(getattr(fm, k).to_numpy(), v), k
        assert same(fm.absorb_coeff.to_numpy(), oracle.sponge(nx, nz, fw))


@pytest.mark.parametrize("kernel", KERNELS)
def test_callback_chunks_and_determinism(kernel):
    """A display callback splits the run into chunks; results must not change,
    and two runs must agree bit for bit."""
    _case("callback chunking kernel=%s" % kernel)
    cfg = small_psv(nt=123)
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    seen = []
    cb = lambda u, v, w, it, nx, nz, dx, dz: seen.append((it, float(np.abs(w).sum())))
    a = run_hip(kw, wav, isnap=10, kernel=kernel, callback=cb)
    b = run_hip(kw, wav, isnap=10, kernel=kernel)
    o = run_oracle(kw, wav, isnap=10)
    assert_bitwise(a, o)
    assert_bitwise(b, o)
    fw_its = [it for it, _ in seen[:13]]
    assert fw_its == list(range(0, 123, 10))


def test_psv_anomaly_full_size_parity():
    """Plain P-SV RTM, 200x200, 1500 steps, isnap = 1: HIP against the float32
    oracle (bitwise), and the float32 scheme's error against float64 stated."""
    _case("psv_anomaly 200x200x1500")
    cfg = configs.psv_anomaly()
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    h = run_hip(kw, wav, isnap=1)
    o = run_oracle(kw, wav, isnap=1)
    assert_bitwise(h, o)
    o64 = run_oracle(kw, wav, isnap=1, dtype=np.float64)
    e_seis = oracle.relative_l2(h["seis"][2], o64["seis"][2])
    e_img = oracle.relative_l2(h["image"][2], o64["image"][2])
    print("[fp64] seismogram w rel-L2 %.3e  image w rel-L2 %.3e" % (e_seis, e_img), flush=True)
    assert e_seis < 1e-4 and e_img < 1e-3


def test_psv_anomaly_full_size_bench_kernels_bitwise():
    """The 200x200x1500 P-SV RTM with the blocked shape the autotuner picks for
    it on MI355X in both directions (16 x 16 tiles, 1024 threads, 8 steps per
    launch; sources and receivers looked up once per launch) against the
    oracle bit for bit."""
    _case("psv_anomaly 200x200x1500 blocked 16x16x8 vs oracle")
    cfg = configs.psv_anomaly()
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    opts = dict(tile_x=16, tile_z=16, block_steps=8)
    h = run_hip(kw, wav, isnap=1, kernel="blocked", fw_opts=opts, bw_opts=opts)
    assert h["fw"].block_config() == (16, 16, 1024, 8)
    assert h["bw"].block_config() == (16, 16, 1024, 8)
    o = run_oracle(kw, wav, isnap=1)
    assert_bitwise(h, o)


def test_sh_layered_reduced_parity_and_float64_error():
    """SH 1200x600: bitwise against the oracle over 300 steps, and the float32
    scheme's error against the oracle in float64 there."""
    _case("sh_layered 1200x600x300")
    cfg = configs.sh_layered(nt=300)
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    h = run_hip(kw, wav, isnap=10)
    o = run_oracle(kw, wav, isnap=10)
    assert_bitwise(h, o)
    o64 = run_oracle(kw, wav, isnap=10, dtype=np.float64)
    e_seis = oracle.relative_l2(h["seis"][1], o64["seis"][1])
    e_img = oracle.relative_l2(h["image"][1], o64["image"][1])
    print("[fp64] SH 1200x600x300 seismogram v rel-L2 %.3e  image v rel-L2 %.3e"
          % (e_seis, e_img), flush=True)
    assert e_seis < 1e-4 and e_img < 1e-3


def test_sh_layered_full_length_bench_kernels_bitwise():
    """The single-shot bench workload at full length, 1200 x 600 x 4000: the
    blocked shapes the autotuner picks for it on MI355X (48 x 64 tiles with
    1024 threads and 8 steps per launch both ways; before the per-launch step
    schedule the backward took 32 x 48 tiles with 512 threads) against the
    oracle, bit for bit: seismograms, final states, the image (isnap 10, so
    the oracle keeps 400 snapshots) and the codes."""
    _case("sh_layered 1200x600x4000 blocked 48x64x8 / 48x64x8 and 32x48x8 vs oracle")
    cfg = configs.sh_layered()
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    o = run_oracle(kw, wav, isnap=10)
    for bw in ((48, 64, 1024, 8), (32, 48, 512, 8)):
        h = run_hip(kw, wav, isnap=10, kernel="blocked",
                    fw_opts=dict(tile_x=48, tile_z=64, block_steps=8),
                    bw_opts=dict(tile_x=bw[0], tile_z=bw[1], block_steps=bw[3]))
        assert h["fw"].block_config() == (48, 64, 1024, 8)
        assert h["bw"].block_config() == bw
        assert_bitwise(h, o)
        assert np.abs(h["image"][1]).max() > 0 and np.abs(h["seis"][1]).max() > 0


def test_sh_layered_full_length_image_error_against_float64():
    """Error of the float32 scheme at full length (1200 x 600 x 4000,
    isnap 10): the GPU's seismogram and image against the oracle run in
    float64."""
    _case("sh_layered 1200x600x4000 vs float64 oracle")
    cfg = configs.sh_layered()
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])
    h = run_hip(kw, wav, isnap=10, kernel="tuned")
    o64 = run_oracle(kw, wav, isnap=10, dtype=np.float64)
    e_seis = oracle.relative_l2(h["seis"][1], o64["seis"][1])
    e_img = oracle.relative_l2(h["image"][1], o64["image"][1])
    print("[fp64] SH 1200x600x4000 seismogram v rel-L2 %.3e  image v rel-L2 %.3e"
          % (e_seis, e_img), flush=True)
    assert e_seis < 1e-4 and e_img < 1e-3


BLOCK_CFGS = {"psv": [(16, 64, 2), (16, 64, 4), (16, 128, 4), (16, 128, 8), (32, 128, 4),
                      (48, 64, 4), (32, 64, 4), (24, 56, 4), (16, 32, 4), (16, 32, 8),
                      (16, 16, 8)],
              "sh": [(16, 64, 4), (16, 64, 8), (16, 128, 4), (16, 128, 8), (32, 128, 4),
                     (32, 128, 8), (24, 128, 8), (32, 96, 8), (48, 64, 8), (32, 48, 8),
                     (24, 64, 8), (16, 32, 4), (16, 32, 8), (16, 16, 8)],
              "both": [(16, 64, 2), (16, 64, 4), (16, 128, 4), (16, 32, 4)]}


@pytest.mark.parametrize("system,cfg", [(s, c) for s, cs in BLOCK_CFGS.items() for c in cs])
def test_blocked_configurations_bitwise(system, cfg):
    """Every compiled tile / time-block shape of the blocked kernel, with nt not
    a multiple of the block depth, duplicate sources and border receivers."""
    tx, tz, k = cfg
    _case("blocked %s tile_x=%d tile_z=%d block_steps=%d" % (system, tx, tz, k))
    c = small_psv(nt=97) if system != "sh" else small_sh(nt=97)
    kw = c.kwargs(c.shots[0])
    kw["src_loc_field"] = np.array([[30, 5], [30, 5]], np.int32)
    kw["num_sources"] = 2
    w = configs.ricker(40.0, c.nt, c.dt)
    z = np.zeros_like(w)
    comps = {"psv": (w, z, w), "sh": (z, w, z), "both": (w, w, w)}[system]
    wav = tuple(np.stack([x, 0.5 * x]) for x in comps)
    hk = dict(kw, tile_x=tx, tile_z=tz, block_steps=k)
    h = run_hip(hk, wav, isnap=3, kernel="blocked")
    o = run_oracle(kw, wav, isnap=3)
    assert_bitwise(h, o)


@pytest.mark.parametrize("system,shape", [("psv", (16, 16, 8)), ("psv", (32, 128, 4)),
                                          ("sh", (16, 16, 8)), ("sh", (32, 48, 8)),
                                          ("psv", None)])
def test_irregular_isnaps_bitwise(system, shape):
    """Backward imaging with an isnaps table that disagrees with the forward's
    own snapshot schedule (entries of another time, or -1): a step images only
    where its slot's entry equals its time index (backward_modeling.py's
    isnap check).  The blocked kernels then read the table instead of taking
    the slots of the regular schedule."""
    _case("irregular isnaps %s shape=%s" % (system, shape))
    c = small_psv(nt=97) if system == "psv" else small_sh(nt=97)
    kw = c.kwargs(c.shots[0])
    wav = c.wavelets(c.shots[0])
    isnap = 2
    kernel, hk = "fused", dict(kw)
    if shape:
        kernel = "blocked"
        hk.update(tile_x=shape[0], tile_z=shape[1], block_steps=shape[2])
    h = run_hip(hk, wav, isnap=isnap, kernel=kernel, backward=False)
    m = oracle.Model(kw["mu_field"], kw["lam_field"], kw["rho_field"], kw["dx"], kw["dz"],
                     kw["fs"], kw["nt"], kw["src_loc_field"], kw["recv_loc_field"],
                     kw["absorbing_frame"], isnap=isnap, sci=100)
    fo = m.forward(wav, save=True)
    snaps = fo["snaps"]
    assert same(h["fw"].isnaps_field.to_numpy()[:fo["num_snaps"]], snaps[3][:fo["num_snaps"]])
    isn = np.array(snaps[3][:fo["num_snaps"]], np.int32)
    isn[::3] = -1          # no image at these slots
    isn[4] += isnap        # the time of another slot
    isn[7] = isn[6]
    bo = m.backward(fo["seis"], snaps[:3], isn, fo["num_snaps"], isnap)
    regular = m.backward(fo["seis"], snaps[:3], snaps[3], fo["num_snaps"], isnap)

    import taichirtm_amd as trtm
    fw = h["fw"]
    seis = fw.get_seismogram()
    bw = trtm.BackwardModeling(obsdata_u_field=seis[0], obsdata_v_field=seis[1],
                               obsdata_w_field=seis[2], kernel=kernel,
                               **{k: v for k, v in hk.items() if k != "kernel"})
    code = bw.run_calc(fw.u_save_field, fw.v_save_field, fw.w_save_field, isn, fw.num_snaps,
                       isnap)
    img = bw.get_results()
    assert code == bo["code"]
    comps = (0, 2) if system == "psv" else (1,)
    for k in comps:
        assert same(img[k], bo["result"][k]), "image %d" % k
        assert not same(bo["result"][k], regular["result"][k])


@pytest.mark.parametrize("system", ["psv", "sh"])
def test_small_tiles_long_injection_lists_bitwise(system):
    """The 16 x 16 tiles, 8 steps, with receivers on every cell of three rows
    under the surface: a tile's region then holds 96 receiver groups, more
    than its per-launch table takes, so the backward walks its list every
    step (the loop variant with the per-step list walks)."""
    _case("blocked 16x16x8 %s, 168 receivers in three rows" % system)
    c = small_psv(nt=97) if system == "psv" else small_sh(nt=97)
    kw = c.kwargs(c.shots[0])
    nx = kw["mu_field"].shape[0]
    rec = np.array([[i, j] for j in (1, 2, 3) for i in range(2, nx - 2)], np.int32)
    kw["recv_loc_field"] = rec
    kw["num_receivers"] = len(rec)
    wav = c.wavelets(c.shots[0])
    opts = dict(tile_x=16, tile_z=16, block_steps=8)
    h = run_hip(kw, wav, isnap=2, kernel="blocked", fw_opts=opts, bw_opts=opts)
    assert h["bw"].block_config() == (16, 16, 1024, 8)
    o = run_oracle(kw, wav, isnap=2)
    assert_bitwise(h, o)


TOPO_CFGS = [("psv", (16, 16, 8)), ("psv", (24, 56, 4)), ("psv", (48, 64, 4)),
             ("psv", (32, 128, 4)), ("sh", (16, 16, 8)), ("sh", (48, 64, 8)),
             ("sh", (32, 48, 8)), ("both", (16, 32, 4))]


@pytest.mark.parametrize("system,cfg,long_lists",
                         [(s, c, False) for s, c in TOPO_CFGS] +
                         [("psv", (16, 16, 8), True), ("sh", (16, 16, 8), True)])
def test_blocked_configurations_topography_bitwise(system, cfg, long_lists):
    """The step loop compiled for a surface matrix (forward_modeling.py:412-419),
    for the shape families of the blocked kernel: one or two cells per thread
    (per-launch tables, and with receivers on three full rows the per-step
    list walks), branch-free up to four cells per thread, per-cell branches
    beyond; SH and both systems too (syz is masked as well)."""
    tx, tz, k = cfg
    _case("blocked topography %s %s long_lists=%s" % (system, cfg, long_lists))
    c = small_psv(nt=97) if system != "sh" else small_sh(nt=97)
    kw = c.kwargs(c.shots[0])
    nx, nz = kw["mu_field"].shape
    sm = np.ones((nx, nz), np.float32)
    for i in range(nx):
        sm[i, : 2 + (i // 7) % 3] = 0.0
    rows = (4, 5, 6) if long_lists else (5,) if system == "sh" else ()
    if system == "sh":
        # source and receivers below the masked rows: SH rows under a zero
        # syz on both sides exchange nothing with the rows below
        kw["src_loc_field"] = np.array([[nx // 2, 8]], np.int32)
    if rows:
        rec = np.array([[i, j] for j in rows for i in range(2, nx - 2)], np.int32)
        kw["recv_loc_field"] = rec
        kw["num_receivers"] = len(rec)
    w = configs.ricker(40.0, c.nt, c.dt)
    z = np.zeros_like(w)
    comps = {"psv": (w, z, w), "sh": (z, w, z), "both": (w, w, w)}[system]
    wav = tuple(x[None] for x in comps)
    opts = dict(tile_x=tx, tile_z=tz, block_steps=k)
    h = run_hip(kw, wav, isnap=3, kernel="blocked", surface=sm, fw_opts=opts, bw_opts=opts)
    assert h["bw"].block_config()[0:2] == (tx, tz) and h["bw"].block_config()[3] == k
    o = run_oracle(kw, wav, isnap=3, surface=sm)
    assert_bitwise(h, o)
    comp = 2 if system != "sh" else 1
    assert np.abs(o["image"][comp]).max() > 0
    assert not same(o["image"][comp], run_oracle(kw, wav, isnap=3)["image"][comp])


WAVE_CFGS = {"psv": [(16, 4), (8, 4), (16, 2), (8, 2)],
             "sh": [(16, 4), (8, 4), (24, 4), (16, 2)],
             "both": [(8, 4), (8, 2)]}


@pytest.mark.parametrize("system,cfg", [(s, c) for s, cs in WAVE_CFGS.items() for c in cs])
def test_wave_configurations_bitwise(system, cfg):
    """Every compiled shape of the one-wavefront-per-tile kernel, with nt not a
    multiple of the block depth, duplicate sources, border receivers and a grid
    that is not a multiple of the tile (partial tiles at the bottom and right)."""
    tx, k = cfg
    _case("wave %s tile_x=%d block_steps=%d" % (system, tx, k))
    c = small_psv(nt=97) if system != "sh" else small_sh(nt=97)
    kw = c.kwargs(c.shots[0])
    kw["src_loc_field"] = np.array([[30, 5], [30, 5]], np.int32)
    kw["num_sources"] = 2
    w = configs.ricker(40.0, c.nt, c.dt)
    z = np.zeros_like(w)
    comps = {"psv": (w, z, w), "sh": (z, w, z), "both": (w, w, w)}[system]
    wav = tuple(np.stack([x, 0.5 * x]) for x in comps)
    hk = dict(kw, tile_x=tx, tile_z=64 - 2 * k, block_steps=k)
    h = run_hip(hk, wav, isnap=3, kernel="wave")
    o = run_oracle(kw, wav, isnap=3)
    assert_bitwise(h, o)


@pytest.mark.parametrize("workload", ["sh_layered", "psv_faulted", "psv_salt"])
def test_full_size_forward_error_against_float64(workload):
    """The float32 scheme's error against its float64 run at full size
    (SH 1200x600x4000; the first shot of the 2000x1000 P-SV survey; the first
    shot of the 4000x2000x8000 salt model, snapshots every 100 steps): the
    forward seismograms of ForwardModeling (blocked kernels) against the FWI
    forward step run in double, which is the same scheme (its float32 records
    equal ForwardModeling's bit for bit)."""
    from taichirtm_amd.batch import ShotSpec
    from taichirtm_amd.fwi import ElasticFWI
    import torch
    _case("float64 error %s" % workload)
    if workload == "sh_layered":
        cfg = configs.sh_layered()
    elif workload == "psv_faulted":
        cfg = configs.psv_faulted(n_shots=1)
    else:
        cfg = configs.psv_salt(n_shots=1)[0]
    shot = cfg.shots[0]
    kw = cfg.kwargs(shot)
    wav = cfg.wavelets(shot)
    h = run_hip(kw, wav, isnap=100 if workload == "psv_salt" else 10, kernel="auto",
                backward=False)
    spec = ShotSpec(np.array([shot.src], np.int32), tuple(np.asarray(w, np.float64) for w in wav),
                    np.asarray(shot.rec, np.int32))
    p = ElasticFWI(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, [spec], precision="float64",
                   checkpoint_every=shot.nt)
    m = [torch.as_tensor(np.asarray(x, np.float64), device="cuda") for x in (cfg.vp, cfg.vs, cfg.rho)]
    rec = p.forward(*m)[0].cpu().numpy()
    comps = [1] if workload == "sh_layered" else [0, 2]
    for c in comps:
        e = oracle.relative_l2(h["seis"][c], rec[c])
        print("[fp64] %s %dx%dx%d seismogram %s rel-L2 %.3e"
              % (workload, cfg.nx, cfg.nz, shot.nt, "uvw"[c], e), flush=True)
        assert e < 1e-3
