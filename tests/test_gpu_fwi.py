"""FWI gradient through the exact discrete adjoint.

Pins: Richardson-extrapolated central finite differences in float64 along 8
seeded directions (the gradient of J = 1/2 |d(m) - d_obs|^2 with respect to
(vp, vs, rho)), the float32 gradient against the float64 one, invariance to the
checkpoint interval, the autograd.Function against the explicit call, and
bitwise equality of the FWI forward records with ForwardModeling's.

The anomaly sits 45 m deep so that its P and S reflections arrive inside the
72 ms record: the misfit is then large next to the third-order terms of J, and
the finite differences resolve the gradient to ~1e-11 (a deep anomaly whose
reflections never arrive leaves J ~ 1e-16 and a difference quotient dominated
by curvature)."""
import numpy as np
import pytest
import torch

from taichirtm_amd import configs
from taichirtm_amd.batch import ShotSpec
from taichirtm_amd.fwi import ElasticFWI, Seismograms, l2_misfit_and_gradient
from tests.parity import run_hip, same

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case]", name, flush=True)


def problem(system="psv", nx=56, nz=44, nt=180, n_shots=2, seed=0):
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    vp = (2000.0 + 12.0 * jj + 60.0 * np.sin(ii / 7.0)).astype(np.float64)
    vs = vp / 1.8
    rho = 0.31 * vp ** 0.25
    true = [x.copy() for x in (vp, vs, rho)]
    blob = np.exp(-(((ii - nx / 2) / 6.0) ** 2 + ((jj - 9.0) / 4.0) ** 2))
    true[0] = vp * (1 + 0.06 * blob)
    true[1] = vs * (1 - 0.05 * blob)
    dt = 4e-4
    shots = []
    for q in range(n_shots):
        w = configs.ricker(40.0, nt, dt).astype(np.float64)[None]
        z = np.zeros_like(w)
        wav = (w, z, 0.5 * w) if system == "psv" else (z, w, z) if system == "sh" else (w, w, w)
        src = np.array([[12 + q * 25, 2]], np.int32)
        rec = np.stack([np.arange(2, nx - 2), np.ones(nx - 4, int)], 1).astype(np.int32)
        shots.append(ShotSpec(src, wav, rec))
    return (nx, nz, 5.0, 5.0, 1.0 / dt, 8), shots, (vp, vs, rho), true


def misfit(p, vp, vs, rho, obs):
    """J = 1/2 |d(m) - d_obs|^2 from the forward records alone."""
    recs = p.forward(vp, vs, rho)
    return 0.5 * sum(float(((r - o) ** 2).sum()) for r, o in zip(recs, obs))


def _t(x, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device="cuda")


@pytest.mark.parametrize("system", ["psv", "sh"])
def test_gradient_matches_central_differences_fp64(system):
    _case("FWI FD check %s" % system)
    grid, shots, m0, mt = problem(system)
    p = ElasticFWI(*grid, shots, precision="float64", checkpoint_every=23)
    obs = p.forward(*[_t(x) for x in mt])
    obs = [o.clone() for o in obs]
    m0t = [_t(x) for x in m0]
    J, (gvp, gvs, grho) = l2_misfit_and_gradient(p, *m0t, obs)
    assert J > 0
    rng = np.random.default_rng(42)
    nx, nz = grid[0], grid[1]
    from scipy.ndimage import gaussian_filter
    worst = 0.0
    for k in range(8):
        d = [gaussian_filter(rng.standard_normal((nx, nz)), 3.0) for _ in range(3)]
        d = [x / np.abs(x).max() for x in d]
        scale = [m0[0].mean(), m0[1].mean(), m0[2].mean()]
        dm = [_t(x * s) for x, s in zip(d, scale)]
        if system == "sh":  # vp does not enter the SH system
            dm[0] = torch.zeros_like(dm[0])

        def central(eps):
            Jp = misfit(p, *[a + eps * b for a, b in zip(m0t, dm)], obs)
            Jm = misfit(p, *[a - eps * b for a, b in zip(m0t, dm)], obs)
            return (Jp - Jm) / (2 * eps)

        # Richardson: the eps^2 term of the central difference cancels
        fd = (4 * central(0.5e-5) - central(1e-5)) / 3
        an = float(sum((g * b).sum() for g, b in zip((gvp, gvs, grho), dm)))
        rel = abs(fd - an) / max(abs(fd), 1e-300)
        print("[fd] dir %d: fd %.10e adjoint %.10e rel %.2e" % (k, fd, an, rel), flush=True)
        worst = max(worst, rel)
    assert worst < 1e-8


def test_float32_gradient_close_to_float64():
    _case("FWI float32 vs float64 gradient")
    grid, shots, m0, mt = problem("psv")
    out = {}
    for prec in ("float64", "float32"):
        p = ElasticFWI(*grid, shots, precision=prec, checkpoint_every=40)
        dtype = torch.float64 if prec == "float64" else torch.float32
        obs = [o.clone() for o in p.forward(*[_t(x, dtype) for x in mt])]
        out[prec] = l2_misfit_and_gradient(p, *[_t(x, dtype) for x in m0], obs)[1]
    for a, b in zip(out["float32"], out["float64"]):
        rel = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        print("[fp32 vs fp64] gradient rel-L2 %.2e" % rel, flush=True)
        assert rel < 1e-3


@pytest.mark.parametrize("workload", ["fwi", "salt"])
def test_full_size_gradient_error_against_float64(workload):
    """The float32 path against the float64 run of the same scheme at full
    size, one shot: the first of the 600x300x2500 FWI problem, and the middle
    one of the 4000x2000x8000 salt survey (the first shot's records never
    reach the salt body, so its residual is zero).  Observed data come from
    the true model, misfit and gradient from the starting model.

    Pinned: the records (1e-4) and the float32 adjoint given the float64
    residual (1e-4; l2_misfit_and_gradient with records_from).  Stated: the
    residual's own float32 error and the end-to-end gradient error it causes
    (the residual is a small difference of two large record sets, so float32
    rounding of the forward runs, ~1e-6 of the records, is a large part of
    it)."""
    _case("FWI full-size float32 vs float64 %s" % workload)
    true, start = configs.psv_fwi(n_shots=1) if workload == "fwi" else configs.psv_salt(n_shots=32)
    s = true.shots[len(true.shots) // 2]  # salt: a shot above the salt body
    size = "%dx%dx%d" % (true.nx, true.nz, s.nt)

    def rel(a, b):
        return float(torch.linalg.norm((a.double() - b.double()).ravel())
                     / torch.linalg.norm(b.double().ravel()))

    run = {}
    for prec, dtype, npt in (("float64", torch.float64, np.float64),
                             ("float32", torch.float32, np.float32)):
        shot = ShotSpec(np.array([s.src], np.int32),
                        tuple(np.asarray(w, np.float32).astype(npt) for w in true.wavelets(s)),
                        np.asarray(s.rec, np.int32))
        p = ElasticFWI(true.nx, true.nz, true.dx, true.dz, true.fs, true.fw, [shot],
                       precision=prec, history_bytes=0)
        obs = p.forward(*[_t(x, dtype) for x in (true.vp, true.vs, true.rho)])[0].clone()
        syn = p.forward(*[_t(x, dtype) for x in (start.vp, start.vs, start.rho)])[0].clone()
        run[prec] = (p, obs, syn, syn - obs)
    p64, obs64, syn64, res64 = run["float64"]
    p32, obs32, syn32, res32 = run["float32"]
    e_rec = max(rel(obs32, obs64), rel(syn32, syn64))
    e_res = rel(res32, res64)
    print("[fp32 vs fp64] %s %s records %.2e, residual %.2e (|residual| / |records| %.2e)"
          % (workload, size, e_rec, e_res, float(res64.norm() / obs64.norm())),
          flush=True)
    g64 = [x.double().clone() for x in p64.gradient([res64])]
    g32 = [x.double().clone() for x in p32.gradient([res32])]
    # float32 adjoint, residual formed from float64 records (rounded once)
    g32s = [x.double().clone() for x in l2_misfit_and_gradient(
        p32, *[_t(x, torch.float32) for x in (start.vp, start.vs, start.rho)], [obs64],
        records_from=p64)[1]]
    assert e_rec < 1e-4
    for name, a, b, c in zip(("vp", "vs", "rho"), g32, g64, g32s):
        same_res, end = rel(c, b), rel(a, b)
        print("[fp32 vs fp64] %s %s gradient %s: float64 residual %.2e, end to end %.2e"
              % (workload, size, name, same_res, end), flush=True)
        assert same_res < 1e-4
    # float32 end to end with the residual propagated as a scattered field
    # (reference = the true model): no float64 run
    m32 = [_t(x, torch.float32) for x in (start.vp, start.vs, start.rho)]
    ref32 = [_t(x, torch.float32) for x in (true.vp, true.vs, true.rho)]
    J_s, g_s = l2_misfit_and_gradient(p32, *m32, reference=ref32)
    res_s = p32.residual(*m32, ref32)[0]
    e_res_s = rel(res_s, res64)
    print("[fp32 scattered vs fp64] %s %s residual %.2e" % (workload, size, e_res_s), flush=True)
    errs = []
    for name, a, b in zip(("vp", "vs", "rho"), g_s, g64):
        errs.append(rel(a, b))
        print("[fp32 scattered vs fp64] %s %s gradient %s end to end %.2e"
              % (workload, size, name, errs[-1]), flush=True)
    assert e_res_s < 1e-3 and max(errs) < 1e-3


@pytest.mark.parametrize("system", ["psv", "sh", "both"])
def test_scattered_residual_against_float64_difference(system):
    """ElasticFWI.residual (the scattered-field forward): this model's
    checkpoints equal the plain forward's bit for bit; the residual is zero
    exactly when the reference is the model itself; against the float64
    difference of the two models' records it is far closer than the float32
    difference of two float32 record sets; the gradient from it matches the
    float64 gradient."""
    _case("FWI scattered residual %s" % system)
    grid, shots, m0, mt = problem(system, nt=200)

    def rel(a, b):
        return float(torch.linalg.norm((a.double() - b.double()).ravel())
                     / torch.linalg.norm(b.double().ravel()))

    p64 = ElasticFWI(*grid, shots, precision="float64", checkpoint_every=40, history_bytes=0)
    obs64 = [o.clone() for o in p64.forward(*[_t(x) for x in mt])]
    J64, g64 = l2_misfit_and_gradient(p64, *[_t(x) for x in m0], obs64)
    syn64 = [o.clone() for o in p64.forward(*[_t(x) for x in m0])]
    res64 = [a - b for a, b in zip(syn64, obs64)]

    p = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=40, history_bytes=0,
                   kernel="blocked")
    m32 = [_t(x, torch.float32) for x in m0]
    r32 = [_t(x, torch.float32) for x in mt]
    obs32 = [o.clone() for o in p.forward(*r32)]
    syn32 = [o.clone() for o in p.forward(*m32)]
    ck_plain = p._ck.clone()
    res_s = p.residual(*m32, r32)
    assert torch.equal(p._ck, ck_plain)
    e_plain = max(rel(a - b, c) for a, b, c in zip(syn32, obs32, res64))
    e_scat = max(rel(a, c) for a, c in zip(res_s, res64))
    print("[scattered] %s residual vs float64: plain float32 difference %.2e, scattered %.2e"
          % (system, e_plain, e_scat), flush=True)
    assert e_scat < 0.5 * e_plain
    zero = p.residual(*m32, m32)
    assert all(float(z.abs().max()) == 0.0 for z in zero)
    J, g = l2_misfit_and_gradient(p, *m32, reference=r32)
    assert abs(J - J64) <= 1e-4 * J64
    for name, a, b in zip(("vp", "vs", "rho"), g, g64):
        if system == "sh" and name == "vp":  # SH does not see vp: both are zero
            assert float(a.abs().max()) == 0.0 and float(b.abs().max()) == 0.0
            continue
        e = rel(a, b)
        print("[scattered] %s gradient %s vs float64 %.2e" % (system, name, e), flush=True)
        assert e < 1e-3


def test_scattered_residual_same_bits_for_every_shape():
    """Every compiled scattered-kernel shape that fits gives the same residual
    bits (the same expressions per cell in the same order)."""
    _case("FWI scattered shapes")
    grid, shots, m0, mt = problem("psv", nx=90, nt=97, n_shots=3)
    m32 = [_t(x, torch.float32) for x in m0]
    r32 = [_t(x, torch.float32) for x in mt]
    ref = None
    from taichirtm_amd import _lib, autotune
    listed = [c[1:] for c in autotune.fwi_candidates(_lib.PSV, "scattered")]
    # every shape the library lists for the pass (the tuner tries them all)
    assert (48, 48, 8) in listed and (24, 64, 4) in listed
    for tx, tz, k in listed:
        p = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=30, history_bytes=0,
                       kernel="blocked", tile_x=tx, tile_z=tz, block_steps=k)
        res = torch.stack(p.residual(*m32, r32))
        if ref is None:
            ref = res
            assert float(res.abs().max()) > 0
        else:
            assert torch.equal(res, ref), (tx, tz, k)


def test_autograd_with_reference_model():
    """Seismograms.apply(..., reference=...) outputs the residual; the
    gradient of 1/2 |output|^2 equals l2_misfit_and_gradient's."""
    _case("FWI autograd, scattered residual")
    grid, shots, m0, mt = problem("psv", nt=120)
    p = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=30, kernel="blocked")
    params = [_t(x, torch.float32).requires_grad_(True) for x in m0]
    ref = [_t(x, torch.float32) for x in mt]
    d = Seismograms.apply(*params, p, None, ref)
    (0.5 * (d ** 2).sum()).backward()
    J, g = l2_misfit_and_gradient(p, *[x.detach() for x in params], reference=ref)
    for a, b in zip(params, g):
        assert torch.equal(a.grad, b.to(torch.float32))


def test_gradient_independent_of_checkpoint_interval():
    _case("FWI checkpoint interval invariance")
    grid, shots, m0, mt = problem("psv", nt=150)
    res = []
    for L in (7, 50, 150):
        p = ElasticFWI(*grid, shots, precision="float64", checkpoint_every=L, history_bytes=0)
        obs = [o.clone() for o in p.forward(*[_t(x) for x in mt])]
        res.append(l2_misfit_and_gradient(p, *[_t(x) for x in m0], obs)[1])
    for r in res[1:]:
        for a, b in zip(r, res[0]):
            assert torch.equal(a, b)


def test_autograd_function_matches_explicit_gradient():
    _case("FWI autograd")
    grid, shots, m0, mt = problem("psv", nt=140)
    p = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=30)
    obs = torch.stack(p.forward(*[_t(x, torch.float32) for x in mt])).clone()
    params = [_t(x, torch.float32).requires_grad_(True) for x in m0]
    d = Seismograms.apply(*params, p)
    loss = 0.5 * ((d - obs) ** 2).sum()
    loss.backward()
    J, g = l2_misfit_and_gradient(p, *[x.detach() for x in params], list(obs))
    assert abs(float(loss) - J) <= 1e-5 * J
    for a, b in zip(params, g):
        assert torch.equal(a.grad, b.to(torch.float32))


def test_autograd_refuses_a_problem_run_forward_again():
    """Two applications of one problem before backward (a line-search trial,
    say): the first loss's backward would read the checkpoints of the second
    model, so it raises; a fresh application then gives the right gradient."""
    _case("FWI autograd, stale forward")
    grid, shots, m0, mt = problem("psv", nt=120)
    p = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=30)
    obs = torch.stack(p.forward(*[_t(x, torch.float32) for x in mt])).clone()
    params = [_t(x, torch.float32).requires_grad_(True) for x in m0]
    d1 = Seismograms.apply(*params, p)
    with torch.no_grad():
        Seismograms.apply(*[(1.01 * x).detach() for x in params], p)
    loss = 0.5 * ((d1 - obs) ** 2).sum()
    with pytest.raises(RuntimeError, match="ran forward again"):
        loss.backward()
    d = Seismograms.apply(*params, p)
    (0.5 * ((d - obs) ** 2).sum()).backward()
    J, g = l2_misfit_and_gradient(p, *[x.detach() for x in params], list(obs))
    for a, b in zip(params, g):
        assert torch.equal(a.grad, b.to(torch.float32))


def test_autograd_with_float64_records_matches_explicit_gradient():
    """Seismograms with records_from: float64 records out, the float32
    adjoint on the float64 residual; equals l2_misfit_and_gradient with the
    same records_from, and is closer to the float64 gradient than the plain
    float32 path."""
    _case("FWI autograd, float64 records")
    grid, shots, m0, mt = problem("psv", nt=140)
    p = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=30)
    p64 = ElasticFWI(*grid, shots, precision="float64", checkpoint_every=30)
    obs = torch.stack(p64.forward(*[_t(x) for x in mt])).clone()
    params = [_t(x, torch.float32).requires_grad_(True) for x in m0]
    d = Seismograms.apply(*params, p, p64)
    assert d.dtype == torch.float64
    loss = 0.5 * ((d - obs) ** 2).sum()
    loss.backward()
    J, g = l2_misfit_and_gradient(p, *[x.detach() for x in params], list(obs), records_from=p64)
    assert abs(float(loss) - J) <= 1e-12 * J
    for a, b in zip(params, g):
        assert torch.equal(a.grad, b.to(torch.float32))
    g64 = l2_misfit_and_gradient(p64, *[_t(x) for x in m0], list(obs))[1]
    g32 = l2_misfit_and_gradient(p, *[_t(x, torch.float32) for x in m0], list(obs))[1]
    for a, b, c in zip(g, g64, g32):
        e_mixed = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        e_f32 = float(torch.linalg.norm(c - b) / torch.linalg.norm(b))
        print("[fwi] gradient rel-L2 to float64: float64 records %.2e, float32 %.2e"
              % (e_mixed, e_f32), flush=True)
        assert e_mixed <= e_f32 and e_mixed < 1e-4
    with pytest.raises(ValueError, match="float64"):
        Seismograms.apply(*params, p, p)


def test_fwi_forward_records_equal_forward_modeling():
    _case("FWI forward vs ForwardModeling")
    grid, shots, m0, mt = problem("psv", nt=120, n_shots=1)
    nx, nz, dx, dz, fs, fw = grid
    vp, vs, rho = [np.asarray(x, np.float32) for x in m0]
    mu, lam = configs.moduli(vp, vs, rho)
    s = shots[0]
    kw = dict(nx=nx, nz=nz, dx=dx, dz=dz, nt=s.nt, fs=fs, mu_field=mu, lam_field=lam,
              rho_field=rho, absorbing_frame=fw, src_loc_field=s.src_loc,
              recv_loc_field=s.rec_loc, num_sources=1, num_receivers=len(s.rec_loc))
    h = run_hip(kw, tuple(np.asarray(w, np.float32) for w in s.wavelets), backward=False)
    p = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=33)
    rec = p.forward(*[_t(x, torch.float32) for x in (vp, vs, rho)])[0].cpu().numpy()
    for c in range(3):
        assert same(rec[c], h["seis"][c])


BLOCK_SHAPES = [(24, 64, 4), (32, 64, 4), (16, 64, 4), (16, 64, 8), (32, 64, 2), (16, 64, 2),
                (16, 48, 4), (24, 32, 4), (20, 40, 4)]


def _uneven(shots, nts, spreads):
    """Shots with their own record lengths and receiver spreads."""
    out = []
    for s, nt, (a, b) in zip(shots, nts, spreads):
        out.append(ShotSpec(s.src_loc, tuple(np.asarray(w)[:, :nt] for w in s.wavelets),
                            np.asarray(s.rec_loc)[a:b]))
    return out


def _records_and_gradient(grid, shots, m0, mt, adjoint_shape=None, **kw):
    kw.setdefault("history_bytes", 0)  # every segment recomputed unless asked otherwise
    p = ElasticFWI(*grid, shots, precision="float32", **kw)
    if adjoint_shape is not None:
        p.shapes["adjoint"] = adjoint_shape
    obs = [o.clone() for o in p.forward(*[_t(x, torch.float32) for x in mt])]
    recs = [r.clone() for r in p.forward(*[_t(x, torch.float32) for x in m0])]
    res = [r - o for r, o in zip(recs, obs)]
    g = p.gradient(res)
    return recs, g, {k: v.clone() for k, v in p.coefficient_gradient.items()}


FORWARD_ONLY_SHAPES = [(32, 64, 8), (24, 64, 8), (48, 64, 8), (48, 48, 8), (40, 56, 8)]


@pytest.mark.parametrize("shape", FORWARD_ONLY_SHAPES, ids=lambda s: "%dx%dk%d" % s)
@pytest.mark.parametrize("system", ["psv", "sh", "both"])
def test_forward_only_shapes_bitwise_equal_step_kernels(system, shape):
    """Shapes listed for the forward passes only (their adjoint would exceed
    the LDS): checkpointing and history forward with them, the adjoint with
    the default shape; records and gradients equal the per-step kernels' bit
    for bit."""
    _case("FWI forward-only shape %s %s" % (system, shape))
    from taichirtm_amd import _lib, autotune
    systems = {"psv": _lib.PSV, "sh": _lib.SH, "both": _lib.PSV | _lib.SH}[system]
    if ("fwi",) + tuple(shape) not in autotune.fwi_candidates(systems, "forward"):
        pytest.skip("shape exceeds the LDS of a workgroup for these wave systems")
    assert ("fwi",) + tuple(shape) not in autotune.fwi_candidates(systems, "adjoint")
    grid, shots, m0, mt = problem(system, nt=130, n_shots=3)
    shots = _uneven(shots, (130, 97, 121), ((0, 52), (5, 30), (20, 45)))
    shots[2] = ShotSpec(np.array([[1, 1]], np.int32), shots[2].wavelets, shots[2].rec_loc)
    ref = _records_and_gradient(grid, shots, m0, mt, kernel="step", checkpoint_every=29)
    p = ElasticFWI(*grid, shots, precision="float32", kernel="blocked", checkpoint_every=29,
                   history_bytes=0)
    p.shapes["forward"] = p.shapes["history"] = shape
    obs = [o.clone() for o in p.forward(*[_t(x, torch.float32) for x in mt])]
    recs = [r.clone() for r in p.forward(*[_t(x, torch.float32) for x in m0])]
    g = p.gradient([r - o for r, o in zip(recs, obs)])
    for a, b in zip(recs, ref[0]):
        assert torch.equal(a, b)
    for a, b in zip(g, ref[1]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape", BLOCK_SHAPES, ids=lambda s: "%dx%dk%d" % s)
@pytest.mark.parametrize("system", ["psv", "sh", "both"])
def test_blocked_kernels_bitwise_equal_step_kernels(system, shape):
    """K steps per launch (LDS tiles, halo recomputed) against one step per
    launch: records, coefficient gradients and (vp, vs, rho) gradients bit for
    bit, for 3 shots of uneven record length and receiver spread, a checkpoint
    interval that is no multiple of K, and a source next to the grid edge."""
    _case("FWI blocked %s %s" % (system, shape))
    from taichirtm_amd import _lib, autotune
    systems = {"psv": _lib.PSV, "sh": _lib.SH, "both": _lib.PSV | _lib.SH}[system]
    if ("fwi",) + tuple(shape) not in autotune.fwi_candidates(systems, "forward"):
        pytest.skip("shape exceeds the LDS of a workgroup for these wave systems")
    # a shape whose adjoint (with its stress windows) exceeds the LDS runs the
    # forward passes; the adjoint takes the library default
    adj = None if ("fwi",) + tuple(shape) in autotune.fwi_candidates(systems, "adjoint") \
        else (0, 0, 0)
    grid, shots, m0, mt = problem(system, nt=130, n_shots=3)
    shots = _uneven(shots, (130, 97, 121), ((0, 52), (5, 30), (20, 45)))
    shots[2] = ShotSpec(np.array([[1, 1]], np.int32), shots[2].wavelets, shots[2].rec_loc)
    tx, tz, k = shape
    ref = _records_and_gradient(grid, shots, m0, mt, kernel="step", checkpoint_every=29)
    got = _records_and_gradient(grid, shots, m0, mt, kernel="blocked", tile_x=tx, tile_z=tz,
                                block_steps=k, checkpoint_every=29, adjoint_shape=adj)
    for a, b in zip(got[0], ref[0]):
        assert torch.equal(a, b)
    for a, b in zip(got[1], ref[1]):
        assert torch.equal(a, b)
    for key in ref[2]:
        assert torch.equal(got[2][key], ref[2][key]), key


INNER_SHAPES = [(16, 48, 4), (24, 32, 4), (20, 40, 4), (32, 64, 4), (16, 64, 8)]


def _inner_tiles(nx, nz, fw, shape, points):
    """Tiles whose K-cell region lies inside the grid interior, outside the
    sponge strips and holds none of the given source / receiver cells (no
    per-cell mask, sponge factor or injection applies anywhere in them)."""
    tx, tz, k = shape
    n = 0
    for i0 in range(0, nx, tx):
        for j0 in range(0, nz, tz):
            lo_i, hi_i, lo_j, hi_j = i0 - k, i0 + tx + k - 1, j0 - k, j0 + tz + k - 1
            if lo_i < max(1, fw) or hi_i > min(nx - 2, nx - 1 - fw):
                continue
            if lo_j < 1 or hi_j > min(nz - 2, nz - 1 - fw):
                continue
            if any(lo_i <= i <= hi_i and lo_j <= j <= hi_j for i, j in points):
                continue
            n += 1
    return n


@pytest.mark.parametrize("shape", INNER_SHAPES, ids=lambda s: "%dx%dk%d" % s)
@pytest.mark.parametrize("system", ["psv", "sh"])
def test_blocked_inner_tiles_bitwise_equal_step_kernels(system, shape):
    """A grid large enough for tiles away from the edges, the sponge, the
    sources and the receivers, with a source and a receiver line inside the
    grid and the model difference between them, so that the forward and
    adjoint wavefields cross those tiles: records and gradients of K steps
    per launch equal one step per launch bit for bit."""
    _case("FWI blocked inner tiles %s %s" % (system, shape))
    nx, nz, nt = 200, 150, 500
    grid, shots, m0, mt = problem(system, nx=nx, nz=nz, nt=nt, n_shots=2)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    blob = np.exp(-(((ii - 85) / 5.0) ** 2 + ((jj - 75) / 5.0) ** 2))
    mt = (m0[0] * (1 + 0.08 * blob), m0[1] * (1 - 0.06 * blob), m0[2])
    rec = np.stack([np.arange(20, 70), np.full(50, 60)], 1).astype(np.int32)
    shots = [ShotSpec(np.array([[50, 70]], np.int32), shots[0].wavelets, rec),
             ShotSpec(np.array([[150, 110]], np.int32), shots[1].wavelets, rec[10:40])]
    assert _inner_tiles(nx, nz, grid[5], shape, [(50, 70)] + [tuple(r) for r in rec]) > 0
    from taichirtm_amd import _lib, autotune
    systems = _lib.PSV if system == "psv" else _lib.SH
    # a shape whose adjoint exceeds the LDS runs the forward passes only
    adj = None if ("fwi",) + tuple(shape) in autotune.fwi_candidates(systems, "adjoint") \
        else (0, 0, 0)
    tx, tz, k = shape
    ref = _records_and_gradient(grid, shots, m0, mt, kernel="step", checkpoint_every=60)
    got = _records_and_gradient(grid, shots, m0, mt, kernel="blocked", tile_x=tx, tile_z=tz,
                                block_steps=k, checkpoint_every=60, adjoint_shape=adj)
    assert max(float(x.abs().max()) for x in ref[1]) > 0
    for a, b in zip(got[0], ref[0]):
        assert torch.equal(a, b)
    for a, b in zip(got[1], ref[1]):
        assert torch.equal(a, b)
    for key in ref[2]:
        assert torch.equal(got[2][key], ref[2][key]), key


def test_blocked_kernels_many_shots_bitwise():
    """Ten shots: two launches of up to eight shots per step block."""
    _case("FWI blocked 10 shots")
    grid, shots, m0, mt = problem("psv", nt=90, n_shots=10)
    shots = [ShotSpec(np.array([[4 + 5 * q, 2]], np.int32), s.wavelets, s.rec_loc)
             for q, s in enumerate(shots)]
    shots = _uneven(shots, [90 - 3 * q for q in range(10)], [(q, 40 + q) for q in range(10)])
    ref = _records_and_gradient(grid, shots, m0, mt, kernel="step", checkpoint_every=20)
    got = _records_and_gradient(grid, shots, m0, mt, kernel="blocked", checkpoint_every=20)
    for a, b in zip(got[0], ref[0]):
        assert torch.equal(a, b)
    for a, b in zip(got[1], ref[1]):
        assert torch.equal(a, b)


def test_tuned_kernels_bitwise_equal_step_kernels():
    """kernel="tuned" picks a shape per pass by timing trial windows on the
    problem's own buffers and undoes them: results equal one step per launch."""
    _case("FWI tuned")
    grid, shots, m0, mt = problem("psv", nt=150, n_shots=3)
    shots = _uneven(shots, (150, 111, 137), ((0, 52), (5, 30), (20, 45)))
    shots[2] = ShotSpec(np.array([[45, 2]], np.int32), shots[2].wavelets, shots[2].rec_loc)
    ref = _records_and_gradient(grid, shots, m0, mt, kernel="step", checkpoint_every=70)
    got = _records_and_gradient(grid, shots, m0, mt, kernel="tuned", checkpoint_every=70)
    for a, b in zip(got[0], ref[0]):
        assert torch.equal(a, b)
    for a, b in zip(got[1], ref[1]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("kernel", ["blocked", "tuned", "step"])
def test_history_kept_by_forward_bitwise(kernel):
    """history_bytes: the trailing segments' history written by the forward
    pass (no checkpoint, no recompute) gives the recomputing run's records and
    gradients bit for bit, with every segment kept and with some (3 shots of
    uneven record length, the last segment shorter than the interval)."""
    _case("FWI kept history %s" % kernel)
    grid, shots, m0, mt = problem("psv", nt=150, n_shots=3)
    shots = _uneven(shots, (150, 111, 137), ((0, 52), (5, 30), (20, 45)))
    shots[2] = ShotSpec(np.array([[45, 2]], np.int32), shots[2].wavelets, shots[2].rec_loc)
    nx, nz = grid[0], grid[1]
    ref = _records_and_gradient(grid, shots, m0, mt, kernel=kernel, checkpoint_every=40)
    # history of one 40-step segment of the 3 shots: every step's 5 fields
    # (one step per launch), or the blocked kernels' compact layout (2
    # velocities per step, 3 stresses per 4 steps)
    fields = 40 * 5 if kernel == "step" else 40 * 2 + (40 // 4) * 3
    seg = fields * nx * nz * 4 * 3
    for kept, nbytes in ((4, 10 * seg), (2, 2 * seg + seg // 2)):
        p = ElasticFWI(*grid, shots, precision="float32", kernel=kernel, checkpoint_every=40,
                       history_bytes=nbytes)
        obs = [o.clone() for o in p.forward(*[_t(x, torch.float32) for x in mt])]
        recs = [r.clone() for r in p.forward(*[_t(x, torch.float32) for x in m0])]
        assert p.kept_segments == kept
        g = p.gradient([r - o for r, o in zip(recs, obs)])
        for a, b in zip(recs, ref[0]):
            assert torch.equal(a, b)
        for a, b in zip(g, ref[1]):
            assert torch.equal(a, b)
        for key in ref[2]:
            assert torch.equal(p.coefficient_gradient[key], ref[2][key]), key
        with pytest.raises(RuntimeError):
            p.gradient([r - o for r, o in zip(recs, obs)])


@pytest.mark.parametrize("system", ["psv", "sh"])
def test_gradient_correlates_with_rtm_image(system):
    """The exact-adjoint gradient against the reference's RTM: back-propagating
    the residuals with the RTM backward (backward_modeling.py:539-576) and
    imaging them gives a picture of the same reflector as the gradient with
    respect to 1/rho on the velocity grids (g_iru ~ sum p_u du/dt, the image
    sum F_u B_u): spatial correlation over the interior above 0.8 (measured
    0.90 / 0.96 for u / w, 0.81 for SH v)."""
    from taichirtm_amd.batch import BatchRTM
    _case("FWI gradient vs RTM image %s" % system)
    grid, shots, m0, mt = problem(system, n_shots=2)
    nx, nz, dx, dz, fs, fw = grid
    p = ElasticFWI(*grid, shots, precision="float64", checkpoint_every=23)
    obs = [o.clone() for o in p.forward(*[_t(x) for x in mt])]
    res = [r - o for r, o in zip(p.forward(*[_t(x) for x in m0]), obs)]
    p.gradient(res)
    cg = {k: v.double().cpu().numpy() for k, v in p.coefficient_gradient.items()}
    vp, vs, rho = [np.asarray(x, np.float32) for x in m0]
    mu, lam = configs.moduli(vp, vs, rho)
    rshots = [ShotSpec(s.src_loc, tuple(np.asarray(w, np.float32) for w in s.wavelets), s.rec_loc,
                       tuple(r[c].cpu().numpy().astype(np.float32) for c in range(3)))
              for s, r in zip(shots, res)]
    b = BatchRTM(nx, nz, dx, dz, fs, mu, lam, rho, fw, rshots, isnap=1, kernel="blocked")
    assert b.run() == [0, 0]
    img = {k: v.cpu().numpy() for k, v in b.stacked_image().items()}
    inner = (slice(fw, nx - fw), slice(3, nz - fw))

    def corr(a, c):
        a, c = a[inner].ravel(), c[inner].ravel()
        a, c = a - a.mean(), c - c.mean()
        return float((a * c).sum() / np.sqrt((a * a).sum() * (c * c).sum()))

    pairs = [("u", "g_iru", 0.8), ("w", "g_irw", 0.9)] if system == "psv" else [("v", "g_ir", 0.7)]
    for ic, gk, lo in pairs:
        r = corr(img[ic], cg[gk])
        print("[rtm] image_%s vs %s corr %.3f" % (ic, gk, r), flush=True)
        assert r > lo


def test_dot_product_adjoint_fp64():
    """<d'(m) dm, r> = <dm, d'(m)^T r> for a random data vector r: the
    derivative of the records along dm by Richardson-extrapolated central
    differences against the adjoint applied to r (the gradient of <d(m), r>)."""
    _case("FWI dot product")
    grid, shots, m0, mt = problem("psv", nt=140)
    p = ElasticFWI(*grid, shots, precision="float64", checkpoint_every=31)
    rng = np.random.default_rng(7)
    m0t = [_t(x) for x in m0]
    recs = p.forward(*m0t)
    r = [torch.as_tensor(rng.standard_normal(tuple(x.shape)), device="cuda") for x in recs]
    g = p.gradient(r)
    nx, nz = grid[0], grid[1]
    from scipy.ndimage import gaussian_filter
    dm = [_t(gaussian_filter(rng.standard_normal((nx, nz)), 3.0) * s)
          for s in (m0[0].mean(), m0[1].mean(), m0[2].mean())]
    dm = [x / float(x.abs().max()) * s for x, s in zip(dm, (50.0, 30.0, 0.05))]

    def proj(eps):
        d = p.forward(*[a + eps * b for a, b in zip(m0t, dm)])
        return sum(float((x * y).sum()) for x, y in zip(d, r))

    def central(eps):
        return (proj(eps) - proj(-eps)) / (2 * eps)

    lhs = (4 * central(0.5e-4) - central(1e-4)) / 3
    rhs = float(sum((a * b).sum() for a, b in zip(g, dm)))
    rel = abs(lhs - rhs) / abs(lhs)
    print("[dot] <F dm, r> %.12e  <dm, F^T r> %.12e  rel %.2e" % (lhs, rhs, rel), flush=True)
    assert rel < 1e-8
