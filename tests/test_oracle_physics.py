"""Pins the oracle's time stepping, backward propagation and imaging.

There are no reference tests or golden vectors for this path, so the oracle is
pinned by mathematics: analytic Green's functions (forward step), a reflection
identity (backward step against the forward one), closed forms for the
injection and imaging, and exact invariants.  Each pin is also run against an
oracle build with one planted mistake, which it must reject.
"""
import numpy as np
import pytest

import oracle
from tests.analytic import Green, ricker

VP, VS, RHO = 3464.0, 2000.0, 2.0   # density in g/cm^3 keeps mu*mu*mu*mu finite in float32
OFFSETS = [(60, 0), (0, 60), (42, 42), (-30, 50)]


def _homogeneous(n, vp=VP, vs=VS, rho=RHO):
    mu = np.full((n, n), rho * vs ** 2, np.float32)
    lam = np.full((n, n), ((vp / vs) ** 2 - 2) * rho * vs ** 2, np.float32)
    return mu, lam, np.full((n, n), rho, np.float32)


def analytic_errors(variant="strict", refine=1, dtype=np.float32):
    """Relative L2 misfit of oracle seismograms against the analytic solution,
    SH (v from a v-force) and P-SV (w and u from a w-force), in one run (the
    two systems are decoupled)."""
    dx = 5.0 / refine
    dt = 4e-4 / refine
    nt = 1125 * refine
    n = 400 * refine
    c = n // 2
    mu, lam, rho = _homogeneous(n)
    rec = [(c + a * refine, c + b * refine) for a, b in OFFSETS]
    w = ricker(10.0, nt, dt)
    zero = np.zeros(nt)
    m = oracle.Model(mu, lam, rho, dx, dx, 1.0 / dt, nt, [(c, c)], rec, fw=0, sci=100,
                     dtype=dtype, variant=variant)
    out = m.forward([zero, w, w])
    if out["code"] != 0:  # blew up: infinitely far from the analytic solution
        return {k: np.full(len(OFFSETS), np.inf) for k in ("v", "w", "u")}
    su, sv, sw = out["seis"]
    g = Green(w, dt, dx, dx, VP, VS, RHO)
    errs = {"v": [], "w": [], "u": []}
    for l, (a, b) in enumerate(OFFSETS):
        A, B = a * refine, b * refine
        v_an = g.sh_velocity(A * dx, B * dx)
        w_an = g.psv_velocity(A * dx, B * dx, 1, 1)
        # u[i,j] sits half a cell before the normal stresses in x and w[i,j]
        # half a cell before them in z (the velocity updates use backward
        # differences of stress, forward_modeling.py:332-335)
        u_an = g.psv_velocity((A - 0.5) * dx, (B + 0.5) * dx, 0, 1)
        errs["v"].append(oracle.relative_l2(sv[l], v_an))
        errs["w"].append(oracle.relative_l2(sw[l], w_an))
        errs["u"].append(float(np.linalg.norm(su[l] - u_an) / np.linalg.norm(w_an)))
    return {k: np.array(v) for k, v in errs.items()}


def analytic_ok(e):
    return bool(np.all(e["v"] < 0.04) and np.all(e["w"] < 0.06) and np.all(e["u"] < 0.25))


@pytest.fixture(scope="module")
def base_errors():
    return analytic_errors()


def test_forward_matches_green_function(base_errors):
    e = base_errors
    assert analytic_ok(e), e


def test_forward_float64_matches_green_function():
    e = analytic_errors(dtype=np.float64)
    assert analytic_ok(e), e


@pytest.mark.slow
def test_forward_converges_on_refinement(base_errors):
    """Halving dx and dt cuts the v and w misfit about 4x (second order); the
    u misfit, which goes through the scheme's offset w-sxz coupling, at least
    halves."""
    e2 = analytic_errors(refine=2)
    assert np.all(base_errors["v"] / e2["v"] > 3.0), (base_errors, e2)
    assert np.all(base_errors["w"] / e2["w"] > 3.0), (base_errors, e2)
    big = base_errors["u"] > 0.01
    assert np.all(base_errors["u"][big] / e2["u"][big] > 1.6), (base_errors, e2)


@pytest.mark.parametrize("plant", ["plant1", "plant2", "plant3"])
def test_green_function_pin_rejects_planted_mistakes(plant):
    """plant1: wrong sign of the shear-stress term; plant2: source scaled by
    dt/(dx dz) instead of dt*dx*dz; plant3: u_z differenced along x."""
    assert not analytic_ok(analytic_errors(variant=plant))


# ---------------------------------------------------------------------------
# backward step: reflection identity


def _reflect(a):
    return np.ascontiguousarray(a[::-1, ::-1])


def _random_state(rng, n, dtype):
    st = {}
    for name in oracle.ST_NAMES:
        a = np.zeros((n, n), dtype)
        a[1:-1, 1:-1] = rng.standard_normal((n - 2, n - 2))
        st[name] = a
    return st


def backward_reflection_mismatch(variant="strict", dtype=np.float32, steps=7):
    """backward_modeling.py:312-380 against forward_modeling.py:309-384.

    With i -> nx-1-i and j -> nz-1-j, a forward difference becomes minus a
    backward one, so one backward step (forward differences of stress,
    backward differences of velocity, -dt) equals one forward step on the
    reflected grid, bit for bit, when the medium is homogeneous, the sponge is
    off (FW=0) and the grid border starts at zero.
    """
    n = 23
    rng = np.random.default_rng(11)
    mu, lam, rho = _homogeneous(n)
    m = oracle.Model(mu, lam, rho, 5.0, 5.0, 2500.0, steps, [(5, 5)], [(6, 6)], fw=0,
                     sci=0, dtype=dtype, variant=variant)
    st_b = _random_state(rng, n, dtype)
    st_f = {k: _reflect(v) for k, v in st_b.items()}
    zero = np.zeros((1, steps))
    m.backward([zero, zero, zero], None, None, 0, 1, state=st_b)
    m.forward([zero, zero, zero], state=st_f)
    return max(float(np.max(np.abs(_reflect(st_f[k]) - st_b[k]))) for k in oracle.ST_NAMES)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_backward_is_reflected_forward(dtype):
    assert backward_reflection_mismatch(dtype=dtype) == 0.0


def test_reflection_pin_rejects_planted_mistake():
    """plant4: backward velocity update differences sxx backwards."""
    assert backward_reflection_mismatch(variant="plant4") > 1e-3


# ---------------------------------------------------------------------------
# injection closed forms


def _layered(nx, nz, seed=0):
    rng = np.random.default_rng(seed)
    vs = rng.uniform(800, 1500, (nx, nz)).astype(np.float32)
    rho = rng.uniform(1.8, 2.4, (nx, nz)).astype(np.float32)
    vp = (1.8 * vs).astype(np.float32)
    mu = (rho * vs ** 2).astype(np.float32)
    lam = (((vp / vs) ** 2 - 2) * mu).astype(np.float32)
    return mu, lam, rho


def source_first_step(variant="strict"):
    """forward_modeling.py:421-443: from rest, one step leaves
    (w[k,0] * f32(dt*dx*dz) * inv_rho_c) * sponge at the source cell."""
    nx, nz, nt = 30, 24, 4
    mu, lam, rho = _layered(nx, nz)
    dx, dz, fs = 5.0, 4.0, 2500.0
    src = (9, 7)
    m = oracle.Model(mu, lam, rho, dx, dz, fs, nt, [src], [src], fw=8, variant=variant)
    wav = [np.array([[1.5, 0, 0, 0]]), np.array([[-2.25, 0, 0, 0]]), np.array([[0.75, 0, 0, 0]])]
    out = m.forward(wav, it1=1)
    factor = np.float32((1.0 / fs) * dx * dz)
    c = m.coef
    A = c["absorb"][src]
    exp = {
        "u": np.float32(np.float32(np.float32(1.5) * factor) * c["inv_rho_u"][src]) * A,
        "v": np.float32(np.float32(np.float32(-2.25) * factor) * c["inv_rho"][src]) * A,
        "w": np.float32(np.float32(np.float32(0.75) * factor) * c["inv_rho_w"][src]) * A,
    }
    return {k: (out["state"][k][src], exp[k]) for k in exp}, out


def test_source_injection_closed_form():
    vals, out = source_first_step()
    for k, (got, exp) in vals.items():
        assert got == exp, (k, got, exp)
    # recorded after the sponge, at column it = 0
    assert out["seis"][2][0, 0] == vals["w"][0]


def test_source_pin_rejects_planted_mistake():
    vals, _ = source_first_step("plant2")
    assert any(got != exp for got, exp in vals.values())


def obs_first_step(variant="strict"):
    """backward_modeling.py:452-460: from rest, backward step it=0 (t = nt-1)
    adds obs[k, nt-1] unscaled, then the stress update and sponge run; the
    receiver cell's velocity is obs[k, nt-1] * sponge."""
    nx, nz, nt = 30, 24, 5
    mu, lam, rho = _layered(nx, nz, 1)
    rec = [(4, 1), (0, 1), (17, 9), (17, 9)]  # border cell and a duplicate
    m = oracle.Model(mu, lam, rho, 5.0, 5.0, 2500.0, nt, [(3, 3)], rec, fw=6, variant=variant)
    rng = np.random.default_rng(5)
    obs = [rng.standard_normal((len(rec), nt)).astype(np.float32) for _ in range(3)]
    out = m.backward(obs, None, None, 0, 1, it1=1)
    A = m.coef["absorb"]
    checks = []
    for comp, o in zip(("u", "v", "w"), obs):
        got = out["state"][comp]
        checks.append((got[4, 1], o[0, -1] * A[4, 1]))
        checks.append((got[0, 1], o[1, -1] * A[0, 1]))
        checks.append((got[17, 9], np.float32(o[2, -1] + o[3, -1]) * A[17, 9]))
    return checks


def test_observed_data_injection_closed_form():
    for got, exp in obs_first_step():
        assert got == np.float32(exp)


def test_injection_pin_rejects_planted_mistake():
    assert any(got != np.float32(exp) for got, exp in obs_first_step("plant5"))


# ---------------------------------------------------------------------------
# imaging condition


def image_by_explicit_sum(variant="strict", isnap=3, nt=31):
    """backward_modeling.py:481-488,556-564: I_c = sum over t in T, t
    descending, of F_c(t) * B_c(t) in float32, T = {1 <= t <= nt-1, t % isnap == 0,
    t/isnap - 1 < nt//isnap}; F(t) the forward state after step t, B(t) the
    backward state after step nt-1-t."""
    nx, nz = 26, 22
    mu, lam, rho = _layered(nx, nz, 2)
    rec = [(i, 1) for i in range(2, 24, 3)]
    m = oracle.Model(mu, lam, rho, 5.0, 5.0, 2500.0, nt, [(12, 3)], rec, fw=5, isnap=isnap)
    m_var = oracle.Model(mu, lam, rho, 5.0, 5.0, 2500.0, nt, [(12, 3)], rec, fw=5, isnap=isnap,
                         variant=variant)
    w = ricker(60.0, nt, 4e-4)
    zero = np.zeros(nt)
    # forward history, one step at a time
    st = m.new_state()
    seis = [np.zeros((len(rec), nt), np.float32) for _ in range(3)]
    F = {}
    for it in range(nt):
        m.forward([zero, w, w], state=st, seis=seis, it0=it, it1=it + 1)
        F[it] = {k: st[k].copy() for k in ("u", "v", "w")}
    obs = [s * np.float32(0.5) for s in seis]
    stb = m.new_state()
    B = {}
    for it in range(nt):
        m.backward(obs, None, None, 0, isnap, state=stb, it0=it, it1=it + 1)
        B[nt - 1 - it] = {k: stb[k].copy() for k in ("u", "v", "w")}
    num_snaps = nt // isnap
    I = {k: np.zeros((nx, nz), np.float32) for k in ("u", "v", "w")}
    for t in range(nt - 1, 0, -1):
        if t % isnap == 0 and t // isnap - 1 < num_snaps:
            for k in I:
                I[k] = I[k] + F[t][k] * B[t][k]
    run = oracle.rtm(m_var, [zero, w, w], obs=obs, isnap=isnap)
    return I, run


@pytest.mark.parametrize("isnap,nt", [(1, 25), (3, 31), (4, 32)])
def test_image_is_explicit_sum(isnap, nt):
    I, run = image_by_explicit_sum(isnap=isnap, nt=nt)
    for k, r in zip(("u", "v", "w"), run["bw"]["result"]):
        assert np.array_equal(r, I[k]), k
        assert np.any(r != 0)


def test_imaging_pin_rejects_planted_mistake():
    I, run = image_by_explicit_sum(variant="plant6", isnap=3, nt=31)
    assert not all(np.array_equal(r, I[k]) for k, r in zip(("u", "v", "w"), run["bw"]["result"]))


# ---------------------------------------------------------------------------
# invariants


def _small_problem(nt=120, variant="strict"):
    nx, nz = 48, 40
    mu, lam, rho = _layered(nx, nz, 4)
    rec = [(i, 1) for i in range(0, nx, 5)] + [(30, 20)]
    return oracle.Model(mu, lam, rho, 5.0, 5.0, 2500.0, nt, [(20, 6)], rec, fw=8, variant=variant)


def test_linearity_power_of_two():
    m = _small_problem()
    w = ricker(40.0, m.nt, 4e-4)
    z = np.zeros(m.nt)
    a = m.forward([w, w * 0.5, -w])
    b = m.forward([w * 4, w * 2, -w * 4])
    for sa, sb in zip(a["seis"], b["seis"]):
        # exact except where float32 subnormals (below 1.2e-38) round differently
        assert np.max(np.abs(sa * 4 - sb)) < 1e-35
        assert np.max(np.abs(sb)) > 1e-9


def test_time_shift_invariance():
    m = _small_problem()
    k = 17
    w = ricker(40.0, m.nt, 4e-4)
    ws = np.concatenate([np.zeros(k), w[:-k]])
    z = np.zeros(m.nt)
    a = m.forward([z, w, w])
    b = m.forward([z, ws, ws])
    for sa, sb in zip(a["seis"], b["seis"]):
        assert np.array_equal(sa[:, : m.nt - k], sb[:, k:])
        assert np.all(sb[:, :k] == 0)


def test_psv_sh_decoupled():
    m = _small_problem()
    w = ricker(40.0, m.nt, 4e-4)
    z = np.zeros(m.nt)
    out = m.forward([w, z, w])
    for k in ("v", "syx", "syz"):
        assert np.all(out["state"][k] == 0)
    out = m.forward([z, w, z])
    for k in ("u", "w", "sxx", "szz", "sxz"):
        assert np.all(out["state"][k] == 0)


def test_finite_check_codes():
    """forward_modeling.py:455-488,563-566 / backward_modeling.py:417-450,573-576."""
    m = _small_problem(nt=300)
    w = ricker(40.0, m.nt, 4e-4)
    z = np.zeros(m.nt)
    bad = w.copy()
    bad[5] = np.nan
    out = m.forward([bad, z, w])
    # checked at it % 100 == 0: at it=100 the NaN has reached w through sxz
    assert (out["code"], out["it_stop"]) == (3, 100)
    m.sci = 1
    m._prob.sci = 1
    out = m.forward([bad, z, w])
    assert (out["code"], out["it_stop"]) == (1, 5)
    big = w.copy()
    big[3] = 1e36
    out = m.forward([z, big, z])
    assert (out["code"], out["it_stop"]) == (2, 3)
    obs = [np.zeros((len(m.rec_loc), m.nt), np.float32) for _ in range(3)]
    obs[2][0, m.nt - 1 - 7] = np.inf
    out = m.backward(obs, None, None, 0, 1)
    assert (out["code"], out["it_stop"]) == (6, 7)


def test_fp32_error_and_fma_spread_are_small():
    """Rounding sensitivity of the scheme: float32 against the float64 run, and
    operand-by-operand evaluation against FMA contraction, on a forward +
    backward + image problem."""
    res = {}
    for tag, dtype, variant in (("f32", np.float32, "strict"), ("f64", np.float64, "strict"),
                                ("fma", np.float32, "fma")):
        m = _small_problem(nt=400, variant=variant)
        m = oracle.Model(m.mu, m.lam, m.rho, 5.0, 5.0, 2500.0, 400, m.src_loc, m.rec_loc, fw=8,
                         dtype=dtype, variant=variant)
        w = ricker(40.0, m.nt, 4e-4)
        z = np.zeros(m.nt)
        res[tag] = oracle.rtm(m, [z, w, w], isnap=1)
    for comp in (0, 2):
        e64 = oracle.relative_l2(res["f32"]["fw"]["seis"][comp], res["f64"]["fw"]["seis"][comp])
        efma = oracle.relative_l2(res["fma"]["fw"]["seis"][comp], res["f32"]["fw"]["seis"][comp])
        i64 = oracle.relative_l2(res["f32"]["bw"]["result"][comp], res["f64"]["bw"]["result"][comp])
        ifma = oracle.relative_l2(res["fma"]["bw"]["result"][comp], res["f32"]["bw"]["result"][comp])
        assert e64 < 1e-5 and efma < 1e-5 and i64 < 1e-5 and ifma < 1e-5, (e64, efma, i64, ifma)


# -- topography (surface matrix) ---------------------------------------------

def _topo_problem(nx=40, nz=32, nt=70, variant="strict", surface=None):
    rng = np.random.default_rng(11)
    mu = (RHO * (VS * (1 + 0.1 * rng.random((nx, nz)))) ** 2).astype(np.float32)
    lam = (((VP / VS) ** 2 - 2) * mu).astype(np.float32)
    rho = np.full((nx, nz), RHO, np.float32)
    src = [(nx // 2, 9)]
    rec = np.stack([np.arange(2, nx - 2), np.full(nx - 4, 9)], 1)
    m = oracle.Model(mu, lam, rho, 5.0, 5.0, 2500.0, nt, src, rec, 6, isnap=5, sci=1000,
                     surface_matrix=surface, variant=variant)
    w = ricker(60.0, nt, 4e-4)[None].astype(np.float32)
    return m, [0.5 * w, w, w]


def topo_flat_mismatch(variant="strict"):
    """A surface matrix that is 1 everywhere except the row j = 0, where it is
    0, must give the flat free surface (forward_modeling.py:404-410, zeroing
    syz, sxz, szz on j = 0 before each step) value for value, since x * 1 = x
    and x * 0 = 0 exactly.  Returns the number of differing values."""
    m0, wav = _topo_problem(variant=variant)
    sm = np.ones((m0.nx, m0.nz), np.float32)
    sm[:, 0] = 0
    m1, _ = _topo_problem(variant=variant, surface=sm)
    a = m0.forward(wav, save=True)
    b = m1.forward(wav, save=True)
    bad = 0
    for k in oracle.ST_NAMES:
        bad += int((a["state"][k] != b["state"][k]).sum())
    for c in range(3):
        bad += int((a["seis"][c] != b["seis"][c]).sum())
        bad += int((a["snaps"][c] != b["snaps"][c]).sum())
    return bad


def topo_stepwise_mismatch(variant="strict"):
    """forward_modeling.py:412-419: before every step, syz, sxz and szz (and
    nothing else) are multiplied by the surface matrix.  The oracle run with a
    topography mask (0 above a stepped surface, 1 below) must equal, value for
    value, the same steps taken one at a time with those three fields masked
    in numpy first and a surface matrix of ones (multiplying by 1 is exact).
    Returns the number of differing values."""
    m, wav = _topo_problem(variant=variant)
    nx, nz = m.nx, m.nz
    depth = (3 + 3 * (np.arange(nx) > nx // 2)).astype(int)
    sm = (np.arange(nz)[None, :] >= depth[:, None]).astype(np.float32)
    mt, _ = _topo_problem(variant=variant, surface=sm)
    m1, _ = _topo_problem(variant=variant, surface=np.ones((nx, nz), np.float32))
    a = mt.forward(wav)
    st = m1.new_state()
    seis = [np.zeros((len(m1.rec_loc), m1.nt), np.float32) for _ in range(3)]
    for it in range(m1.nt):
        for k in ("syz", "sxz", "szz"):
            st[k] *= sm
        m1.forward(wav, state=st, seis=seis, it0=it, it1=it + 1)
    bad = sum(int((a["state"][k] != st[k]).sum()) for k in oracle.ST_NAMES)
    bad += sum(int((a["seis"][c] != seis[c]).sum()) for c in range(3))
    assert np.abs(a["seis"][2]).max() > 0 and np.abs(a["state"]["sxz"]).max() > 0
    return bad


def test_topography_row_zero_equals_flat_surface():
    assert topo_flat_mismatch() == 0


def test_topography_masks_stresses_before_each_step():
    assert topo_stepwise_mismatch() == 0


@pytest.mark.parametrize("plant", ["plant7", "plant8"])
def test_topography_pins_reject_planted_mistakes(plant):
    """plant7: the mask also scales sxx; plant8: the mask is applied after the
    velocity update instead of before it."""
    assert topo_flat_mismatch(plant) > 0 or topo_stepwise_mismatch(plant) > 0
    assert topo_stepwise_mismatch(plant) > 0
