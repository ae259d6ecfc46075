"""FWI gradient and RTM image together under a device-memory budget.

Pins: the grouping the budget forces (shots per group, checkpoint spacing)
does not change a bit of the gradient or the image (per-shot gradients summed
in shot order in float32, images in float64); the image from the shared
forward history (sponge x post-injection velocity) equals the image of the
RTM's own forward snapshots, and both results equal the direct ElasticFWI /
BatchRTM ones, bit for bit; on the 4000 x 2000 salt model the job stays
within its budget."""
import numpy as np
import pytest
import torch

from taichirtm_amd import budget
from taichirtm_amd.batch import BatchRTM, ShotSpec
from taichirtm_amd.fwi import ElasticFWI, l2_misfit_and_gradient

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case]", name, flush=True)


def _survey(nx, nz, nt, n_shots, nr_range):
    true, start = budget.configs.psv_salt(nx=nx, nz=nz, nt=nt, n_shots=n_shots,
                                          nr_range=nr_range, fw=12)
    if not (true.vp == 4500.0).any():  # small grids: a salt disk below the sources
        ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
        disk = (ii - nx / 2) ** 2 + (jj - nz / 3) ** 2 < (nz / 8) ** 2
        true.vp[disk], true.vs[disk], true.rho[disk] = 4500.0, 2600.0, 2.15
    shots = [ShotSpec(np.array([s.src], np.int32), true.wavelets(s), np.asarray(s.rec, np.int32))
             for s in true.shots]
    obs = budget.observed_data(true, shots)
    with_obs = [ShotSpec(s.src_loc, s.wavelets, s.rec_loc, tuple(o)) for s, o in zip(shots, obs)]
    return true, start, shots, obs, with_obs


def test_grouping_does_not_change_gradient_or_image():
    _case("budget grouping 160x100x300, 5 shots")
    true, start, shots, obs, with_obs = _survey(160, 100, 300, 5, (40, 120))
    grid = (start.nx, start.nz, start.dx, start.dz, start.fs, start.fw)
    m = (start.vp, start.vs, start.rho)
    p1 = budget.plan(start.nx, start.nz, with_obs, 1 << 40)
    assert p1.group == 5
    tight = p1.shared_bytes + int(2.5 * max(p1.shot_bytes_fwi, p1.shot_bytes_rtm))
    p2 = budget.plan(start.nx, start.nz, with_obs, tight)
    assert p2.group == 2 and p2.groups == 3
    outs = []
    for b, share in ((1 << 40, True), (tight, True), (1 << 40, False)):
        job = budget.GradientAndImage(*grid, with_obs, b, kernel="blocked", share_forward=share)
        outs.append(job.run(*m))
    (J1, g1, i1) = outs[0]
    assert J1 > 0
    for J2, g2, i2 in outs[1:]:  # tighter budget; separate forward passes
        assert J1 == J2
        for a, b in zip(g1, g2):
            assert torch.equal(a, b)
        for c in "uvw":
            assert torch.equal(i1[c], i2[c])
    # the direct calls
    fwi = ElasticFWI(*grid, shots, precision="float32", checkpoint_every=p1.L_fwi, kernel="blocked")
    t = lambda x: torch.as_tensor(np.asarray(x, np.float32), device="cuda")
    J, g = l2_misfit_and_gradient(fwi, *[t(x) for x in m], [torch.as_tensor(o) for o in obs])
    for a, b in zip(g, g1):
        assert torch.equal(a, b)
    mu, lam = start.moduli()
    rtm = BatchRTM(start.nx, start.nz, start.dx, start.dz, start.fs, mu, lam, start.rho, start.fw,
                   with_obs, imaging="checkpoint", checkpoint_every=p1.L_rtm, kernel="blocked")
    assert rtm.run() == [0] * 5
    st = rtm.stacked_image()
    for c in "uw":
        assert torch.equal(st[c], i1[c])


def test_salt_full_grid_within_budget():
    """4000 x 2000 salt model, 2 shots of 240 steps, budget for one shot per
    group: peak device memory of the job stays within the budget."""
    _case("budget salt 4000x2000x240, 2 shots")
    true, start, shots, obs, with_obs = _survey(4000, 2000, 240, 2, (1000, 3000))
    p = budget.plan(start.nx, start.nz, with_obs, 1 << 40)
    # sharing the forward history, a shot holds its FWI and RTM-backward memory
    b = p.shared_bytes + int(1.2 * (p.shot_bytes_fwi + p.shot_bytes_rtm))
    job = budget.GradientAndImage(start.nx, start.nz, start.dx, start.dz, start.fs, start.fw,
                                  with_obs, b)
    assert job.plan.group == 1
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    # 240 steps do not reach the salt: take the gradient in a 2 % slower model
    J, g, img = job.run(0.98 * start.vp, 0.98 * start.vs, start.rho)
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    print("[budget] plan %s peak %.2f GB budget %.2f GB" % (job.plan, peak / 1e9, b / 1e9),
          flush=True)
    assert peak <= b
    assert J > 0
    for x in g:
        assert bool(torch.isfinite(x).all()) and float(x.abs().max()) > 0
    for c in "uw":
        assert bool(torch.isfinite(img[c]).all()) and float(img[c].abs().max()) > 0


@pytest.mark.parametrize("shape", [(24, 56, 512, 4), (16, 16, 1024, 8)],
                         ids=lambda s: "%dx%dk%d" % (s[0], s[1], s[3]))
def test_shared_forward_image_branch_free_shape(monkeypatch, shape):
    """The RTM backward of the shared-forward job in the salt job's shape (24 x 56
    tiles, 512 threads, 4 steps: four cells per thread, branch-free, its forward
    wavefield the history's velocities times the sponge, sponged where the
    image uses them) and in the small-grid shape (16 x 16, 8 steps: the
    forward wavefield loaded a step ahead) gives the image of the separate
    checkpointed RTM bit for bit."""
    tx, tz, nt_, k = shape
    _case("budget shared forward, RTM backward %s, 160x100x300, 3 shots" % (shape,))
    true, start, shots, obs, with_obs = _survey(160, 100, 300, 3, (40, 120))
    grid = (start.nx, start.nz, start.dx, start.dz, start.fs, start.fw)
    m = (start.vp, start.vs, start.rho)
    made = []
    k_ = k

    def pinned(*a, **k):
        made.append(BatchRTM(*a, tile_x=tx, tile_z=tz, block_steps=k_, **k))
        return made[-1]
    monkeypatch.setattr(budget, "BatchRTM", pinned)
    job = budget.GradientAndImage(*grid, with_obs, 1 << 40, kernel="blocked", share_forward=True)
    _, _, img = job.run(*m)
    monkeypatch.undo()
    assert made and made[0].bw[0].block_config() == shape
    mu, lam = start.moduli()
    p = budget.plan(start.nx, start.nz, with_obs, 1 << 40)
    rtm = BatchRTM(start.nx, start.nz, start.dx, start.dz, start.fs, mu, lam, start.rho, start.fw,
                   with_obs, imaging="checkpoint", checkpoint_every=p.L_rtm, kernel="blocked")
    assert rtm.run() == [0] * 3
    st = rtm.stacked_image()
    for c in "uw":
        assert torch.equal(st[c], img[c])
        assert float(img[c].abs().max()) > 0
