"""Captured time stepping (trtm_plan_*): a run recorded once into a HIP graph
and replayed gives the eager run's seismograms, snapshots, final state and
images bit for bit, on every replay, with eager work and synchronisations
between the replays."""
import numpy as np
import pytest
import torch

import oracle
import taichirtm_amd as trtm
from taichirtm_amd import _lib, configs
from taichirtm_amd.batch import BatchRTM
from tests.parity import same
from tests.test_gpu_batch import survey

pytestmark = pytest.mark.gpu


def _case(name):
    print("\n[case]", name, flush=True)


def _forward(cfg, kernel, **extra):
    shot = cfg.shots[0]
    wu, wv, ww = cfg.wavelets(shot)
    return trtm.ForwardModeling(wavelet_u_field=wu, wavelet_v_field=wv, wavelet_w_field=ww,
                                isnap=3, kernel=kernel, **cfg.kwargs(shot), **extra)


def _outputs(fw):
    snaps = [np.array(getattr(fw, c + "_save_field").to_numpy()) for c in ("u", "v", "w")]
    state = {k: getattr(fw, k).to_numpy() for k in oracle.ST_NAMES}
    return fw.get_seismogram(), snaps, state, fw.isnaps_field.to_numpy().copy()


@pytest.mark.parametrize("system,kernel,shape", [("psv", "blocked", (16, 64, 4)),
                                                 ("sh", "blocked", (32, 48, 8)),
                                                 ("sh", "wave", (0, 0, 0)),
                                                 ("psv", "wave", (0, 0, 0))])
def test_captured_forward_replays_equal_eager_run(system, kernel, shape):
    """ForwardModeling.capture: three replays of a 157-step run with
    snapshots (isnap 3; the last block shorter than K), each from a reset
    state, with unrelated eager work in between: every replay equals run()."""
    _case("captured forward %s %s %s" % (system, kernel, shape))
    nt = 157
    cfg = (configs.psv_anomaly(nx=60, nz=50, nt=nt, fw=10, radius=8, src=(30, 5),
                               rec_range=(2, 58), f0=40.0) if system == "psv"
           else configs.sh_layered(nx=96, nz=70, nt=nt, fw=12, f0=40.0, seed=3))
    tx, tz, k = shape
    ref_fw = _forward(cfg, kernel, tile_x=tx, tile_z=tz, block_steps=k)
    assert ref_fw.run(save=True) == 0
    ref = _outputs(ref_fw)
    fw = _forward(cfg, kernel, tile_x=tx, tile_z=tz, block_steps=k)
    run = fw.capture(save=True)
    try:
        for rep in range(3):
            fw.reset_state()
            x = torch.randn(1 << 20, device="cuda")  # eager work between replays
            float(x.square().sum())
            torch.cuda.synchronize()
            run.launch()
            assert run.result() == 0
            got = _outputs(fw)
            print("[plan] replay %d: seismogram max %.3e" % (rep, np.abs(got[0][2 if system == "psv" else 1]).max()),
                  flush=True)
            for a, b in zip(got[0], ref[0]):
                assert same(a, b), rep
            for a, b in zip(got[1], ref[1]):
                assert same(a, b), rep
            for key in ref[2]:
                assert same(got[2][key], ref[2][key]), (rep, key)
            assert same(got[3], ref[3])
    finally:
        run.close()


def test_capture_rejects_per_step_kernels():
    _case("capture of the fused kernel")
    cfg = configs.psv_anomaly(nx=60, nz=50, nt=40, fw=10, radius=8, src=(30, 5),
                              rec_range=(2, 58), f0=40.0)
    fw = _forward(cfg, "fused")
    with pytest.raises(_lib.HIPError, match="blocked and wave"):
        fw.capture()


@pytest.mark.parametrize("kernel", ["blocked", "wave"])
def test_batch_graph_equals_eager(kernel):
    """BatchRTM(graph=True): every batched range (checkpointing forward,
    forward with kept segments, recompute, backward with imaging) recorded
    and launched as a graph gives the eager batch's images and seismograms."""
    _case("batch graph %s" % kernel)
    cfg, mu, lam, shots = survey("psv")
    seg = (16 + 1) * 2 * cfg.nx * cfg.nz * 4

    def run(graph):
        b = BatchRTM(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, mu, lam, cfg.rho, cfg.fw, shots,
                     isnap=1, imaging="checkpoint", kernel=kernel, checkpoint_every=16,
                     snapshot_bytes=4 * seg, graph=graph)
        assert b.run() == [0] * len(shots)
        return b.images(), b.seismograms()

    ref, got = run(False), run(True)
    for q in range(len(shots)):
        for c in range(3):
            assert same(got[0][q][c], ref[0][q][c]), (q, c)
            assert same(got[1][q][c], ref[1][q][c]), (q, c)


@pytest.mark.parametrize("system,kernel", [("psv", "blocked"), ("sh", "blocked"), ("sh", "wave")])
def test_captured_rtm_replays_equal_eager(system, kernel):
    """Forward and backward both captured (the single-shot RTM of one shot as
    two graph launches), replayed three times from reset state with the
    source wavelet changed in place in between (x1, x2, x1): every replay
    gives the seismograms and image of run() / run_calc() with that wavelet,
    bit for bit, so the graphs read wavelets, forward snapshots and observed
    data where they live."""
    _case("captured RTM %s %s" % (system, kernel))
    nt = 131
    cfg = (configs.psv_anomaly(nx=60, nz=50, nt=nt, fw=10, radius=8, src=(30, 5),
                               rec_range=(2, 58), f0=40.0) if system == "psv"
           else configs.sh_layered(nx=96, nz=70, nt=nt, fw=12, f0=40.0, seed=3))
    kw = cfg.kwargs(cfg.shots[0])
    wav = cfg.wavelets(cfg.shots[0])

    def forward(scale):
        w = [np.asarray(x, np.float32) * np.float32(scale) for x in wav]
        return trtm.ForwardModeling(wavelet_u_field=w[0], wavelet_v_field=w[1],
                                    wavelet_w_field=w[2], isnap=3, kernel=kernel, **kw)

    def backward(fw):
        seis = fw.get_seismogram()
        return trtm.BackwardModeling(obsdata_u_field=seis[0], obsdata_v_field=seis[1],
                                     obsdata_w_field=seis[2], kernel=kernel, **kw)

    ref = {}
    for scale in (1.0, 2.0):
        f = forward(scale)
        assert f.run(save=True) == 0
        b = backward(f)
        assert b.run_calc(f.u_save_field, f.v_save_field, f.w_save_field, f.isnaps_field,
                          f.num_snaps, 3) == 0
        ref[scale] = (f.get_seismogram(), b.get_results())

    fw = forward(1.0)
    cf = fw.capture(save=True)
    cf.launch()
    assert cf.result() == 0
    bw = backward(fw)
    cb = bw.capture_calc(fw.u_save_field, fw.v_save_field, fw.w_save_field, fw.isnaps_field,
                         fw.num_snaps, 3)
    wav0 = [t.clone() for t in fw._inj]
    try:
        for rep, scale in enumerate((1.0, 2.0, 1.0)):
            for t, t0 in zip(fw._inj, wav0):  # the wavelets change in place
                t.copy_(t0 * scale)
            fw.reset_state()
            bw.reset_state()
            float(torch.randn(1 << 18, device="cuda").abs().max())  # eager work between replays
            cf.launch()
            for c, t in zip(("u", "v", "w"), bw._inj):  # observed data = this forward's records
                if t.numel() and c in fw._seis:
                    t.copy_(fw._seis[c].t())
            cb.launch()
            assert cf.result() == 0 and cb.result() == 0
            img = bw.get_results()
            print("[plan] RTM replay %d (source x%g): image max %.3e"
                  % (rep, scale, max(np.abs(x).max() for x in img)), flush=True)
            seis_ref, img_ref = ref[scale]
            for a, b in zip(img, img_ref):
                assert same(a, b), rep
            for a, b in zip(fw.get_seismogram(), seis_ref):
                assert same(a, b), rep
    finally:
        cf.close()
        cb.close()
