"""trtm_forward_steps_noalloc (workspace in the library's static device
scratch, no allocation) against trtm_forward_steps with a caller workspace:
bitwise equal records, final state, snapshots and check flags, on two calls in
a row, the second on a smaller grid (the scratch still holds the first call's
lists and state), for the blocked, fused and automatically chosen kernels."""
import numpy as np
import pytest
import torch

from taichirtm_amd import _lib, configs
from taichirtm_amd.field import stream_handle

pytestmark = pytest.mark.gpu


def _problem(nx, nz, nt, kernel, both=False):
    import taichirtm_amd as trtm
    cfg = configs.psv_anomaly(nx=nx, nz=nz, nt=nt, fw=10, radius=6, src=(nx // 2, 5),
                              rec_range=(2, nx - 2), f0=40.0)
    shot = cfg.shots[0]
    kw = cfg.kwargs(shot)
    wu, wv, ww = cfg.wavelets(shot)
    if both:  # an SH source as well, so both wave systems run
        wv = ww.copy()
    return trtm.ForwardModeling(wavelet_u_field=wu, wavelet_v_field=wv, wavelet_w_field=ww,
                                isnap=3, kernel=kernel, **kw)


def _outputs(fw, flags):
    torch.cuda.synchronize()
    out = {"flags": flags.cpu().numpy().copy()}
    for c, t in fw._seis.items():
        out["rec_" + c] = t.cpu().numpy().copy()
    for k in ("u", "v", "w", "sxx", "szz", "sxz", "syx", "syz"):
        t = getattr(fw, k).tensor
        if t is not None:
            out[k] = t.cpu().numpy().copy()
    for c in ("u", "v", "w"):
        f = getattr(fw, c + "_save_field", None)
        if f is not None and f.tensor is not None:
            out["snap_" + c] = f.tensor.cpu().numpy().copy()
    return out


def _run(nx, nz, nt, kernel, noalloc, both=False):
    fw = _problem(nx, nz, nt, kernel, both)
    d, flags, ws, sci = fw._prepare_run(True, 50)
    s = stream_handle()
    if noalloc:
        d.workspace = None
        d.workspace_bytes = 0
        # two ranges, so the second call starts from the state the first left
        _lib.call("trtm_forward_steps_noalloc", d, 0, nt // 2, 1, s)
        _lib.call("trtm_forward_steps_noalloc", d, nt // 2, nt, 1, s)
    else:
        _lib.call("trtm_forward_steps", d, 0, nt // 2, 1, s)
        _lib.call("trtm_forward_steps", d, nt // 2, nt, 1, s)
    out = _outputs(fw, flags)
    del ws
    return out


def _assert_equal(a, b):
    assert sorted(a) == sorted(b)
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    assert np.abs(a["rec_w"]).max() > 0


@pytest.mark.parametrize("kernel", ["blocked", "fused", "auto"])
def test_noalloc_equals_workspace_call_twice_shrinking(kernel):
    for nx, nz, nt, both in ((96, 80, 121, True), (64, 48, 90, False)):
        print("noalloc %s %dx%dx%d both=%s" % (kernel, nx, nz, nt, both), flush=True)
        ref = _run(nx, nz, nt, kernel, noalloc=False, both=both)
        got = _run(nx, nz, nt, kernel, noalloc=True, both=both)
        _assert_equal(got, ref)


def test_noalloc_rejects_grid_beyond_scratch_and_caller_workspace():
    fw = _problem(64, 48, 20, "blocked")
    d, flags, ws, sci = fw._prepare_run(True, 50)
    s = stream_handle()
    with pytest.raises(ValueError, match="workspace"):  # a caller workspace is refused
        _lib.call("trtm_forward_steps_noalloc", d, 0, 10, 1, s)
    big = _lib.StepDesc.from_buffer_copy(d)
    big.workspace, big.workspace_bytes = None, 0
    big.nx, big.nz = 4000, 4000  # rejected on the host, before any launch
    with pytest.raises(ValueError, match="static scratch"):
        _lib.call("trtm_forward_steps_noalloc", big, 0, 10, 1, s)
    assert _lib.NOALLOC_SCRATCH_BYTES == 64 * 1024 * 1024
    torch.cuda.synchronize()
    assert int(flags.sum()) == 0
