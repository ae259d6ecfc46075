"""Host tests of the results-file pipeline (imaging.py:29-129,251-277;
rtm.py:1031-1070): stacking against sums written out element by element,
files written the reference's way (surface_matrix=None, a pickled entry),
and files written by save_result read back as the reference would."""
import os

import numpy as np
import pytest

from taichirtm_amd import imaging
from taichirtm_amd.rtm import ReverseTimeMigration

REF_KEYS = {"u", "v", "w", "receiver_loc", "src_loc_step", "offset", "xmin", "xmax", "zmin",
            "zmax", "dx", "dz", "nx", "nz", "surface_matrix"}


def _reference_file(path, u, v, w, offset, dx=2.5, dz=2.5, surface_matrix=None):
    """A result file laid out as rtm.py:1043-1069 writes it."""
    nx, nz = u.shape
    np.savez_compressed(path, u=u, v=v, w=w, receiver_loc=np.array([1.0, 2.0]),
                        src_loc_step=np.array([[nx // 2, 1]]), offset=offset,
                        xmin=float(-offset), xmax=float(dx * nx - offset), zmin=0.0,
                        zmax=float(dz * nz), dx=dx, dz=dz, nx=nx, nz=nz,
                        surface_matrix=surface_matrix)


def _images(rng, n, shape=(6, 4)):
    return [tuple(rng.standard_normal(shape).astype(np.float32) for _ in range(3))
            for _ in range(n)]


def test_stack_is_float32_running_sum_transposed_minus_mean(tmp_path):
    rng = np.random.default_rng(3)
    imgs = _images(rng, 3)
    # values whose float32 sum depends on the order: 1e8 + 1 - 1e8 = 0 in
    # float32 when summed left to right, 1 in exact arithmetic
    imgs[0][0][0, 0], imgs[1][0][0, 0], imgs[2][0][0, 0] = 1e8, 1.0, -1e8
    results = []
    for k, (u, v, w) in enumerate(imgs):
        results.append({"u": u, "v": v, "w": w, "offset": 10.0 + k, "dx": 2.5, "dz": 2.0,
                        "nx": 6, "nz": 4})
    out = imaging.stack_rtm_images(results)
    for c, name in enumerate("uvw"):
        s = np.zeros((6, 4), np.float32)
        for img in imgs:  # element by element, in shot order, float32
            for i in range(6):
                for j in range(4):
                    s[i, j] = np.float32(s[i, j] + img[c][i, j])
        t = s.T
        assert out[name].shape == (4, 6)
        assert np.array_equal(out[name], t - np.mean(t))
    # extent and grid from the last file
    assert out["xmin"] == -12.0 and out["xmax"] == 2.5 * 6 - 12.0
    assert out["zmin"] == 0.0 and out["zmax"] == 2.0 * 4
    assert out["extent"] == [out["xmin"], out["xmax"], out["zmax"], out["zmin"]]
    assert (out["dx"], out["dz"], out["nx"], out["nz"]) == (2.5, 2.0, 6, 4)
    raw = imaging.stack_rtm_images(results, subtract_mean=False)
    assert np.array_equal(raw["w"], (imgs[0][2] + imgs[1][2] + imgs[2][2]).T)
    assert raw["u"][0, 0] == 0.0  # (1e8 + 1) - 1e8 in float32, shot order


def test_stack_empty_raises():
    with pytest.raises(ValueError, match="No results to stack"):
        imaging.stack_rtm_images([])


def test_load_reference_written_files_and_stack(tmp_path):
    rng = np.random.default_rng(4)
    imgs = _images(rng, 2)
    for k, (u, v, w) in enumerate(imgs):
        _reference_file(str(tmp_path / ("shot%d.npz" % k)), u, v, w, offset=5.0 * k)
    res = sorted(imaging.load_rtm_results(str(tmp_path)), key=lambda r: r["filepath"])
    assert len(res) == 2
    for r, (u, v, w) in zip(res, imgs):
        assert set(r) == REF_KEYS | {"filepath"}
        assert r["surface_matrix"] is None  # the pickled None, not unpickled
        assert np.array_equal(r["u"], u) and np.array_equal(r["w"], w)
        assert os.path.basename(r["filepath"]).startswith("shot")
    out = imaging.stack_rtm_images(res)
    t = (imgs[0][1] + imgs[1][1]).T
    assert np.array_equal(out["v"], t - np.mean(t))
    assert out["xmin"] == -5.0


def test_load_reference_file_with_a_surface_matrix(tmp_path):
    rng = np.random.default_rng(5)
    (u, v, w), = _images(rng, 1)
    sm = (rng.random(u.shape) > 0.5).astype(np.float32)
    _reference_file(str(tmp_path / "a.npz"), u, v, w, offset=1.0, surface_matrix=sm)
    r, = imaging.load_rtm_results(str(tmp_path))
    assert np.array_equal(r["surface_matrix"], sm)


def _rtm_with_result(surface=None):
    r = ReverseTimeMigration().set_receivers([0.0, 1.0, 2.0])
    r.image_u, r.image_v, r.image_w = (np.full((5, 3), k, np.float32) for k in (1, 2, 3))
    r.dx, r.dz, r.nx, r.nz, r.offset = 0.5, 0.25, 5, 3, 2.0
    r.src_loc_step = [[2, 1]]
    if surface is not None:
        r.surface_matrix = surface
    return r


@pytest.mark.parametrize("with_surface", [False, True])
def test_save_result_matches_reference_layout(tmp_path, with_surface):
    """save_result's file holds the keys, dtypes and values the reference's
    save_result would write for the same state; the reference's loader
    (allow_pickle=True, run here on this build's own file) sees the same."""
    sm = np.ones((5, 3), np.float32) if with_surface else None
    r = _rtm_with_result(sm)
    r.save_result(str(tmp_path), "ours")
    _reference_file(str(tmp_path / "ref.npz"), r.image_u, r.image_v, r.image_w, r.offset,
                    r.dx, r.dz, surface_matrix=sm)
    ours = np.load(str(tmp_path / "ours.npz"), allow_pickle=True)
    ref = np.load(str(tmp_path / "ref.npz"), allow_pickle=True)
    assert set(ours.files) == set(ref.files) == REF_KEYS
    for k in REF_KEYS - {"receiver_loc", "src_loc_step"}:
        a, b = ours[k], ref[k]
        assert a.dtype == b.dtype and a.shape == b.shape, k
        if a.dtype != object:
            assert np.array_equal(a, b), k
        else:
            assert a.item() is None and b.item() is None
    back, = [x for x in imaging.load_rtm_results(str(tmp_path), "ours.npz")]
    assert (back["surface_matrix"] is None) == (sm is None)


def test_save_stacked_results_round_trip(tmp_path):
    rng = np.random.default_rng(6)
    imgs = _images(rng, 2)
    out = imaging.stack_rtm_images([{"u": u, "v": v, "w": w, "offset": 3.0, "dx": 1.0,
                                     "dz": 1.0, "nx": 6, "nz": 4} for u, v, w in imgs])
    path = str(tmp_path / "sub" / "stack.npz")
    imaging.save_stacked_results(out, path)
    f = np.load(path, allow_pickle=False)
    assert set(f.files) == {"u", "v", "w", "xmin", "xmax", "zmin", "zmax"}
    for k in ("u", "v", "w"):
        assert np.array_equal(f[k], out[k])
    assert float(f["xmin"]) == -3.0 and float(f["zmax"]) == 4.0
