"""Shared drivers for the parity tests: one problem run through the HIP path
(taichirtm_amd) and through the oracle with identical float32 inputs."""
import numpy as np

import oracle


def run_hip(kw, wavelets, isnap=1, kernel="fused", obs=None, sci=100, surface=None,
            callback=None, backward=True, fw_opts=None, bw_opts=None):
    """fw_opts / bw_opts: extra keywords of the forward / backward propagator
    only (e.g. a pinned tile shape per direction)."""
    import taichirtm_amd as trtm
    kw = dict(kw)
    kw.pop("kernel", None)
    if surface is not None:
        kw["surface_matrix_field"] = surface
    wu, wv, ww = wavelets
    fw = trtm.ForwardModeling(wavelet_u_field=wu, wavelet_v_field=wv, wavelet_w_field=ww,
                              isnap=isnap, kernel=kernel, **kw, **(fw_opts or {}))
    code_f = fw.run(save=True, display_callback=callback, stability_check_interval=sci)
    seis = fw.get_seismogram()
    out = {"code_f": code_f, "seis": seis, "fw": fw,
           "state_f": {k: getattr(fw, k).to_numpy() for k in oracle.ST_NAMES}}
    if not backward:
        return out
    ob = seis if obs is None else obs
    bw = trtm.BackwardModeling(obsdata_u_field=ob[0], obsdata_v_field=ob[1],
                               obsdata_w_field=ob[2], kernel=kernel, **kw, **(bw_opts or {}))
    out["code_b"] = bw.run_calc(fw.u_save_field, fw.v_save_field, fw.w_save_field,
                                fw.isnaps_field, fw.num_snaps, isnap,
                                stability_check_interval=sci)
    out["image"] = bw.get_results()
    out["state_b"] = {k: getattr(bw, k).to_numpy() for k in oracle.ST_NAMES}
    out["bw"] = bw
    return out


def run_oracle(kw, wavelets, isnap=1, obs=None, sci=100, surface=None, dtype=np.float32,
               backward=True):
    m = oracle.Model(kw["mu_field"], kw["lam_field"], kw["rho_field"], kw["dx"], kw["dz"],
                     kw["fs"], kw["nt"], kw["src_loc_field"], kw["recv_loc_field"],
                     kw["absorbing_frame"], isnap=isnap, sci=sci, surface_matrix=surface,
                     dtype=dtype)
    fw = m.forward(wavelets, save=True)
    out = {"code_f": fw["code"], "it_f": fw["it_stop"], "seis": fw["seis"],
           "state_f": fw["state"], "isnaps": fw["snaps"][3]}
    if not backward:
        return out
    ob = fw["seis"] if obs is None else obs
    snaps = fw["snaps"]
    bw = m.backward(ob, snaps[:3], snaps[3], fw["num_snaps"], isnap)
    out.update(code_b=bw["code"], it_b=bw["it_stop"], image=bw["result"], state_b=bw["state"],
               synsrc=bw["synsrc"])
    return out


def same(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def assert_bitwise(h, o, backward=True):
    assert h["code_f"] == o["code_f"]
    for c in range(3):
        assert same(h["seis"][c], o["seis"][c]), "seismogram %d" % c
    for k in oracle.ST_NAMES:
        assert same(h["state_f"][k], o["state_f"][k]), "forward " + k
    if backward:
        assert h["code_b"] == o["code_b"]
        for c in range(3):
            assert same(h["image"][c], o["image"][c]), "image %d" % c
        for k in oracle.ST_NAMES:
            assert same(h["state_b"][k], o["state_b"][k]), "backward " + k
