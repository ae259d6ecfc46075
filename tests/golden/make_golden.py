"""Writes tests/golden/tiny_rtm.npz: a tiny P-SV + SH problem with its oracle
outputs (float32 and float64), used as a committed regression fixture.

    python tests/golden/make_golden.py

Inputs: 24 x 20 grid, absorbing_frame 4, 60 steps, isnap 3, a source with all
three force components at (12, 3) and a second one on the same cell,
receivers on the surface row including both border cells plus a duplicate.
The medium is a seeded random layered model in g/cm^3 / kPa units.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

PATH = os.path.join(HERE, "tiny_rtm.npz")


def inputs():
    rng = np.random.default_rng(2024)
    nx, nz, nt = 24, 20, 60
    vs = rng.uniform(900.0, 1400.0, (nx, nz)).astype(np.float32)
    vp = (vs * rng.uniform(1.7, 2.0, (nx, nz))).astype(np.float32)
    rho = rng.uniform(1.8, 2.4, (nx, nz)).astype(np.float32)
    mu = (rho * vs ** 2).astype(np.float32)
    lam = (((vp / vs) ** 2 - 2) * mu).astype(np.float32)
    t = np.arange(nt) * 4e-4
    a = (np.pi * 60.0 * (t - 1 / 60.0)) ** 2
    w = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)
    wav = np.stack([np.stack([w, 0.5 * w]), np.stack([-0.3 * w, w]), np.stack([w, -w])])
    src = np.array([[12, 3], [12, 3]], np.int32)
    rec = np.array([[i, 1] for i in range(0, nx, 3)] + [[nx - 1, 1], [6, 1], [10, 12]], np.int32)
    return dict(nx=nx, nz=nz, nt=nt, dx=5.0, dz=4.0, fs=2500.0, fw=4, isnap=3,
                mu=mu, lam=lam, rho=rho, wavelets=wav.astype(np.float32), src=src, rec=rec)


def outputs(inp, dtype):
    m = oracle.Model(inp["mu"], inp["lam"], inp["rho"], inp["dx"], inp["dz"], inp["fs"],
                     inp["nt"], inp["src"], inp["rec"], inp["fw"], isnap=inp["isnap"], sci=1,
                     dtype=dtype)
    r = oracle.rtm(m, list(inp["wavelets"]), isnap=inp["isnap"])
    fw, bw = r["fw"], r["bw"]
    out = {"seis": np.stack(fw["seis"]), "state_f": np.stack([fw["state"][k] for k in oracle.ST_NAMES]),
           "image": np.stack(bw["result"]),
           "state_b": np.stack([bw["state"][k] for k in oracle.ST_NAMES]),
           "isnaps": fw["snaps"][3]}
    return out


def generate():
    inp = inputs()
    data = {k: np.asarray(v) for k, v in inp.items()}
    for tag, dt in (("f32", np.float32), ("f64", np.float64)):
        for k, v in outputs(inp, dt).items():
            data["%s_%s" % (tag, k)] = v
    return data


if __name__ == "__main__":
    np.savez_compressed(PATH, **generate())
    print("wrote", PATH)
