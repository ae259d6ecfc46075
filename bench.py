"""Benchmark: elastic FD / RTM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sh_layered]

One step = one single-shot RTM of the workload: forward modeling with a
snapshot every isnap steps, then backward propagation of the recorded
seismogram with the zero-lag image accumulated in the step kernel
(the reference's rtm.py:935-1000 sequence).  Inputs (medium, wavelets) are
resident in HBM before the timed region; the state is reset between steps.

value  = useful cell-updates / second (2 * nx * nz * nt per shot: forward and
         backward steps; nothing is recomputed).
roofline: the dominant kernel is the backward step with imaging; its average
         launch time is measured with HIP events on the launch stream over the
         timed steps; algorithmic bytes per cell per launch are listed in
         ALGO_BYTES.  frac is against the 8 TB/s HBM peak; frac_of_copy
         against copy_gbs, the bandwidth of a 1 GiB device-to-device copy
         measured in the same process (read + write bytes).
f64_rel_l2: the timed run's forward seismograms against the same scheme in
         double (relative L2), computed after the timed region.
cpu_baseline: the oracle (C/OpenMP restatement of the reference's unfused
         per-step kernel sequence, float32, both wave systems updated as the
         reference does) timed on a bounded sample of the same workload.
kernel:  "tuned" (default) lets the autotuner pick the blocked kernel family
         and shape per direction during the warmup step.
graph:   after the warmup the forward and the backward run are each recorded
         into a HIP graph (trtm_plan_*) and every later step replays them (one
         more untimed step replays them first); --no-graph issues the step
         launches eagerly.

Other workloads (--workload psv_batch | fwi | salt): the batched
checkpointed P-SV RTM of the 2000 x 1000 faulted model (64 shots in batches
of 16), the P-SV FWI gradient of the 600 x 300 x 2500 problem (16 shots), and
the FWI gradient plus RTM image of the 4000 x 2000 x 8000 salt model (32
shots) under a device-memory budget (--budget-gb; --shots N for fewer
shots); one step is the whole job, the roofline is job-level (useful
cell-updates x algorithmic bytes / wall).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCell-updates/s + RTM shots/hour on 1 MI355X; % HBM roofline; vs Taichi CPU"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per cell per launch (float32 words; sponge from the
# cell index, 0 bytes; source/receiver traffic ignored):
#   SH forward step with snapshot:   v, syx, syz read+write 24 + myx, myz, ir 12 + snapshot write 4
#   SH backward step with imaging:   24 + 12 + forward snapshot read 4 + image read+write 8
#   P-SV forward with snapshot:      u, w, sxx, szz, sxz r+w 40 + lam, mu, mxz, iru, irw 20 + u, w snapshot 8
#   P-SV backward with imaging:      40 + 20 + u, w snapshot 8 + u, w image r+w 16
ALGO_BYTES = {
    ("sh", "forward"): 40, ("sh", "backward"): 48,
    ("psv", "forward"): 68, ("psv", "backward"): 84,
}


# measured HBM bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, copy-calibrated)
MEASURED_TRAFFIC = {
    ("sh_layered_1200x600x4000", "blocked", (32, 128, 1024, 8)): 57.4e6,
    ("sh_layered_1200x600x4000", "blocked", (24, 128, 1024, 8)): 59.4e6,
    ("sh_layered_1200x600x4000", "blocked", (32, 48, 512, 8)): 59.8e6,
    ("sh_layered_1200x600x4000", "blocked", (48, 64, 1024, 8)): 60.0e6,
}


def build_workload(name):
    from taichirtm_amd import configs
    if name == "sh_layered":
        return configs.sh_layered(), "sh"
    if name == "psv_anomaly":
        return configs.psv_anomaly(), "psv"
    raise SystemExit("unknown workload %r" % name)


# algorithmic bytes per useful cell-update of a whole job (forward and
# backward/adjoint updates averaged): P-SV RTM with imaging every step
# (68 + 84) / 2; P-SV FWI forward 60 and adjoint-with-gradient 120
JOB_BYTES = {"psv_batch": (68 + 84) / 2.0, "fwi": (60 + 120) / 2.0,
             "salt": (60 + 120 + 68 + 84) / 4.0}


def _shot_specs(cfg, n):
    from taichirtm_amd import configs
    from taichirtm_amd.batch import ShotSpec
    out = []
    for s in cfg.shots[:n]:
        w = configs.ricker(cfg.f0, s.nt, cfg.dt)[None]
        z = np.zeros_like(w)
        out.append(ShotSpec(np.array([s.src], np.int32), (z, z, w), np.asarray(s.rec, np.int32)))
    return out


def job_main(args):
    """Batched checkpointed P-SV RTM (psv_batch), the P-SV FWI gradient (fwi)
    or the budgeted gradient + image of the salt model (salt): one step = the
    whole job over --shots shots."""
    import torch
    from taichirtm_amd import autotune, configs
    copy_gbs = copy_bandwidth()  # before the job holds its memory budget
    # defaults: the full survey of each workload and a memory budget that
    # fits one MI355X beside it
    full = {"psv_batch": (64, 16, 200.0), "fwi": (16, 0, 200.0), "salt": (32, 0, 128.0)}
    if args.workload in full:
        n, g, gb = full[args.workload]
        args.shots = args.shots or n
        args.group = args.group if args.group is not None else g
        args.budget_gb = args.budget_gb if args.budget_gb is not None else gb
    if args.workload == "psv_batch":
        from taichirtm_amd.batch import BatchRTM
        cfg = configs.psv_faulted()
        shots = _shot_specs(cfg, args.shots)
        mu, lam = cfg.moduli()

        group = args.group or len(shots)

        def step():
            # shots in groups of --group (each group one batch), images
            # stacked in float64 in shot order
            stack = None
            for g0 in range(0, len(shots), group):
                sub = shots[g0:g0 + group]
                b = BatchRTM(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, mu, lam, cfg.rho, cfg.fw, sub,
                             isnap=args.isnap, imaging="checkpoint", kernel=args.kernel,
                             tile_x=args.tile_x, tile_z=args.tile_z, block_steps=args.block_steps,
                             snapshot_bytes=int(args.budget_gb * 1e9))
                if b.run() != [0] * len(sub):
                    raise RuntimeError("non-finite wavefield")
                img = b.stacked_image()
                stack = img if stack is None else {c: stack[c].add_(img[c]) for c in stack}
                tuned = dict(b.tuned)
                tuned["kept_segments"] = [len(b.edges) - 1 - b.first_kept, len(b.edges) - 1]
                del b, img
            return tuned
        desc = ("batched checkpointed P-SV RTM, %d shots of %s (uneven record lengths and "
                "receiver spreads) in batches of %d, isnap=%d, stacked image, segment "
                "wavefields of up to %.0f GB kept from the forward pass"
                % (len(shots), cfg.name, group, args.isnap, args.budget_gb))
    elif args.workload == "fwi":
        from taichirtm_amd.fwi import ElasticFWI, l2_misfit_and_gradient
        cfg, start = configs.psv_fwi()
        shots = _shot_specs(cfg, args.shots)
        fwi_kernel = args.kernel if args.kernel in ("tuned", "blocked", "step") else "auto"
        # the forward pass keeps the history of as many trailing checkpoint
        # segments as --budget-gb holds; the adjoint recomputes the rest
        prob = ElasticFWI(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, shots,
                          precision="float32", kernel=fwi_kernel,
                          history_bytes=int(args.budget_gb * 1e9))
        t = lambda x: torch.as_tensor(np.asarray(x, np.float32), device="cuda")
        obs = [o.clone() for o in prob.forward(t(cfg.vp), t(cfg.vs), t(cfg.rho))]
        m0 = [t(x) for x in (start.vp, start.vs, start.rho)]

        def step():
            l2_misfit_and_gradient(prob, *m0, obs)
            out = dict(prob.shapes)
            out["kept_segments"] = [prob.kept_segments, len(prob.edges) - 1]
            return out
        desc = ("P-SV FWI gradient (forward + checkpointed exact adjoint), %d shots of %s, "
                "forward history of up to %.0f GB kept" % (len(shots), cfg.name, args.budget_gb))
    elif args.workload == "salt":
        from taichirtm_amd import budget
        true, cfg, shot0 = budget.salt_survey(n_shots=32)
        pick = np.linspace(0, len(shot0) - 1, args.shots).round().astype(int)
        shots = [shot0[q] for q in pick]
        obs = budget.observed_data(true, shots)  # input data, made before the timed region
        with_obs = [budget.ShotSpec(s.src_loc, s.wavelets, s.rec_loc, tuple(o))
                    for s, o in zip(shots, obs)]
        del obs
        job = budget.GradientAndImage(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, with_obs,
                                      args.budget_gb * 1e9, kernel=args.kernel)

        def step():
            job.run(cfg.vp, cfg.vs, cfg.rho)
            p = job.plan
            return {"group": [p.group], "L_fwi": [p.L_fwi], "L_rtm": [p.L_rtm]}
        desc = ("P-SV FWI gradient + RTM image together under a %.0f GB budget, %d of the 32 "
                "shots of %s (receiver spreads 1000-3000), layered start model"
                % (args.budget_gb, len(shots), true.name))
    else:
        raise SystemExit("unknown workload %r" % args.workload)
    for k in range(args.warmup):
        t1 = time.perf_counter()
        tuned = step()
        torch.cuda.synchronize()
        print("warmup step %d done, %.1f s" % (k, time.perf_counter() - t1), file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        tuned = step()
        print("timed step %d launched, %.1f s since start" % (k, time.perf_counter() - t0),
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    passes = 4.0 if args.workload == "salt" else 2.0  # salt: FWI and RTM, forward + backward each
    useful = passes * cfg.nx * cfg.nz * sum(s.nt for s in shots)
    g = useful / wall / 1e9
    achieved = useful * JOB_BYTES[args.workload] / wall / 1e9
    out = {
        "metric": METRIC, "value": round(g, 3), "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "grid": [cfg.nx, cfg.nz], "shots": len(shots),
                   "kernel": args.kernel,
                   "tuned": {k: list(v) for k, v in (tuned or {}).items()}},
        "shots_per_hour": round(3600.0 * len(shots) / wall, 1),
        "roofline": {"bound": "hbm", "basis": "job (useful updates x algorithmic bytes / wall)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_update": JOB_BYTES[args.workload]},
    }
    out["roofline"]["copy_gbs"] = round(copy_gbs, 1)
    out["roofline"]["frac_of_copy"] = round(achieved / copy_gbs, 4)
    if not args.no_cpu_baseline:
        # the oracle's RTM of the first shot on a shorter sample (it keeps every
        # snapshot in host memory); per-cell rate of the CPU path on this grid
        out["cpu_baseline"] = cpu_baseline(cfg, target_s=4.0 if args.workload == "salt" else 6.0)
    print(json.dumps(out), flush=True)


class RTMShot:
    """Forward + backward propagators of one shot, built once."""

    def __init__(self, cfg, isnap, kernel, tile_x=0, block_steps=0, tile_z=0):
        import taichirtm_amd as trtm
        shot = cfg.shots[0]
        self.kw = cfg.kwargs(shot)
        self.kw.update(tile_x=tile_x, block_steps=block_steps, tile_z=tile_z)
        wu, wv, ww = cfg.wavelets(shot)
        self.isnap = isnap
        self.fw = trtm.ForwardModeling(wavelet_u_field=wu, wavelet_v_field=wv,
                                       wavelet_w_field=ww, isnap=isnap, kernel=kernel, **self.kw)
        nr, nt = len(shot.rec), shot.nt
        zero = np.zeros((nr, nt), np.float32)
        # observed data: the forward's own seismogram, copied on the device each step
        obs = [zero if c not in self.fw._seis else self.fw._seis[c].new_zeros((nr, nt))
               for c in ("u", "v", "w")]
        self.obs = obs
        self.nt = nt
        # one forward pass fills the observed data, so the backward propagator
        # sees which wave systems carry data when it is built
        if self.fw.run(save=True):
            raise RuntimeError("forward modeling became non-finite")
        self._copy_obs()
        self.bw = trtm.BackwardModeling(obsdata_u_field=obs[0], obsdata_v_field=obs[1],
                                        obsdata_w_field=obs[2], kernel=kernel, **self.kw)
        self.cf = self.cb = None

    def capture(self):
        """Record the forward and the backward run each into a HIP graph
        (after the kernels are tuned); later steps replay them."""
        self.cf = self.fw.capture(save=True)
        self.cb = self.bw.capture_calc(self.fw.u_save_field, self.fw.v_save_field,
                                       self.fw.w_save_field, self.fw.isnaps_field,
                                       self.fw.num_snaps, self.isnap)

    def _copy_obs(self):
        from taichirtm_amd import _lib
        for c, t in self.fw._seis.items():
            o = self.obs[("u", "v", "w").index(c)]
            _lib.call("trtm_transpose", self.nt, o.shape[0], ctypes_ptr(t), ctypes_ptr(o),
                      stream())

    def run(self, ev=None):
        self.fw.reset_state()
        self.bw.reset_state()
        if ev:
            ev[0].record()
        if self.cf:
            self.cf.launch()
        else:
            code = self.fw.run(save=True)
        if ev:
            ev[1].record()
        # forward seismogram -> observed data of the backward pass (device transpose)
        self._copy_obs()
        if ev:
            ev[2].record()
        if self.cb:
            self.cb.launch()
        else:
            code_b = self.bw.run_calc(self.fw.u_save_field, self.fw.v_save_field,
                                      self.fw.w_save_field, self.fw.isnaps_field,
                                      self.fw.num_snaps, self.isnap)
        if ev:
            ev[3].record()
        if self.cf:
            code, code_b = self.cf.result(), self.cb.result()
        if code or code_b:
            raise RuntimeError("non-finite wavefield (codes %d, %d)" % (code, code_b))


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def stream():
    from taichirtm_amd.field import stream_handle
    return stream_handle()


def cpu_baseline(cfg, target_s=12.0):
    """Oracle forward + backward (with imaging every step) on a truncated
    record of the same model; rate in cell-updates per second."""
    import oracle
    shot = cfg.shots[0]
    mu, lam = cfg.moduli()

    def timed(nt):
        m = oracle.Model(mu, lam, cfg.rho, cfg.dx, cfg.dz, cfg.fs, nt, [shot.src], shot.rec,
                         cfg.fw, isnap=1, sci=100)
        wu, wv, ww = (w[:, :nt] for w in cfg.wavelets(shot))
        t0 = time.perf_counter()
        run = oracle.rtm(m, [wu, wv, ww], isnap=1)
        dt = time.perf_counter() - t0
        assert run["fw"]["code"] == 0 and run["bw"]["code"] == 0
        return dt

    probe = 20
    t_probe = timed(probe)
    nt = int(min(shot.nt, max(probe, probe * target_s / max(t_probe, 1e-3))))
    t = timed(nt)
    cells = 2.0 * cfg.nx * cfg.nz * nt
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": cells / t / 1e9, "unit": "GCell-updates/s", "cores": cores,
            "kind": "port",
            "sample": "oracle float32 forward+backward+image (isnap=1), %dx%d grid, %d of %d "
                      "steps each way, %.1f s" % (cfg.nx, cfg.nz, nt, shot.nt, t),
            "shots_per_hour": 3600.0 / (t * shot.nt / nt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sh_layered")
    ap.add_argument("--kernel", default="tuned")
    ap.add_argument("--shots", type=int, default=0,
                    help="jobs: number of shots (default: the full survey, 64 / 16 / 32)")
    ap.add_argument("--group", type=int, default=None,
                    help="psv_batch: shots per batch (0 = all in one batch)")
    ap.add_argument("--tile-x", type=int, default=0)
    ap.add_argument("--block-steps", type=int, default=0)
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--isnap", type=int, default=1)
    ap.add_argument("--budget-gb", type=float, default=None,
                    help="salt: device memory of the job; fwi: forward history kept; "
                         "psv_batch: segment wavefields kept")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="single-shot RTM: issue the step launches eagerly instead of replaying "
                         "the captured forward and backward runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(shots shard across GPUs with one image all-reduce)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    torch.cuda.set_device(0)
    if args.workload in JOB_BYTES:
        return job_main(args)
    cfg, system = build_workload(args.workload)
    shot = RTMShot(cfg, args.isnap, args.kernel, args.tile_x, args.block_steps, args.tile_z)
    for _ in range(args.warmup):
        shot.run()
    graph = not args.no_graph and args.kernel in ("blocked", "wave", "auto", "tuned")
    if graph:  # each pass one HIP graph launch from here on
        shot.capture()
        shot.run()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        shot.run(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fw_ms = [e[0].elapsed_time(e[1]) for e in evs]
    bw_ms = [e[2].elapsed_time(e[3]) for e in evs]

    nt = cfg.nt
    cells = cfg.nx * cfg.nz
    ms_per_step = wall * 1000.0 / args.steps
    gcells = 2.0 * cells * nt / (ms_per_step / 1000.0) / 1e9
    # time steps per launch of the step kernel (1 for the fused and simple kernels)
    blocked = args.kernel in ("blocked", "wave", "auto", "tuned")
    k = shot.bw.block_config()[3] if blocked else 1
    launches = (nt + k - 1) // k
    bw_launch_us = 1000.0 * float(np.mean(bw_ms)) / launches
    fw_launch_us = 1000.0 * float(np.mean(fw_ms)) / launches
    b_bytes = ALGO_BYTES[(system, "backward")] * cells * k
    f_bytes = ALGO_BYTES[(system, "forward")] * cells * k
    achieved = b_bytes / (bw_launch_us * 1e-6) / 1e9
    family = shot.bw._tuned[0] if shot.bw._tuned else args.kernel
    if family == "wave":
        kname = ("k_wave backward, %d steps per launch, one wavefront per tile, state and "
                 "image in registers" % k)
    elif k > 1:
        kname = "k_block backward, %d steps per launch, image accumulated in registers" % k
    else:
        kname = "k_fused backward step with imaging"
    # HBM bytes per launch of the default blocked SH configuration, from the
    # FETCH_SIZE / WRITE_SIZE passes recorded under profiles/
    traffic = MEASURED_TRAFFIC.get((cfg.name, family,
                                    shot.bw.block_config() if blocked else None))
    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "GCell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s single-shot RTM (forward + backward + zero-lag image, isnap=%d)"
                               % (cfg.name, args.isnap),
                   "grid": [cfg.nx, cfg.nz], "nt": nt, "dx": cfg.dx, "dt": cfg.dt,
                   "absorbing_frame": cfg.fw, "receivers": len(cfg.shots[0].rec),
                   "kernel": args.kernel, "tile_x": args.tile_x, "tile_z": args.tile_z,
                   "block_steps": args.block_steps, "system": system, "graph": graph,
                   "chosen": {"forward": list(shot.fw._tuned or ()),
                              "backward": list(shot.bw._tuned or ())}},
        "shots_per_hour": round(3600.0 / (ms_per_step / 1000.0), 1),
        "phases_ms": {"forward": round(float(np.mean(fw_ms)), 3),
                      "backward": round(float(np.mean(bw_ms)), 3)},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": b_bytes,
                     "steps_per_launch": k,
                     "avg_launch_us": round(bw_launch_us, 3),
                     "forward_launch_us": round(fw_launch_us, 3),
                     "forward_achieved": round(f_bytes / (fw_launch_us * 1e-6) / 1e9, 1)},
    }
    copy_gbs = copy_bandwidth()
    out["roofline"]["copy_gbs"] = round(copy_gbs, 1)
    out["roofline"]["frac_of_copy"] = round(achieved / copy_gbs, 4)
    out["f64_rel_l2"] = float64_error(cfg, shot)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Sustained HBM bandwidth of a device-to-device copy of nbytes (read +
    write counted), median over reps, timed with events on the current
    stream, outside the timed region: the measured ceiling the roofline
    fraction is also quoted against."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    del a, b
    return 2.0 * nbytes / (float(np.median(ms)) * 1e-3) / 1e9


def float64_error(cfg, shot):
    """Relative L2 error (||a - b|| / ||b||) of the last timed forward
    seismograms against the same scheme in double (the FWI forward step, whose
    float32 records equal ForwardModeling's bit for bit), outside the timed
    region."""
    import torch
    from taichirtm_amd.batch import ShotSpec
    from taichirtm_amd.fwi import ElasticFWI
    s = cfg.shots[0]
    spec = ShotSpec(np.array([s.src], np.int32),
                    tuple(np.asarray(w, np.float64) for w in cfg.wavelets(s)),
                    np.asarray(s.rec, np.int32))
    p = ElasticFWI(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, [spec], precision="float64",
                   checkpoint_every=s.nt)
    m = [torch.as_tensor(np.asarray(x, np.float64), device="cuda") for x in (cfg.vp, cfg.vs, cfg.rho)]
    ref = p.forward(*m)[0].cpu().numpy()
    got = shot.fw.get_seismogram()
    out = {}
    for c, name in enumerate("uvw"):
        if np.abs(ref[c]).max() > 0:
            a, b = np.asarray(got[c], np.float64), np.asarray(ref[c], np.float64)
            out["seismogram_" + name] = float("%.3e" % (np.linalg.norm(a - b) / np.linalg.norm(b)))
    return out


if __name__ == "__main__":
    main()
