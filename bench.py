"""Benchmark: elastic FD / RTM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload salt]

Default workload (salt): the FWI gradient and the RTM image of the 32-shot
salt survey, P-SV 4000 x 2000 x 8000, computed together under a 128 GB
device-memory budget.  One step is one shot group of the budget's plan (3
shots), taken round robin over the survey's groups; per group, one forward
pass keeps checkpoints, and walking the segments last to first each
segment's history is recomputed once and feeds the exact adjoint (gradient)
and the reference's back-propagation with zero-lag imaging
(backward_modeling.py:539-576).  Observed data (the true model's records) are
computed before the timed region and stay resident on the device.

value  = useful cell-updates / second over the timed groups: 3 per cell-step
         and shot (FWI forward, FWI adjoint, RTM backward with imaging; the
         forward is shared by both backward passes, and the checkpoint pass
         and the recompute are overhead, not counted).
roofline: job-level achieved = useful updates x (60 + 120 + 84) / 3 bytes /
         wall, against the 8 TB/s HBM peak.  roofline.kernel is the dominant
         kernel, the blocked FWI adjoint: its average launch time from HIP
         events on the launch stream around every adjoint call of the timed
         groups, its algorithmic bytes (120 per cell-step and shot), and its
         real HBM bytes per launch (roofline.traffic) from two rocprofv3
         --pmc passes (FETCH_SIZE, WRITE_SIZE) over a short run of the same
         kernel shape on the same grid and shot count, each scaled by what the
         counter reads for a 256 MiB elementwise copy in the same pass.
single_shot: the plain single-shot SH RTM of the 1200 x 600 x 4000 layered
         model (forward + backward + image, isnap 1), graph-replayed.
cpu_baseline: the oracle (C/OpenMP restatement of the reference's unfused
         per-step kernel sequence, float32) timed on a bounded sample of the
         first shot's RTM and extrapolated per cell-update.

Other workloads (--workload psv_batch | fwi | sh_layered | psv_anomaly): the
64-shot batched checkpointed P-SV RTM of the 2000 x 1000 faulted model, the
16-shot P-SV FWI gradient of the 600 x 300 x 2500 problem with sqrt(N)
checkpoints (one step = the whole job), and the single-shot RTMs.  Prints one
JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCell-updates/s + RTM shots/hour on 1 MI355X; % HBM roofline; vs Taichi CPU"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per cell per launch of the single-shot RTM kernels
# (float32 words; sponge from the cell index, 0 bytes; source/receiver
# traffic ignored):
#   SH forward step with snapshot:   v, syx, syz read+write 24 + myx, myz, ir 12 + snapshot write 4
#   SH backward step with imaging:   24 + 12 + forward snapshot read 4 + image read+write 8
#   P-SV forward with snapshot:      u, w, sxx, szz, sxz r+w 40 + lam, mu, mxz, iru, irw 20 + u, w snapshot 8
#   P-SV backward with imaging:      40 + 20 + u, w snapshot 8 + u, w image r+w 16
ALGO_BYTES = {
    ("sh", "forward"): 40, ("sh", "backward"): 48,
    ("psv", "forward"): 68, ("psv", "backward"): 84,
}

# Algorithmic bytes per useful cell-update of a whole job:
#   P-SV forward step 60 (state r+w 40, medium 20); backward with imaging 84
#   (+ forward wavefield 8, image r+w 16); exact adjoint with gradient 120
#   (adjoint state r+w 40, medium 20, history 20, 5 gradients r+w 40).
#   psv_batch: forward + backward with imaging; fwi: forward + adjoint;
#   salt: forward + adjoint + backward with imaging.
JOB_BYTES = {"psv_batch": (60 + 84) / 2.0, "fwi": (60 + 120) / 2.0,
             "salt": (60 + 120 + 84) / 3.0}
JOB_SWEEPS = {"psv_batch": 2, "fwi": 2, "salt": 3}
ADJOINT_BYTES = 120  # per cell-step and shot of k_fwi_adj_blk


def build_workload(name):
    from taichirtm_amd import configs
    if name == "sh_layered":
        return configs.sh_layered(), "sh"
    if name == "psv_anomaly":
        return configs.psv_anomaly(), "psv"
    raise SystemExit("unknown workload %r" % name)


def _shot_specs(cfg, n):
    from taichirtm_amd import configs
    from taichirtm_amd.batch import ShotSpec
    out = []
    for s in cfg.shots[:n]:
        w = configs.ricker(cfg.f0, s.nt, cfg.dt)[None]
        z = np.zeros_like(w)
        out.append(ShotSpec(np.array([s.src], np.int32), (z, z, w), np.asarray(s.rec, np.int32)))
    return out


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


# -- the salt survey: one shot group per step ----------------------------------

def salt_main(args):
    import torch
    from taichirtm_amd import budget, configs
    from taichirtm_amd.batch import ShotSpec
    copy_gbs = copy_bandwidth()  # before the job holds its memory budget
    single = None
    if not args.no_single_shot:
        single = single_shot_line(args, steps=10, warmup=1)
        single.pop("_cfg")
        torch.cuda.empty_cache()
    budget_gb = args.budget_gb if args.budget_gb is not None else 128.0
    true, start, shot0 = budget.salt_survey(n_shots=32)
    n_shots = args.shots or 32
    pick = np.linspace(0, len(shot0) - 1, n_shots).round().astype(int)
    shots = [shot0[q] for q in pick]
    t1 = time.perf_counter()
    obs = budget.observed_data(true, shots, on_device=True)  # inputs, resident in HBM
    _log("observed data of %d shots: %.1f s" % (len(shots), time.perf_counter() - t1))
    with_obs = [ShotSpec(s.src_loc, s.wavelets, s.rec_loc, tuple(o)) for s, o in zip(shots, obs)]
    del obs
    job = budget.GradientAndImage(start.nx, start.nz, start.dx, start.dz, start.fs, start.fw,
                                  with_obs, budget_gb * 1e9, kernel=args.kernel)
    job.prepare(start.vp, start.vs, start.rho)
    n_groups = job.plan.groups
    cells = start.nx * start.nz
    for k in range(args.warmup):
        t1 = time.perf_counter()
        job.run_group(k % n_groups)
        torch.cuda.synchronize()
        _log("warmup group %d done, %.1f s" % (k % n_groups, time.perf_counter() - t1))
    events = []
    job.adjoint_events = events
    order = [(args.warmup + k) % n_groups for k in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, g in enumerate(order):
        job.run_group(g)
        _log("timed step %d (group %d) queued, %.1f s since start"
             % (k, g, time.perf_counter() - t0))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    job.adjoint_events = None
    J, grads, image = job.result()
    if not (np.isfinite(J) and all(bool(torch.isfinite(g).all()) for g in grads)
            and all(bool(torch.isfinite(t).all()) for t in image.values())):
        raise RuntimeError("non-finite misfit, gradient or image")
    shot_steps = sum(s.nt for g in order for s in job.group_shots(g))
    useful = JOB_SWEEPS["salt"] * cells * shot_steps
    ms_per_step = wall * 1e3 / args.steps
    g = useful / wall / 1e9
    achieved = useful * JOB_BYTES["salt"] / wall / 1e9
    shots_timed = sum(len(job.group_shots(g_)) for g_ in order)

    # dominant kernel: the blocked FWI adjoint of the timed groups
    ms = sum(e0.elapsed_time(e1) for e0, e1, _ in events)
    launches = sum(n for _, _, n in events)
    adj_shape = list(job.shapes["adjoint"])
    group = job.plan.group
    adj_bytes = ADJOINT_BYTES * cells * shot_steps  # every timed cell-step and shot
    kernel = {"name": "k_fwi_adj_blk<P-SV, %d x %d tile, %d steps per launch>"
                      % (adj_shape[0], adj_shape[1], adj_shape[2]),
              "shape": adj_shape, "launches": launches,
              "avg_launch_us": round(1e3 * ms / max(launches, 1), 2),
              "share_of_wall": round(ms / 1e3 / wall, 4),
              "algorithmic_bytes_per_launch": int(adj_bytes / max(launches, 1)),
              "achieved": round(adj_bytes / (ms * 1e-3) / 1e9, 1), "unit": "GB/s"}
    kernel["frac"] = round(kernel["achieved"] / HBM_PEAK_GBS, 4)
    plan_L = job.plan.L_fwi
    tuned = {"fwi": {k: list(v) for k, v in job.shapes.items()},
             "rtm_backward": list(job.rtm_tuned.get("backward", ()))}
    del job, with_obs, grads, image
    torch.cuda.empty_cache()
    traffic = None
    if not args.no_traffic and adj_shape[2]:
        traffic = counter_traffic(adj_shape, group, start)
        kernel["counters"] = traffic
        if traffic.get("bytes_per_launch"):
            tb = traffic["bytes_per_launch"]
            kernel["traffic"] = tb
            kernel["traffic_gbs"] = round(tb / (kernel["avg_launch_us"] * 1e-6) / 1e9, 1)
            kernel["traffic_frac"] = round(kernel["traffic_gbs"] / HBM_PEAK_GBS, 4)
    desc = ("psv_salt_4000x2000x8000: P-SV FWI gradient + RTM image together under a %.0f GB "
            "budget, %d shots (receiver spreads 1000-3000), groups of %d, one group per step "
            "(round robin), layered start model" % (budget_gb, len(shots), group))
    out = {
        "metric": METRIC, "value": round(g, 3), "unit": "GCell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded salt model; observed data from the true model)",
        "config": {"workload": desc, "grid": [start.nx, start.nz], "nt": start.nt,
                   "dx": start.dx, "dt": start.dt, "shots": len(shots), "group": group,
                   "groups": n_groups, "timed_groups": order, "checkpoint_every": plan_L,
                   "budget_gb": budget_gb, "kernel": args.kernel, "tuned": tuned},
        "shots_per_hour": round(3600.0 * shots_timed / wall, 1),
        "useful_sweeps": "3 per cell-step and shot: FWI forward (60 B), FWI adjoint (120 B), "
                         "RTM backward with imaging (84 B)",
        "roofline": {"bound": "hbm",
                     "basis": "job: useful cell-updates x 88 B (60 + 120 + 84) / 3 / wall; "
                              "traffic and kernel: the dominant kernel",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": kernel.get("traffic"),
                     "bytes_per_update": JOB_BYTES["salt"],
                     "copy_gbs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4),
                     "kernel": kernel},
    }
    out["f64_rel_l2"] = salt_float64_error(start, shots[0])
    if single is not None:
        out["single_shot"] = single
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(start, target_s=12.0, sweeps=3)
    print(json.dumps(out), flush=True)


def salt_float64_error(cfg, spec):
    """Relative L2 of the first shot's float32 records in the start model
    against the same scheme in double."""
    import torch
    from taichirtm_amd.batch import ShotSpec
    from taichirtm_amd.fwi import ElasticFWI
    out = {}
    recs = []
    for prec in ("float32", "float64"):
        dt = np.float32 if prec == "float32" else np.float64
        s = ShotSpec(spec.src_loc, tuple(np.asarray(w, dt) for w in spec.wavelets), spec.rec_loc)
        p = ElasticFWI(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, [s], precision=prec,
                       history_bytes=0, kernel="auto")
        m = [torch.as_tensor(np.asarray(x, dt), device="cuda") for x in (cfg.vp, cfg.vs, cfg.rho)]
        recs.append(p.records(*m)[0].cpu().numpy().astype(np.float64))
        del p, m
    for c, name in enumerate("uvw"):
        b = recs[1][c]
        if np.abs(b).max() > 0:
            out["seismogram_" + name] = float("%.3e" % (np.linalg.norm(recs[0][c] - b)
                                                         / np.linalg.norm(b)))
    return out


# -- real HBM bytes of the dominant kernel, from rocprofv3 counters ------------

def counter_traffic(shape, shots, cfg, nt=48, timeout=240):
    """FETCH_SIZE and WRITE_SIZE per launch of the blocked FWI adjoint with
    ``shape`` (tile_x, tile_z, block_steps) on ``shots`` shots of cfg's grid,
    each from a rocprofv3 --pmc pass over a short run of this script
    (--probe), scaled by the counter's reading for a 256 MiB elementwise copy
    (negation) in the same pass.  Returns a dict; bytes_per_launch is None
    with a reason when the profiler is not available or fails."""
    import shutil
    out = {"counters": ["FETCH_SIZE", "WRITE_SIZE"], "probe_steps": nt, "shots": shots}
    prof = shutil.which("rocprofv3")
    if prof is None:
        out["bytes_per_launch"] = None
        out["note"] = "rocprofv3 not found"
        return out
    env = dict(os.environ, TMPDIR="/tmp")
    total = 0.0
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory(prefix="trtm_pmc_", dir="/tmp") as td:
            cmd = ["timeout", "-s", "KILL", str(timeout), prof, "--pmc", counter,
                   "--output-format", "csv", "-d", td, "-o", "run", "--",
                   sys.executable, os.path.join(ROOT, "bench.py"), "--probe", "adjoint",
                   "--probe-shape", ",".join(str(x) for x in shape),
                   "--shots", str(shots), "--probe-steps", str(nt)]
            t1 = time.perf_counter()
            r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
            _log("counter pass %s: rc %d, %.1f s" % (counter, r.returncode,
                                                    time.perf_counter() - t1))
            vals = _read_counters(td, counter)
        if not vals.get("adjoint") or not vals.get("copy"):
            out["bytes_per_launch"] = None
            out["note"] = "%s pass gave no readings (rc %d): %s" % (
                counter, r.returncode, (r.stderr or r.stdout)[-300:])
            return out
        adj = float(np.mean(vals["adjoint"]))
        cal = float(np.mean(vals["copy"]))
        copy_bytes = (1 << 28) * (1 if counter == "FETCH_SIZE" else 1)
        scale = copy_bytes / (cal * 1024.0)
        key = counter.lower()
        out[key + "_kib"] = round(adj, 1)
        out[key + "_copy_kib"] = round(cal, 1)
        out[key + "_scale"] = round(scale, 4)
        out[key + "_dispatches"] = len(vals["adjoint"])
        total += adj * 1024.0 * scale
    out["bytes_per_launch"] = int(total)
    out["command"] = ("rocprofv3 --pmc <FETCH_SIZE | WRITE_SIZE> --output-format csv -- python3 "
                      "bench.py --probe adjoint --probe-shape %s --shots %d --probe-steps %d"
                      % (",".join(str(x) for x in shape), shots, nt))
    return out


def _read_counters(folder, counter):
    """Per-dispatch values of ``counter`` (KiB) from rocprofv3's counter
    CSV: {"adjoint": [...], "copy": [...]}, summed over the rows of a
    dispatch."""
    import csv
    import glob
    per = {}
    for f in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                kind = ("adjoint" if "k_fwi_adj_blk" in name else
                        "copy" if "neg_kernel" in name else None)
                if kind is None:
                    continue
                key = (kind, row.get("Dispatch_Id") or row.get("Correlation_Id"))
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    out = {"adjoint": [], "copy": []}
    for (kind, _), v in per.items():
        out[kind].append(v)
    return out


def probe_main(args):
    """The counter passes' workload: a 256 MiB elementwise copy (negation,
    the calibration), then the blocked FWI adjoint with the given shape on
    --shots shots of the salt survey's start model over --probe-steps steps
    (forward history kept, no recompute)."""
    import torch
    from taichirtm_amd import budget
    from taichirtm_amd.batch import ShotSpec
    from taichirtm_amd.fwi import ElasticFWI
    torch.cuda.set_device(0)
    a = torch.ones(1 << 26, dtype=torch.float32, device="cuda")  # 256 MiB
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    del a, b
    true, start, shots = budget.salt_survey(n_shots=32)
    nt = args.probe_steps
    sub = [ShotSpec(s.src_loc, tuple(np.asarray(w)[:, :nt] for w in s.wavelets), s.rec_loc)
           for s in shots[:args.shots]]
    tx, tz, k = (int(x) for x in args.probe_shape.split(","))
    p = ElasticFWI(start.nx, start.nz, start.dx, start.dz, start.fs, start.fw, sub,
                   precision="float32", kernel="blocked", tile_x=tx, tile_z=tz, block_steps=k,
                   checkpoint_every=nt, history_bytes=1 << 40)
    m = [torch.as_tensor(np.asarray(x, np.float32), device="cuda")
         for x in (start.vp, start.vs, start.rho)]
    recs = p.forward(*m)
    res = [torch.randn_like(r) for r in recs]
    p.gradient(res)
    torch.cuda.synchronize()
    print("probe done: %d shots, %d steps, adjoint shape %s" % (len(sub), nt, args.probe_shape),
          flush=True)


# -- batched jobs: one step = the whole job -------------------------------------

def job_main(args):
    """Batched checkpointed P-SV RTM (psv_batch) or the P-SV FWI gradient
    with sqrt(N) checkpoints (fwi): one step = the whole job over --shots
    shots."""
    import torch
    from taichirtm_amd import configs
    copy_gbs = copy_bandwidth()  # before the job holds its memory budget
    full = {"psv_batch": (64, 16, 200.0), "fwi": (16, 0, 0.0)}
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
                             snapshot_bytes=int(args.budget_gb * 1e9), graph=args.graph)
                if b.run() != [0] * len(sub):
                    raise RuntimeError("non-finite wavefield")
                img = b.stacked_image()
                stack = img if stack is None else {c: stack[c].add_(img[c]) for c in stack}
                tuned = dict(b.tuned)
                tuned["kept_segments"] = [len(b.edges) - 1 - b.first_kept, len(b.edges) - 1]
                del b, img
            return tuned
        desc = ("psv_faulted_2000x1000: batched checkpointed P-SV RTM, %d shots (uneven record "
                "lengths and receiver spreads) in batches of %d, isnap=%d, stacked image, segment "
                "wavefields of up to %.0f GB kept from the forward pass"
                % (len(shots), group, args.isnap, args.budget_gb))
    else:
        from taichirtm_amd.fwi import ElasticFWI, l2_misfit_and_gradient
        cfg, start = configs.psv_fwi()
        shots = _shot_specs(cfg, args.shots)
        fwi_kernel = args.kernel if args.kernel in ("tuned", "blocked", "step") else "auto"
        t = lambda x: torch.as_tensor(np.asarray(x, np.float32), device="cuda")
        m0 = [t(x) for x in (start.vp, start.vs, start.rho)]
        mt = [t(x) for x in (cfg.vp, cfg.vs, cfg.rho)]
        if args.budget_gb > 0:
            # plain residual against observed records, the history of up to
            # --budget-gb of trailing segments kept by the forward pass
            prob = ElasticFWI(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, shots,
                              precision="float32", kernel=fwi_kernel,
                              history_bytes=int(args.budget_gb * 1e9))
            obs = [o.clone() for o in prob.records(*mt)]
            run = lambda: l2_misfit_and_gradient(prob, *m0, obs)
            how = "forward history of up to %.0f GB kept, residual of float32 records" % args.budget_gb
        else:
            # sqrt(N) checkpoints, every segment recomputed; the residual
            # against the true model's data propagated as a scattered field
            nt_max = max(s.nt for s in shots)
            prob = ElasticFWI(cfg.nx, cfg.nz, cfg.dx, cfg.dz, cfg.fs, cfg.fw, shots,
                              precision="float32", kernel=fwi_kernel, history_bytes=0,
                              checkpoint_every=int(round(math.sqrt(nt_max))))
            run = lambda: l2_misfit_and_gradient(prob, *m0, reference=mt)
            how = ("checkpoints every sqrt(N) steps and every segment recomputed; residual "
                   "against the true model's data as a float32 scattered field")

        def step():
            run()
            out = dict(prob.shapes)
            out["kept_segments"] = [prob.kept_segments, len(prob.edges) - 1]
            out["checkpoint_every"] = [prob.L]
            return out
        desc = ("psv_fwi_600x300x2500: P-SV FWI gradient (forward + exact adjoint), %d shots, %s"
                % (len(shots), how))
    tuned = None
    for k in range(args.warmup):
        t1 = time.perf_counter()
        tuned = step()
        torch.cuda.synchronize()
        _log("warmup step %d done, %.1f s" % (k, time.perf_counter() - t1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        tuned = step()
        _log("timed step %d launched, %.1f s since start" % (k, time.perf_counter() - t0))
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    useful = JOB_SWEEPS[args.workload] * cfg.nx * cfg.nz * sum(s.nt for s in shots)
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
    if args.workload == "fwi":
        del prob
        torch.cuda.empty_cache()
        out["f64_rel_l2"] = fwi_float64_error(cfg, start, scattered=args.budget_gb == 0)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, target_s=12.0, sweeps=2)
    print(json.dumps(out), flush=True)


def fwi_float64_error(true, start, scattered=True):
    """Relative L2 of the first shot's float32 gradient (the timed path: the
    scattered-field residual, or the plain one) against the float64 run of
    the same scheme, per parameter, and of its residual."""
    import torch
    from taichirtm_amd.batch import ShotSpec
    from taichirtm_amd.fwi import ElasticFWI, l2_misfit_and_gradient
    s = true.shots[0]
    out = {}
    grads, ress = {}, {}
    for prec, dtype in (("float64", torch.float64), ("float32", torch.float32)):
        npt = np.float64 if prec == "float64" else np.float32
        w = configs_ricker(true, s).astype(npt)
        z = np.zeros_like(w)
        shot = ShotSpec(np.array([s.src], np.int32), (z, z, w), np.asarray(s.rec, np.int32))
        p = ElasticFWI(true.nx, true.nz, true.dx, true.dz, true.fs, true.fw, [shot],
                       precision=prec, history_bytes=0)
        m = [torch.as_tensor(np.asarray(x, np.float32), dtype=dtype, device="cuda")
             for x in (start.vp, start.vs, start.rho)]
        mt = [torch.as_tensor(np.asarray(x, np.float32), dtype=dtype, device="cuda")
              for x in (true.vp, true.vs, true.rho)]
        if prec == "float32" and scattered:
            ress[prec] = p.residual(*m, mt)[0].double()
            grads[prec] = l2_misfit_and_gradient(p, *m, reference=mt)[1]
        else:
            obs = p.records(*mt)[0].clone()
            ress[prec] = (p.forward(*m)[0] - obs).double()
            grads[prec] = p.gradient([ress[prec].to(dtype)])
        del p
    rel = lambda a, b: float(torch.linalg.norm((a.double() - b.double()).ravel())
                             / torch.linalg.norm(b.double().ravel()))
    out["residual"] = float("%.3e" % rel(ress["float32"], ress["float64"]))
    for name, a, b in zip(("vp", "vs", "rho"), grads["float32"], grads["float64"]):
        out["gradient_" + name] = float("%.3e" % rel(a, b))
    out["sample"] = "shot 0, float32 %s residual, against float64" % (
        "scattered-field" if scattered else "records")
    return out


def configs_ricker(cfg, shot):
    from taichirtm_amd import configs
    return configs.ricker(cfg.f0, shot.nt, cfg.dt)[None]


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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, target_s=12.0, sweeps=2):
    """Oracle forward + backward (with imaging every step) of the first shot
    on a truncated record of the same model (the oracle keeps every snapshot
    in host memory); rate in cell-updates per second, extrapolated to the
    full record.  shots_per_hour: a shot of ``sweeps`` sweeps of nt steps at
    that rate."""
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
    rate = cells / t
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": rate / 1e9, "unit": "GCell-updates/s", "cores": cores,
            "kind": "port", "label": "reference-equivalent CPU restatement",
            "cpu_model": _cpu_model(), "extrapolated": nt < shot.nt,
            "sample": "oracle float32 forward + backward + image (isnap=1) of shot 0, %dx%d grid, "
                      "%d of %d steps each way, %.1f s%s"
                      % (cfg.nx, cfg.nz, nt, shot.nt, t,
                         "; rate extrapolated to the full record" if nt < shot.nt else ""),
            "shots_per_hour": 3600.0 * rate / (sweeps * cfg.nx * cfg.nz * shot.nt)}


def single_shot_line(args, steps, warmup, workload="sh_layered"):
    """One JSON object for the plain single-shot RTM of ``workload``: one
    step = forward with a snapshot every isnap steps + backward with the
    image, each pass replayed as one HIP graph after the warmup."""
    import torch
    cfg, system = build_workload(workload)
    shot = RTMShot(cfg, args.isnap, args.kernel, args.tile_x, args.block_steps, args.tile_z)
    for _ in range(warmup):
        shot.run()
    graph = not args.no_graph and args.kernel in ("blocked", "wave", "auto", "tuned")
    if graph:  # each pass one HIP graph launch from here on
        shot.capture()
        shot.run()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        shot.run(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fw_ms = [e[0].elapsed_time(e[1]) for e in evs]
    bw_ms = [e[2].elapsed_time(e[3]) for e in evs]

    nt = cfg.nt
    cells = cfg.nx * cfg.nz
    ms_per_step = wall * 1000.0 / steps
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
    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "GCell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": b_bytes,
                     "steps_per_launch": k,
                     "avg_launch_us": round(bw_launch_us, 3),
                     "forward_launch_us": round(fw_launch_us, 3),
                     "forward_achieved": round(f_bytes / (fw_launch_us * 1e-6) / 1e9, 1)},
    }
    out["f64_rel_l2"] = float64_error(cfg, shot)
    out["_cfg"] = cfg
    del shot
    return out


def single_main(args):
    out = single_shot_line(args, args.steps, args.warmup, args.workload)
    cfg = out.pop("_cfg")
    copy_gbs = copy_bandwidth()
    out["roofline"]["copy_gbs"] = round(copy_gbs, 1)
    out["roofline"]["frac_of_copy"] = round(out["roofline"]["achieved"] / copy_gbs, 4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, target_s=12.0)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="salt",
                    choices=["salt", "psv_batch", "fwi", "sh_layered", "psv_anomaly"])
    ap.add_argument("--kernel", default="tuned")
    ap.add_argument("--shots", type=int, default=0,
                    help="jobs: number of shots (default: the full survey, 32 / 64 / 16)")
    ap.add_argument("--group", type=int, default=None,
                    help="psv_batch: shots per batch (0 = all in one batch)")
    ap.add_argument("--tile-x", type=int, default=0)
    ap.add_argument("--block-steps", type=int, default=0)
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--isnap", type=int, default=1)
    ap.add_argument("--budget-gb", type=float, default=None,
                    help="salt: device memory of the job (128); fwi: forward history kept (0: "
                         "sqrt(N) checkpoints); psv_batch: segment wavefields kept (200)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single-shot", action="store_true",
                    help="salt: skip the single-shot SH line")
    ap.add_argument("--no-traffic", action="store_true",
                    help="salt: skip the rocprofv3 counter passes of the dominant kernel")
    ap.add_argument("--no-graph", action="store_true",
                    help="single-shot RTM: issue the step launches eagerly instead of replaying "
                         "the captured forward and backward runs")
    ap.add_argument("--graph", action="store_true",
                    help="psv_batch: record every batched step range into a HIP graph (measured "
                         "1 %% slower than eager launches: each range is replayed once)")
    ap.add_argument("--probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--probe-shape", default="", help=argparse.SUPPRESS)
    ap.add_argument("--probe-steps", type=int, default=48, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(shots shard across GPUs with one image all-reduce)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    torch.cuda.set_device(0)
    if args.probe:
        return probe_main(args)
    if args.workload == "salt":
        return salt_main(args)
    if args.workload in JOB_BYTES:
        return job_main(args)
    return single_main(args)


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
