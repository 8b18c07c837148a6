"""Benchmark: forward FDTD throughput of the gfx950 kernels on one MI355X.

Workload (default): one single-shot production run -- a 400 x 400 x 200
FP32 grid (32M cells) holding a 100 mm x 35 mm ice disk of 3000 Voronoi
grains with a Watson girdle fabric (kappa = -3), a 2.25 MHz Ricker from a
6 mm piston on the top face, 20-cell CPML absorber (the configuration's
boundary; --boundary sponge runs the upstream sponge instead, and the
default run also times the sponge shot for comparison), 12000 steps.

A "step" is one FDTD time step over the whole grid (stress half-step,
source, velocity half-step, receiver record), i.e. nz*ny*nx cell-updates.
W untimed warm-up steps run first; the fields are then reset and K steps
are timed from n = 0 with inputs resident in HBM.  The timed run records a
hipEvent pair around every half-step kernel on the launch stream, which
gives each kernel's average duration and so its achieved HBM bandwidth
against algorithmic bytes (stress 68 B/cell, velocity 56 B/cell, step
124 B/cell).

Prints ONE JSON line.  Only one GPU is used: --gpus above 1 exits.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FDTD Gcell-updates/s (fwd, fwd+adj), %HBM peak; 72-azimuth sweep seconds"
HBM_PEAK_GBS = 8000.0
BYTES_STRESS = {"f32": 68, "f64": 132}     # v + s read, s written, label, damp
BYTES_VELOCITY = {"f32": 56, "f64": 108}   # s + v read, v written, label, damp
BYTES_STEP = {"f32": 124, "f64": 240}


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown CPU"


def cpu_baseline(cfg, nz_sample=20, warm=2, timed=20):
    """Upstream array-path loop (NumPy, float64, one thread) on a z-slab of
    the same specimen: cell-updates per second.  Builds its own wavelet of
    warm + timed samples (the workload's Ricker or Hann burst at its dt), so
    it does not depend on how many steps the GPU legs ran."""
    import numpy as np

    from oracle import numpy_forward as NF
    lab = cfg["lab"]
    nz = lab.shape[0]
    z0 = nz // 2 - nz_sample // 2
    sub = lab[z0:z0 + nz_sample]
    Ct, rho_t = cfg["mat_tables"]
    ny, nx = sub.shape[1:]
    c = (nz_sample // 2, ny // 2, nx // 2)
    idx = np.array([np.ravel_multi_index(c, sub.shape)])
    shot = NF.ArrayPathShot(sub + 1, Ct, rho_t, cfg["coeffs"], cfg["h"], cfg["dt"],
                            (idx, np.array([1.0])),
                            (np.array([0, 1]), idx, np.array([1.0])))
    nw = warm + timed
    wav = cfg["wavelet"]
    if len(wav) < nw:
        wav = cfg["wavelet_fn"](cfg["dt"], nw)
    for n in range(warm):
        shot.step(float(wav[n]))
    t0 = time.perf_counter()
    for n in range(warm, warm + timed):
        shot.step(float(wav[n]))
    dt = time.perf_counter() - t0
    cells = sub.size
    return dict(value=cells * timed / dt / 1e9, unit="Gcell-updates/s", cores=1, kind="port",
                sample="%dx%dx%d z-slab of the benchmark volume (%.1fM cells), %d timed steps "
                       "after %d warm-up, float64 NumPy restatement of the upstream array-path "
                       "loop, 1 thread, %s; %.2f s/step"
                       % (nz_sample, ny, nx, cells / 1e6, timed, warm, cpu_model(), dt / timed))


def forward_shot(cfg, boundary, precision, kernel="tiled"):
    """DeviceShot of a forward workload (specimen.config_c1 / config_c0) with
    the 20-cell CPML ("cpml", the configurations' boundary) or sponge."""
    from pulse_echo_azimuthal_simulation_amd import solver
    cpml_w = cfg["sponge_width"] if boundary == "cpml" else 0
    prep = solver.prepare_shot(cfg["lab"], cfg["axes"], cfg["h"], cfg["dt"], cfg["src_pts"],
                               cfg["rec_groups"], coeffs=cfg["coeffs"],
                               sponge_width=0 if boundary == "cpml" else cfg["sponge_width"],
                               damp_mask=cfg["damp_mask"], mat_tables=cfg["mat_tables"])
    cp = None
    if cpml_w:
        from pulse_echo_azimuthal_simulation_amd import cpml
        cp = cpml.Cpml(prep["mats"].shape, cpml_w, cfg["h"], cfg["dt"],
                       cpml.vmax_table(prep["Ct"], prep["rho_t"]),
                       cpml.dominant_frequency(cfg["wavelet"], cfg["dt"]))
    return solver.DeviceShot(prep["mats"], prep["Ct"], prep["rho_t"], prep["damp"],
                             prep["coeffs"], cfg["h"], cfg["dt"], prep["src_idx"],
                             prep["src_w"], prep["rec_ptr"], prep["rec_idx"], prep["rec_w"],
                             precision=precision, kernel=kernel, cpml=cp)


def bench_precision(args):
    """Error of the float32 traces against a float64 run of the same kernels:
    the forward workload (400x400x200 by default) run for --steps steps in
    both precisions with each boundary; rel-L2 over the whole trace and over
    the window after the first 10% of the peak (the echoes)."""
    import numpy as np
    import torch

    from pulse_echo_azimuthal_simulation_amd import specimen
    nt = max(args.steps, 1)
    cfg = specimen.config_c1(nt=nt) if args.config in (None, "c1") else specimen.config_c0(nt=nt)
    res = {}
    for boundary in ("cpml", "sponge"):
        tr = {}
        for prec in ("f32", "f64"):
            sh = forward_shot(cfg, boundary, prec, args.kernel)
            wav = sh.wavelet(cfg["wavelet"])
            t = sh.new_traces(nt)
            t0 = time.perf_counter()
            sh.run(wav, 0, nt, t)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            tr[prec] = t.cpu().numpy().astype(np.float64)[:, 0]
            log("%s %s: %d steps in %.1f s (%.2f Gcell-updates/s)"
                % (boundary, prec, nt, wall, int(np.prod(sh.shape)) * nt / wall / 1e9))
            del sh, t
            torch.cuda.empty_cache()
        a, b = tr["f32"], tr["f64"]
        peak = np.abs(b).max()
        late = np.argmax(np.abs(b) > 0.1 * peak)
        res[boundary] = {
            "rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
            "max_abs_err_over_peak": float(np.abs(a - b).max() / peak),
            "rel_l2_after_onset": float(np.linalg.norm(a[late:] - b[late:])
                                        / np.linalg.norm(b[late:])),
            "peak_abs": float(peak),
            "traces_finite": bool(np.isfinite(a).all() and np.isfinite(b).all()),
        }
        log("%s: f32 vs f64 rel-L2 %.3e" % (boundary, res[boundary]["rel_l2"]))
    nz, ny, nx = cfg["lab"].shape
    print(json.dumps({
        "metric": "f32 trace error against an f64 run of the same kernels (rel-L2)",
        "value": res["cpml"]["rel_l2"], "unit": "relative L2", "higher_is_better": False,
        "steps": nt, "dtype": "f32 vs f64", "data": "synthetic",
        "config": {"workload": "single-shot forward, %dx%dx%d grid, %d grains, pulse-echo "
                               "trace of the top piston" % (nx, ny, nz, len(cfg["axes"])),
                   "kernel": args.kernel},
        "boundaries": res}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--kernel", default="tiled", choices=["tiled", "direct"])
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--config", default=None, choices=["c1", "c0", "c3"],
                    help="workload (default c1 forward; c3 for --mode adjoint)")
    ap.add_argument("--mode", default="forward",
                    choices=["forward", "adjoint", "sweep", "inversion", "precision"],
                    help="forward: single-shot forward run; adjoint: misfit gradient w.r.t. "
                         "per-grain c-axis angles (forward + checkpointed reverse pass); sweep: "
                         "batched azimuth sweep of the 400x400x200 disk plus 8 smaller disks; "
                         "precision: f32 traces against an f64 run of the same kernels")
    ap.add_argument("--checkpoint-every", type=int, default=200)
    ap.add_argument("--hist", default="auto", choices=["auto", "strain", "velocity"],
                    help="adjoint mode: per-step history of the reverse pass")
    ap.add_argument("--adjoint-boundary", default="sponge", choices=["sponge", "cpml"],
                    help="adjoint mode: the workload's 20-cell sponge or a 20-cell CPML")
    ap.add_argument("--inv-full", action="store_true",
                    help="inversion mode: one misfit/gradient evaluation at production size "
                         "(36 azimuths of the 3000-grain 100x35 mm disk)")
    ap.add_argument("--azimuths", type=int, default=72)
    ap.add_argument("--plain-sample", type=int, default=2,
                    help="sweep mode: azimuths timed one at a time (host prep + "
                         "forward_fused_labels each) to estimate the one-at-a-time sweep")
    ap.add_argument("--history-gb", type=float, default=-1.0,
                    help="adjoint mode: HBM budget for the history kept from the first pass "
                         "(segments that fit are not recomputed); -1 = the free device memory "
                         "less 16 GiB, 0 = recompute every segment")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--boundary", default="cpml", choices=["cpml", "sponge"],
                    help="absorber of the forward run (cpml: 20 cells + couplant mask)")
    ap.add_argument("--sponge-steps", type=int, default=2000,
                    help="forward cpml runs: also time this many steps with the 20-cell "
                         "sponge (0 = skip)")
    ap.add_argument("--sweep-boundary", default="sponge", choices=["sponge", "cpml"],
                    help="--mode sweep: the upstream sweep's 20-cell sponge (default) or a "
                         "20-cell CPML")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--launch", default=None,
                    help="fixed launch shapes 'stress_variant,stress_zchunk,vel_variant,vel_zchunk' "
                         "(skips the autotuner)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="also run a plain device copy of this many GiB (PMC calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pulse_echo_azimuthal_simulation_amd import solver, specimen

    if args.mode == "adjoint":
        return bench_adjoint(args)
    if args.mode == "sweep":
        return bench_sweep(args)
    if args.mode == "inversion":
        return bench_inversion(args)
    if args.mode == "precision":
        return bench_precision(args)
    if args.config is None:
        args.config = "c1"
    t_setup = time.perf_counter()
    # the wavelet (and the configuration's record length) must cover every
    # leg that runs from n = 0: the warm-up, the timed steps and the sponge run
    nt = max(args.steps, args.warmup, args.sponge_steps if args.boundary == "cpml" else 0, 1)
    if args.config == "c1":
        cfg = specimen.config_c1(nt=nt)
        workload = ("single-shot forward, 400x400x200 FP32 grid (32M cells), 3000-grain Voronoi "
                    "ice disk 100x35 mm, Watson girdle kappa=-3, 2.25 MHz Ricker, 6 mm top "
                    "piston (pulse-echo), 20-cell %s + couplant damping mask, %d steps"
                    % ("CPML" if args.boundary == "cpml" else "sponge", nt))
    else:
        cfg = specimen.config_c0(nt=nt)
        workload = ("plain forward, 96x96x48 disk raster (+23-cell pads), 200 grains, Watson "
                    "kappa=5, 1 MHz 3-cycle Hann, 6 mm top piston, 20-cell %s, %d steps"
                    % ("CPML" if args.boundary == "cpml" else "sponge", nt))

    def make_shot(boundary):
        return forward_shot(cfg, boundary, args.precision, args.kernel)

    shot = make_shot(args.boundary)
    tuned = None
    if args.launch:
        from pulse_echo_azimuthal_simulation_amd import _abi
        sv, sz_, vv, vz_ = (int(x) for x in args.launch.split(","))
        shot.desc.tile[0], shot.desc.zchunk[0], shot.desc.tile[1], shot.desc.zchunk[1] = sv, sz_, vv, vz_
        tuned = {0: {"tile": _abi.tile_variants(args.precision, 0)[sv], "zchunk": sz_},
                 1: {"tile": _abi.tile_variants(args.precision, 1)[vv], "zchunk": vz_}}
    elif args.kernel == "tiled" and not args.no_tune:
        from pulse_echo_azimuthal_simulation_amd import _abi, autotune
        best, _ = autotune.tune(shot)
        tuned = {h: {"tile": _abi.tile_variants(args.precision, h)[best[h][0]],
                     "zchunk": best[h][1], "ms": round(best[h][2], 4)}
                 for h in (0, 1)}
        log("autotune: stress %s, velocity %s" % (tuned[0], tuned[1]))
    wav = shot.wavelet(cfg["wavelet"])
    traces = shot.new_traces(max(args.steps, args.warmup, 1))
    nz, ny, nx = shot.shape
    ncell = nz * ny * nx
    log("setup %.1f s: grid %s, %d grains, dt %.4g s, %d footprint cells"
        % (time.perf_counter() - t_setup, shot.shape, len(cfg["axes"]), cfg["dt"],
           len(cfg["src_pts"])))

    # warm-up, then reset the fields so the timed run is a whole shot from n = 0
    if args.warmup > 0:
        shot.run(wav, 0, args.warmup, traces)
    torch.cuda.synchronize()
    shot.reset()
    traces.zero_()
    torch.cuda.synchronize()
    log("warm-up %d steps done; timing %d steps" % (args.warmup, args.steps))

    t0 = time.perf_counter()
    ms_stress, ms_vel, ms_sr = shot.run_timed(wav, 0, args.steps, traces)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tr = traces.cpu().numpy()
    finite = bool(np.isfinite(tr).all())
    value = ncell * args.steps / wall / 1e9
    log("timed %.3f s: %.2f Gcell-updates/s; stress %.3f ms, velocity %.3f ms, source/record "
        "%.3f ms per step; traces finite=%s, peak |p| %.3e"
        % (wall, value, ms_stress / args.steps, ms_vel / args.steps, ms_sr / args.steps,
           finite, float(np.abs(tr[:args.steps]).max())))

    p = args.precision
    t_s = ms_stress / args.steps / 1e3
    t_v = ms_vel / args.steps / 1e3
    bw_s = BYTES_STRESS[p] * ncell / t_s / 1e9
    bw_v = BYTES_VELOCITY[p] * ncell / t_v / 1e9
    dominant = "stress" if t_s >= t_v else "velocity"
    achieved = bw_s if dominant == "stress" else bw_v
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s_%s_%s.json"
                         % (args.config, p, args.kernel, args.boundary))
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get(dominant)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": p,
        "data": "synthetic",
        "config": {"workload": workload, "grid_zyx": [nz, ny, nx], "cells": ncell,
                   "grains": int(len(cfg["axes"])),
                   "boundary": ("cpml20" if args.boundary == "cpml" else "sponge20") + "+fluid-mask",
                   "kernel": args.kernel, "mode": "forward",
                   "launch": None if tuned is None else
                   {"stress": tuned[0], "velocity": tuned[1]}},
        "hbm_frac_step": round(value * BYTES_STEP[p] / HBM_PEAK_GBS, 4),
        "kernels_ms_per_step": {"stress": round(t_s * 1e3, 5), "velocity": round(t_v * 1e3, 5),
                                "source_record": round(ms_sr / args.steps, 5)},
        "kernels_gbs": {"stress": round(bw_s, 1), "velocity": round(bw_v, 1)},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "traces_finite": finite,
    }
    if args.boundary == "cpml" and args.sponge_steps > 0:
        # the same workload with the upstream absorber (20-cell sponge)
        del shot, traces
        torch.cuda.empty_cache()
        sp = make_shot("sponge")
        tuned2 = None
        if args.kernel == "tiled" and not args.no_tune:
            from pulse_echo_azimuthal_simulation_amd import _abi, autotune
            tuned2, _ = autotune.tune(sp)
        tr2 = sp.new_traces(args.sponge_steps)
        sp.run(wav, 0, min(args.warmup, args.sponge_steps), tr2)
        torch.cuda.synchronize()
        sp.reset()
        t1 = time.perf_counter()
        ms2 = sp.run_timed(wav, 0, args.sponge_steps, tr2)
        torch.cuda.synchronize()
        w2 = time.perf_counter() - t1
        ts2, tv2 = ms2[0] / args.sponge_steps / 1e3, ms2[1] / args.sponge_steps / 1e3
        dom2 = "stress" if ts2 >= tv2 else "velocity"
        ach2 = (BYTES_STRESS[p] * ncell / ts2 if dom2 == "stress"
                else BYTES_VELOCITY[p] * ncell / tv2) / 1e9
        traffic2 = None
        tfile2 = os.path.join(ROOT, "profiles", "traffic_%s_%s_%s_sponge.json"
                              % (args.config, p, args.kernel))
        if os.path.exists(tfile2):
            with open(tfile2) as fh:
                traffic2 = json.load(fh).get(dom2)
        out["sponge20"] = {"value": round(ncell * args.sponge_steps / w2 / 1e9, 4),
                           "steps": args.sponge_steps,
                           "launch": None if tuned2 is None else
                           {h: {"tile": list(_abi.tile_variants(p, i)[tuned2[i][0]]),
                                "zchunk": tuned2[i][1]}
                            for i, h in ((0, "stress"), (1, "velocity"))},
                           "kernels_ms_per_step": {"stress": round(ts2 * 1e3, 5),
                                                   "velocity": round(tv2 * 1e3, 5)},
                           "roofline": {"kernel": dom2, "bound": "hbm", "achieved": round(ach2, 1),
                                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                        "frac": round(ach2 / HBM_PEAK_GBS, 4),
                                        "traffic": traffic2}}
        log("sponge20: %.2f Gcell-updates/s over %d steps" % (out["sponge20"]["value"],
                                                               args.sponge_steps))
        del sp, tr2
    if args.calibrate_copy > 0:
        n = int(args.calibrate_copy * 2 ** 30 / 4)
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = n * 4
    if not args.no_cpu_baseline:
        log("CPU baseline (NumPy array path, 1 thread) ...")
        out["cpu_baseline"] = cpu_baseline(cfg)
        log("CPU baseline %.4f Gcell-updates/s" % out["cpu_baseline"]["value"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_sweep(args):
    """Batched azimuth sweep against one-at-a-time solves.

    Workload: the 100 mm x 35 mm, 3000-grain, girdle-fabric disk of the
    single-shot benchmark (h = 0.285 mm) rotated under a +x rim transducer at
    N azimuths (360/N deg apart), each a (169, 397, 397) grid, plus 8 smaller
    seeded disks whose grids run from 128x128x64 to 256x256x128; every shot
    runs --steps steps (2.25 MHz Ricker, couplant mask, and the 20-cell sponge
    or with --sweep-boundary cpml a 20-cell CPML).
    Batched: every shot prepared on the GPU (re-grid, distance transform,
    dt) and all shots advanced together in one multi-shot launch per
    half-step, replayed as a hipGraph.  One at a time: the upstream sweep
    pattern, host preparation (rotated_grid, tables, dt, EDT mask,
    footprint) then forward_fused_labels per azimuth; --plain-sample azimuths
    and two of the small disks are timed and the sweep total extrapolated.
    Traces of the timed one-at-a-time shots are compared with the batch."""
    import numpy as np
    import torch

    from pulse_echo_azimuthal_simulation_amd import autotune, host, solver, specimen, sweep

    nt = max(args.steps, 1)
    f0 = 2.25e6
    h = specimen.C_REF / f0 / 6
    cw = 20 if args.sweep_boundary == "cpml" else 0
    co = host.optimised_coeffs(8, kh_max=2.0, multistart=True)
    t0 = time.perf_counter()
    big = sweep.disk_build(100e-3, 35e-3, h, 3000, seed=11, kappa=-3.0,
                           fabric_axis=np.random.default_rng(11).normal(size=3))
    smalls = []
    for k in range(8):
        n = 128 + 128 * k // 7                       # grid side 128 .. 256
        nd = n - 46
        nzd = max(n // 2 - 46, 12)
        smalls.append(sweep.disk_build(nd * h, nzd * h, h, 60 + 40 * k, seed=100 + k,
                                       kappa=5.0, fabric_axis=(np.cos(0.5), np.sin(0.5), 0.0)))
    log("specimens built in %.1f s" % (time.perf_counter() - t0))
    azs = [360.0 * k / args.azimuths for k in range(args.azimuths)]

    # ---------------- batched
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prep_big = sweep.GpuAzimuthPrep(big, h, co, f0, sponge=20, cpml_width=cw)
    tables = sweep.rotated_tables(big["axes"], azs)
    shots, wavs, nts = [], [], []
    for az, tab in zip(azs, tables):
        sh, wav, n_, _ = prep_big.shot(az, nt=nt, tables=tab)
        shots.append(sh)
        wavs.append(wav)
        nts.append(n_)
    for b in smalls:
        sh, wav, n_, _ = sweep.GpuAzimuthPrep(b, h, co, f0, sponge=20,
                                              cpml_width=cw).shot(0.0, nt=nt)
        shots.append(sh)
        wavs.append(wav)
        nts.append(n_)
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    t1 = time.perf_counter()
    best, _ = autotune.tune(shots[0], reps=3)
    for sh in shots:
        for hh in (0, 1):
            sh.desc.tile[hh] = best[hh][0]
            sh.desc.zchunk[hh] = best[hh][1] if sh.shape == shots[0].shape else 0
        for f in sh.v + sh.s:
            f.zero_()
    batch = sweep.ShotBatch(shots, wavs, nts)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t1
    t1 = time.perf_counter()
    t_prog = [time.perf_counter()]

    def progress(done, total):
        if time.perf_counter() - t_prog[0] > 20 or done >= total:
            t_prog[0] = time.perf_counter()
            log("batched run: %d / %d steps" % (done, total))

    traces = batch.run(graph_steps=50, progress=progress, progress_every=250)
    torch.cuda.synchronize()
    t_run = time.perf_counter() - t1
    t_batched = time.perf_counter() - t0
    cells = sum(int(np.prod(sh.shape)) * n_ for sh, n_ in zip(shots, nts))
    log("batched: %d shots, prep %.2f s, tune %.2f s, run %.2f s (%.2f Gcell-updates/s), "
        "total %.2f s" % (len(shots), t_prep, t_tune, t_run, cells / t_run / 1e9, t_batched))
    tr_batch = [t.cpu().numpy() for t in traces]
    del batch, shots, wavs, traces
    torch.cuda.empty_cache()

    # ---------------- one at a time (sample, extrapolated)
    plain_boundary = (dict(sponge_width=0, boundary="cpml", cpml_width=cw, cpml_f0=f0) if cw
                      else dict(sponge_width=20))
    plain_az, plain_small = [], []
    match = True
    for k in range(min(args.plain_sample, len(azs))):
        az = azs[(k * len(azs)) // max(args.plain_sample, 1)]
        torch.cuda.synchronize()
        ts = time.perf_counter()
        a = sweep.azimuth_shot(big, h, az, co, f0, sponge=20, nt=nt)
        tr = solver.forward_fused_labels(a["lab"], a["axes"], h, a["dt"], nt, a["src_pts"],
                                         a["wavelet"], a["rec_groups"], coeffs=co,
                                         damp_mask=a["damp_mask"], mat_tables=a["mat_tables"],
                                         **plain_boundary)
        plain_az.append(time.perf_counter() - ts)
        ref = tr_batch[azs.index(az)]
        match &= bool(np.allclose(tr, ref, rtol=0, atol=1e-6 * np.abs(ref).max()))
    for k in (0, 7):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        a = sweep.azimuth_shot(smalls[k], h, 0.0, co, f0, sponge=20, nt=nt)
        solver.forward_fused_labels(a["lab"], a["axes"], h, a["dt"], nt, a["src_pts"],
                                    a["wavelet"], a["rec_groups"], coeffs=co,
                                    damp_mask=a["damp_mask"], mat_tables=a["mat_tables"],
                                    **plain_boundary)
        plain_small.append(time.perf_counter() - ts)
    t_plain = float(np.mean(plain_az)) * len(azs) + float(np.mean(plain_small)) * len(smalls)
    log("one at a time: %.2f s per azimuth, %.2f s per small disk (sampled) -> %.1f s; "
        "traces match batch: %s" % (np.mean(plain_az), np.mean(plain_small), t_plain, match))
    out = {
        "metric": METRIC,
        "value": round(t_batched, 3),
        "unit": "s (whole sweep: %d azimuths + 8 disks, %d steps each)" % (len(azs), nt),
        "n_gpus": 1, "steps": nt, "warmup": 0,
        "ms_per_step": round(t_run / nt * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched azimuth sweep: 100x35 mm 3000-grain disk at %d azimuths "
                               "(169x397x397 each) + 8 disks 128x128x64..256x256x128, 20-cell %s"
                               % (len(azs), "CPML" if cw else "sponge"),
                   "shots": len(azs) + len(smalls), "mode": "sweep",
                   "launch": {"stress": best[0][:2], "velocity": best[1][:2]}},
        "batched_s": {"prep": round(t_prep, 3), "tune": round(t_tune, 3),
                      "run": round(t_run, 3), "total": round(t_batched, 3)},
        "batched_gcells": round(cells / t_run / 1e9, 3),
        "one_at_a_time_s_estimate": round(t_plain, 2),
        "one_at_a_time_sample_s": {"azimuth": [round(x, 3) for x in plain_az],
                                   "small_disk": [round(x, 3) for x in plain_small]},
        "speedup_vs_one_at_a_time": round(t_plain / t_batched, 3),
        "traces_match": match,
    }
    print(json.dumps(out), flush=True)


def bench_inversion(args):
    """Fabric inversion over an azimuthal sweep (reduced size of the
    production inversion: a 9.6 mm x 4.8 mm disk of 200 grains at 0.1 mm,
    12 azimuths 30 deg apart, 3 mm rim element, 2 MHz Ricker, the default
    600 steps each).
    Synthetic data from the Watson fabric alpha = 30 deg, kappa = 5; start
    alpha = 70 deg, kappa = 1.5; L-BFGS-B on the float64 misfit with the
    adjoint gradient (fabric_fit.py), at most 30 iterations.  value =
    forward-equivalent cell-updates per second over the whole fit."""
    import math

    import numpy as np
    import torch

    from pulse_echo_azimuthal_simulation_amd import fabric_fit as F, host, sweep

    if args.inv_full:
        return bench_inversion_full(args)
    nt = args.steps if args.steps != 12000 else 600
    h = 1e-4
    G = 200
    build = sweep.disk_build(9.6e-3, 4.8e-3, h, G, seed=21)
    lat = F.latents(G, 22)
    co = host.optimised_coeffs(8, kh_max=2.0, multistart=True)
    azs = [30.0 * k for k in range(12)]
    t0 = time.perf_counter()
    inv = F.FabricInversion(build, azs, h, co, 2e6, nt=nt, lat=lat, sponge=10,
                            checkpoint_every=100, precision=args.precision, element_d_mm=3.0)
    truth = (math.radians(30.0), 5.0)
    data = [t.clone() for t in inv.traces(truth)]
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    p0 = (math.radians(70.0), 1.5)
    t1 = time.perf_counter()
    p, J, hist, res = inv.fit(p0, data, max_iter=30, log=log)
    t_fit = time.perf_counter() - t1
    cells = int(np.prod(inv.shots[0].shape))
    evals = len(hist)
    value = evals * len(azs) * cells * nt / t_fit / 1e9
    a_fit = math.degrees(p[0]) % 180.0
    out = {
        "metric": METRIC, "value": round(value, 4),
        "unit": "Gcell-updates/s (forward-equivalent, whole fit)",
        "n_gpus": 1, "steps": nt, "warmup": 0,
        "ms_per_step": round(t_fit / (evals * len(azs) * nt) * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": "fabric inversion: 9.6x4.8 mm disk, %d grains, grid %s, %d azimuths, "
                               "%d steps, Watson truth alpha 30 deg kappa 5, start 70 deg / 1.5"
                               % (G, list(inv.shots[0].shape), len(azs), nt),
                   "mode": "inversion"},
        "truth": {"alpha_deg": 30.0, "kappa": 5.0, "a1": round(F.watson_eigenvalue(5.0), 4)},
        "recovered": {"alpha_deg": round(a_fit, 3), "kappa": round(float(p[1]), 4),
                      "a1": round(F.watson_eigenvalue(float(p[1])), 4)},
        "misfit": {"start": hist[0]["J"], "final": J}, "evaluations": evals,
        "iterations": int(res.nit), "seconds": {"setup": round(t_setup, 2),
                                                 "fit": round(t_fit, 2)},
        "history": [[round(math.degrees(e["alpha"]), 3), round(e["kappa"], 4), e["J"]]
                    for e in hist],
    }
    print(json.dumps(out), flush=True)


def bench_inversion_full(args):
    """One evaluation (misfit and its gradient in the fabric parameters) of
    the fabric inversion at the production size: the 100 mm x 35 mm disk of
    3000 grains at 0.285 mm (ppw 6 at 2.25 MHz) under a rim transducer at 36
    azimuths 10 deg apart, each a 169 x 397 x 397 grid with the 20-cell
    sponge, --steps steps (default 600; the configuration's 12000 are
    extrapolated), forward + checkpointed adjoint per azimuth, float64
    misfit.  The azimuths run one after another, each filling the device."""
    import math

    import numpy as np
    import torch

    from pulse_echo_azimuthal_simulation_amd import fabric_fit as F, host, specimen, sweep

    nt = args.steps if args.steps != 12000 else 600
    f0 = 2.25e6
    h = specimen.C_REF / f0 / 6
    G = 3000
    t0 = time.perf_counter()
    build = sweep.disk_build(100e-3, 35e-3, h, G, seed=11)
    lat = F.latents(G, 22)
    co = host.optimised_coeffs(8, kh_max=2.0, multistart=True)
    azs = [10.0 * k for k in range(36)]
    inv = F.FabricInversion(build, azs, h, co, f0, nt=nt, lat=lat, sponge=20,
                            checkpoint_every=args.checkpoint_every, precision=args.precision,
                            element_d_mm=6.0)
    truth = (math.radians(30.0), -3.0)
    data = [t.clone() for t in inv.traces(truth)]
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log("setup %.1f s: %d azimuths of %s, %d steps" % (t_setup, len(azs),
                                                       list(inv.shots[0].shape), nt))
    p0 = (math.radians(40.0), -2.0)
    # two evaluations: every checkpoint segment recomputed (the
    # configuration's checkpoint/recompute scheme, which a 12000-step shot
    # needs: its history would not fit), then with the first-pass history
    # kept in the free HBM where it fits
    t1 = time.perf_counter()
    inv.history_gb = 0.0
    J, g = inv.misfit_gradient(p0, data, log=log)
    torch.cuda.synchronize()
    t_eval = time.perf_counter() - t1
    cells = int(np.prod(inv.shots[0].shape))
    value = len(azs) * cells * nt / t_eval / 1e9
    log("evaluation (recompute) %.1f s: J %.4e, dJ/d(alpha, kappa) %s" % (t_eval, J, g))
    t1 = time.perf_counter()
    inv.history_gb = None
    J2, g2 = inv.misfit_gradient(p0, data, log=log)
    torch.cuda.synchronize()
    t_eval2 = time.perf_counter() - t1
    log("evaluation (history in HBM) %.1f s: J %.4e, dJ/d(alpha, kappa) %s" % (t_eval2, J2, g2))
    out = {
        "metric": METRIC, "value": round(value, 4),
        "unit": "Gcell-updates/s (forward-equivalent, fwd+adj over all azimuths)",
        "n_gpus": 1, "steps": nt, "warmup": 0,
        "ms_per_step": round(t_eval / (len(azs) * nt) * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": "fabric inversion, one evaluation at production size: 100x35 mm "
                               "disk, %d grains, %d azimuths of %s cells, %d steps, 20-cell "
                               "sponge, checkpoints every %d" % (G, len(azs),
                                                                 list(inv.shots[0].shape), nt,
                                                                 args.checkpoint_every),
                   "mode": "inversion-full"},
        "seconds": {"setup": round(t_setup, 2), "evaluation": round(t_eval, 2),
                    "evaluation_history_in_hbm": round(t_eval2, 2),
                    "evaluation_at_12000_steps_extrapolated": round(t_eval * 12000 / nt, 1),
                    "fit_30_iterations_at_12000_steps_extrapolated_h":
                        round(30 * t_eval * 12000 / nt / 3600, 2)},
        "misfit": J, "gradient": [float(g[0]), float(g[1])],
        "gradient_history_in_hbm": [float(g2[0]), float(g2[1])],
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Misfit gradient w.r.t. per-grain (theta, phi) on the adjoint workload:
    forward with checkpoints, then per segment a recompute storing v and the
    reverse kernels, reducing dJ/dCt projected on dC/dtheta, dC/dphi per grain.
    value = forward cell-updates per second of the whole gradient
    (nz*ny*nx*nt / wall); cost ratio = gradient wall / forward wall."""
    import numpy as np
    import torch

    from pulse_echo_azimuthal_simulation_amd import adjoint, autotune, fabric, solver, specimen

    cfgname = args.config or "c3"
    nt = max(args.steps if args.steps != 12000 else 4000, 1)     # the workload's 4000 steps
    t_setup = time.perf_counter()
    if cfgname != "c3":
        raise SystemExit("--mode adjoint runs the c3 workload")
    cfg = specimen.config_c3(nt=nt)
    theta, phi = fabric.angles_from_axes(cfg["axes"])
    Ct, rho = fabric.stiffness_table(theta, phi)
    cp = args.adjoint_boundary == "cpml"
    prep = solver.prepare_shot(cfg["lab"], cfg["axes"], cfg["h"], cfg["dt"], cfg["src_pts"],
                               cfg["rec_groups"], coeffs=cfg["coeffs"],
                               sponge_width=0 if cp else cfg["sponge_width"],
                               damp_mask=cfg["damp_mask"],
                               mat_tables=(Ct.astype(np.float32), rho.astype(np.float32)))
    pml = None
    if cp:
        from pulse_echo_azimuthal_simulation_amd import cpml
        pml = cpml.Cpml(prep["mats"].shape, cfg["sponge_width"], cfg["h"], cfg["dt"],
                        cpml.vmax_table(prep["Ct"], prep["rho_t"]),
                        cpml.dominant_frequency(cfg["wavelet"], cfg["dt"]))

    def mk(table, kernel="tiled"):
        return solver.DeviceShot(prep["mats"], table, prep["rho_t"], prep["damp"],
                                 prep["coeffs"], cfg["h"], cfg["dt"], prep["src_idx"],
                                 prep["src_w"], prep["rec_ptr"], prep["rec_idx"], prep["rec_w"],
                                 precision=args.precision, kernel=kernel, cpml=pml)

    tt, pt = fabric.angles_from_axes(cfg["axes_true"])
    target = mk(fabric.stiffness_table(tt, pt)[0])
    best, _ = autotune.tune(target)
    for h in (0, 1):
        target.desc.tile[h], target.desc.zchunk[h] = best[h][0], best[h][1]
    wav = target.wavelet(cfg["wavelet"])
    d = target.new_traces(nt)
    target.run(wav, 0, nt, d)
    shot = mk(Ct)
    for h in (0, 1):
        shot.desc.tile[h], shot.desc.zchunk[h] = best[h][0], best[h][1]
    del target
    fwd_shot = shot
    proj = fabric.angle_projections(theta, phi)
    hist = args.hist if args.hist != "auto" else adjoint.choose_history(shot,
                                                                        args.checkpoint_every)
    adj = adjoint.DeviceAdjoint(shot, n_proj=2, proj=proj, hist=hist)
    nz, ny, nx = shot.shape
    ncell = nz * ny * nx
    log("adjoint setup %.1f s: grid %s, %d grains, %d steps, dt %.4g"
        % (time.perf_counter() - t_setup, shot.shape, len(cfg["axes"]), nt, cfg["dt"]))

    def resid(tr):
        r = tr - d[:tr.shape[0]]
        return 0.5 * float((r.double() ** 2).sum()), r

    # warm-up (compiles nothing, but pages in every kernel once)
    adjoint.misfit_gradient(shot, wav, min(nt, 2 * args.checkpoint_every), resid,
                            checkpoint_every=args.checkpoint_every, adjoint=adj)
    torch.cuda.synchronize()
    tr = fwd_shot.new_traces(nt)
    fwd_shot.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd_shot.run(wav, 0, nt, tr)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    if args.history_gb < 0:
        free, _ = torch.cuda.mem_get_info()
        hist_bytes = max(0, int(free) - 16 * 2 ** 30)
    else:
        hist_bytes = int(args.history_gb * 2 ** 30)
    t0 = time.perf_counter()
    J, g, _ = adjoint.misfit_gradient(shot, wav, nt, resid, checkpoint_every=args.checkpoint_every,
                                      history_bytes=hist_bytes, adjoint=adj)
    gnp = g.cpu().numpy()
    t_grad = time.perf_counter() - t0
    log("forward %.3f s, gradient %.3f s (ratio %.2f), J %.4e, |g| max %.3e"
        % (t_fwd, t_grad, t_grad / t_fwd, J, np.abs(gnp).max()))
    out = {
        "metric": METRIC,
        "value": round(ncell * nt / t_grad / 1e9, 4),
        "unit": "Gcell-updates/s (fwd+adj, forward-equivalent)",
        "n_gpus": 1, "steps": nt, "warmup": 0,
        "ms_per_step": round(t_grad / nt * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": "misfit gradient w.r.t. 500 per-grain (theta, phi), 80x160x160 "
                               "grid, 1 MHz Hann, target traces from axes tilted N(0, 5 deg), "
                               "20-cell %s, checkpoints every %d steps, %s history, first-pass "
                               "history budget %.1f GiB"
                               % ("CPML" if cp else "sponge",
                                  args.checkpoint_every, hist, hist_bytes / 2 ** 30),
                   "grid_zyx": [nz, ny, nx], "cells": ncell, "mode": "fwd+adj", "history": hist},
        "forward_s": round(t_fwd, 4), "gradient_s": round(t_grad, 4),
        "cost_ratio_vs_forward": round(t_grad / t_fwd, 3),
        "forward_gcells": round(ncell * nt / t_fwd / 1e9, 3),
        "misfit": J,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
