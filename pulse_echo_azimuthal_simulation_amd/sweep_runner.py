"""Resumable azimuth-sweep runner over the batched GPU engine.

Same filesystem contract as the upstream sweep engine
(sim/pipeline/sweep_runner.py:1-24): a sweep directory holds
``config.json`` (written once), one ``az{ddd}.npz`` per completed azimuth
(written atomically: tmp file + os.replace with retries), a merged
``status.json`` heartbeat, and a cooperative ``STOP`` flag.  Continuing a
sweep skips the azimuths whose files exist.  Per-azimuth records carry the
upstream keys (trace f32, dt, f0, az_deg, E1, t1_s, coda_db, tail, h, order,
ppw, fluid_damp, seed; sweep_runner.py:373-381) and the observables of
``analyse`` (sweep_runner.py:209-225).

The difference is the execution: pending azimuths are prepared on the GPU
(``sweep.GpuAzimuthPrep``) and advanced together in multi-shot batches
(``sweep.ShotBatch``) sized to a device-memory budget; the STOP flag is
honoured between graph replays (the batch in flight is discarded; saved
files are final).  CLI: ``python -m pulse_echo_azimuthal_simulation_amd.sweep_runner
--root DIR --new NAME --set key=json ...`` then ``--run NAME``.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time

import numpy as np

C_REF = 3850.0

# sim/pipeline/sweep_runner.py:74-107
DEFAULTS = dict(
    f0_mhz=5.0, ppw=6.0, order=8, az_start=0, az_stop=360, az_step=1,
    sweep_order="progressive", record_factor=2.7, fluid_damp=0.02, sponge=10,
    element_d_mm=6.35, diameter_mm=100.0, thickness_mm=35.0, n_grains=100, size_cv=0.35,
    concentration=3.93, spatial_corr=0.0, fabric_axis=[1.0, 0.0, 0.0], seed=7,
    axes_convention="rigid2", fd_kh_max=2.0, rasterise="single",
)


class StopRequested(Exception):
    pass


def atomic_replace(tmp, dst, critical, retries=25, wait=0.2):
    """os.replace with retries on PermissionError (sweep_runner.py:52-71):
    data files raise after the retries, status files give up silently."""
    for _ in range(retries):
        try:
            os.replace(tmp, dst)
            return True
        except PermissionError:
            time.sleep(wait)
    if critical:
        raise PermissionError("could not replace %s after retries" % dst)
    try:
        os.remove(tmp)
    except OSError:
        pass
    return False


def sweep_dir(root, name):
    return os.path.join(root, name)


def create(root, name, **overrides):
    cfg = dict(DEFAULTS)
    cfg.update(overrides)
    cfg["name"] = name
    d = sweep_dir(root, name)
    os.makedirs(d, exist_ok=True)
    p = os.path.join(d, "config.json")
    if not os.path.exists(p):
        with open(p, "w", encoding="utf-8") as f:
            json.dump(cfg, f, indent=1)
    return p


def load(root, name_or_path):
    p = (name_or_path if name_or_path.endswith("config.json")
         else os.path.join(sweep_dir(root, name_or_path), "config.json"))
    with open(p, encoding="utf-8") as f:
        return json.load(f)


def bit_reversed_order(n):
    """Van der Corput visiting order of n items (sweep_runner.py:133-145)."""
    if n <= 1:
        return list(range(n))
    bits = (n - 1).bit_length()
    out = []
    for i in range(2 ** bits):
        r = int(format(i, "0%db" % bits)[::-1], 2)
        if r < n:
            out.append(r)
    return out


def azimuths(cfg):
    """Visiting order (sweep_runner.py:148-163): 'progressive' covers the
    first half-circle bit-reversed, then the second; else sequential."""
    base = list(range(int(cfg["az_start"]), int(cfg["az_stop"]), int(cfg["az_step"])))
    if cfg.get("sweep_order", "progressive") != "progressive":
        return base
    mid = cfg["az_start"] + 180
    first = [a for a in base if a < mid]
    second = [a for a in base if a >= mid]
    return ([first[i] for i in bit_reversed_order(len(first))]
            + [second[i] for i in bit_reversed_order(len(second))])


def done_azimuths(root, cfg):
    d = sweep_dir(root, cfg["name"])
    return sorted(int(f[2:5]) for f in os.listdir(d)
                  if f.startswith("az") and f.endswith(".npz") and ".tmp" not in f)


def pending(root, cfg):
    done = set(done_azimuths(root, cfg))
    return [a for a in azimuths(cfg) if a not in done]


def status(root, cfg, **kw):
    """Merge-update status.json atomically (sweep_runner.py:180-206)."""
    p = os.path.join(sweep_dir(root, cfg["name"]), "status.json")
    s = {}
    if os.path.exists(p):
        try:
            with open(p, encoding="utf-8") as f:
                s = json.load(f)
        except Exception:
            s = {}
    s.update(pid=os.getpid(), heartbeat=time.time(), done=len(done_azimuths(root, cfg)),
             total=len(azimuths(cfg)))
    s.update({k: v for k, v in kw.items() if v is not None})
    tmp = p + ".tmp"
    with open(tmp, "w", encoding="utf-8") as f:
        json.dump(s, f)
    atomic_replace(tmp, p, critical=False)
    return s


def analyse(tr, dt, diameter_m=0.100):
    """Trace observables (sweep_runner.py:209-225): first-echo envelope peak
    E1 and its time t1 (searched +-2 us about the back-wall round trip
    2 D / c), coda level in dB (RMS envelope over 24-36 us relative to
    E1) and the stability tail (max |trace| 3 us after E1, over E1)."""
    from scipy.signal import hilbert
    tr = np.asarray(tr, float).ravel()
    e = np.abs(hilbert(tr))
    fs = 1.0 / dt
    k0 = int(2 * diameter_m / C_REF * fs)
    a, b = max(k0 - int(2e-6 * fs), 0), min(k0 + int(2e-6 * fs), len(e))
    k = a + int(np.argmax(e[a:b])) if b > a else int(np.argmax(e))
    E1, t1 = float(e[k]), k * dt
    lo, hi = int(24e-6 * fs), min(int(36e-6 * fs), len(e))
    coda = (20 * np.log10(np.sqrt((e[lo:hi] ** 2).mean()) / E1 + 1e-30)
            if hi > lo else float("nan"))
    k_tail = int((t1 + 3e-6) * fs)
    tail = (float(np.abs(tr[k_tail:]).max() / max(E1, 1e-30))
            if k_tail < len(tr) - 8 else float("nan"))
    return E1, t1, float(coda), tail


def save_azimuth(root, cfg, az, trace, dt, h):
    """One completed azimuth, atomically (sweep_runner.py:373-381)."""
    d = sweep_dir(root, cfg["name"])
    E1, t1, coda, tail = analyse(trace, dt, cfg["diameter_mm"] * 1e-3)
    tmp = os.path.join(d, "az%03d.tmp.npz" % az)
    np.savez(tmp, trace=np.asarray(trace, np.float32).ravel(), dt=dt, f0=cfg["f0_mhz"] * 1e6,
             az_deg=az, E1=E1, t1_s=t1, coda_db=coda, tail=tail, h=h, order=cfg["order"],
             ppw=cfg["ppw"], fluid_damp=cfg["fluid_damp"], seed=cfg["seed"])
    atomic_replace(tmp, os.path.join(d, "az%03d.npz" % az), critical=True)
    return dict(E1=E1, t1=t1, coda=coda, tail=tail)


def build_specimen(cfg, h):
    """The disk of the config at spacing h (sweep_runner.py:290-303): the
    Laguerre-Voronoi DiskSpecimen with the configured grain count, size CV,
    Watson concentration about fabric_axis, spatial correlation and seed."""
    from .disk_specimen import DiskSpecimen
    sp = DiskSpecimen(diameter_m=cfg["diameter_mm"] * 1e-3, thickness_m=cfg["thickness_mm"] * 1e-3,
                      n_grains=int(cfg["n_grains"]), size_cv=float(cfg["size_cv"]),
                      concentration=float(cfg["concentration"]),
                      spatial_corr=float(cfg["spatial_corr"]),
                      fabric_axis=tuple(cfg["fabric_axis"]), seed=int(cfg["seed"]))
    return sp.build(h, adjacency=False)


def run(root, cfg, build=None, mem_budget_gb=None, graph_steps=50, log=print, should_stop=None):
    """Run every pending azimuth of the sweep on the GPU, in batches.

    ``should_stop()`` (default: the STOP file exists) is polled between graph
    replays; returns "complete", "stopped", "legacy" or "no-gpu"."""
    import torch

    from . import _abi, host, sweep
    d = sweep_dir(root, cfg["name"])
    stop_flag = os.path.join(d, "STOP")
    if os.path.exists(stop_flag):
        os.remove(stop_flag)                                     # starting clears old stops
    if cfg.get("axes_convention") != "rigid2":
        if done_azimuths(root, cfg):
            status(root, cfg, phase="LEGACY sweep - cannot continue")
            log("[sweep] config.json lacks axes_convention='rigid2' and has completed "
                "azimuths: start a new sweep instead")
            return "legacy"
        cfg["axes_convention"] = "rigid2"
        p = os.path.join(d, "config.json")
        with open(p + ".tmp", "w", encoding="utf-8") as f:
            json.dump(cfg, f, indent=1)
        atomic_replace(p + ".tmp", p, critical=True)
    if not torch.cuda.is_available():
        status(root, cfg, phase="NO GPU - cannot run")
        log("[sweep] no HIP device: the solver runs on the GPU only")
        return "no-gpu"
    f0 = cfg["f0_mhz"] * 1e6
    h = C_REF / f0 / cfg["ppw"]
    khm = cfg.get("fd_kh_max")
    co = host.optimised_coeffs(cfg["order"], kh_max=1.6 if khm is None else float(khm),
                               multistart=khm is not None)
    if build is None:
        status(root, cfg, phase="building specimen")
        build = build_specimen(cfg, h)
    prep = sweep.GpuAzimuthPrep(build, h, co, f0, sponge=cfg["sponge"],
                                fluid_damp=cfg["fluid_damp"], element_d_mm=cfg["element_d_mm"],
                                record_factor=cfg["record_factor"],
                                single_raster=cfg.get("rasterise") == "single")
    todo = pending(root, cfg)
    log("[sweep] %d of %d azimuths pending" % (len(todo), len(azimuths(cfg))))
    if not todo:
        status(root, cfg, phase="complete")
        return "complete"
    # shots per batch from the memory budget (state, labels, damp ~ 44 B/cell)
    nd, ndz, m, _, _ = prep.geometry(0.0)
    cells = (ndz + 2 * m) * (nd + 2 * m) ** 2
    free = torch.cuda.mem_get_info()[0]
    budget = (mem_budget_gb * 2 ** 30) if mem_budget_gb else 0.7 * free
    per_batch = max(1, int(budget // (cells * 48)))
    tables = sweep.rotated_tables(build["axes"], todo)
    times = []
    for b0 in range(0, len(todo), per_batch):
        group = todo[b0:b0 + per_batch]
        t0 = time.time()
        status(root, cfg, phase="az %s" % group, current_az=group[0],
               sec_per_az=float(np.mean(times)) if times else None)
        shots, wavs, nts, dts = [], [], [], []
        for az, tab in zip(group, tables[b0:b0 + per_batch]):
            sh, wav, nt, info = prep.shot(float(az), tables=tab)
            shots.append(sh)
            wavs.append(wav)
            nts.append(nt)
            dts.append(info["dt"])
        batch = sweep.ShotBatch(shots, wavs, nts)
        batch.setup()
        k = max(1, min(graph_steps, max(nts)))
        stopped = False
        done_steps = 0
        stream = torch.cuda.Stream(device=shots[0].device)
        # the shot preparation and the zero fills of the new buffers ran on the
        # current stream: the graph's stream must start after them
        cur = torch.cuda.current_stream(shots[0].device)
        stream.wait_stream(cur)
        g = C.c_void_p()
        _abi.check(batch.lib.pe_batch_graph(C.byref(batch.desc), k, stream.cuda_stream,
                                            C.byref(g)), "pe_batch_graph")
        try:
            while done_steps < max(nts):
                if should_stop() if should_stop else os.path.exists(stop_flag):
                    stopped = True
                    break
                _abi.check(batch.lib.pe_graph_launch(g, stream.cuda_stream), "pe_graph_launch")
                done_steps += k
                if (done_steps // k) % 20 == 0:
                    stream.synchronize()
        finally:
            stream.synchronize()
            cur.wait_stream(stream)
            batch.lib.pe_graph_destroy(g)
        if stopped:
            log("[sweep] STOP: batch %s discarded; saved azimuths are final" % group)
            status(root, cfg, phase="stopped")
            return "stopped"
        for az, tr, dt in zip(group, batch.traces, dts):
            obs = save_azimuth(root, cfg, az, tr.cpu().numpy(), dt, h)
            log("[sweep] az %3d: coda %6.1f dB, E1 %.4g, ToF %.3f us, tail %.3f"
                % (az, obs["coda"], obs["E1"], obs["t1"] * 1e6, obs["tail"]))
        el = time.time() - t0
        times.extend([el / len(group)] * len(group))
        del batch, shots, wavs
    status(root, cfg, phase="complete")
    log("[sweep] COMPLETE")
    return "complete"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--root", default=os.path.join(os.getcwd(), "sweeps"))
    ap.add_argument("--new", help="create a sweep with this name")
    ap.add_argument("--set", nargs="*", default=[], help="key=value (JSON) overrides for --new")
    ap.add_argument("--run", help="sweep name or config.json path")
    ap.add_argument("--mem-gb", type=float, default=None)
    args = ap.parse_args(argv)
    if args.new:
        kw = {}
        for kv in args.set:
            k, v = kv.split("=", 1)
            try:
                kw[k] = json.loads(v)
            except json.JSONDecodeError:
                kw[k] = v
        print("created %s" % create(args.root, args.new, **kw))
        return 0
    if args.run:
        run(args.root, load(args.root, args.run), mem_budget_gb=args.mem_gb)
        return 0
    ap.print_help()
    return 2


if __name__ == "__main__":
    raise SystemExit(main())
