"""Azimuth sweeps: the upstream per-azimuth shot and a batched multi-shot runner.

The upstream sweep engine (sim/pipeline/sweep_runner.py:322-388) solves one
azimuth at a time: it re-grids the disk rotated about its axis
(``rotated_grid``, sim/core/rotation_test.py:60-137), builds the rotated
stiffness table, a per-azimuth dt and step count, a Ricker wavelet, the
couplant damping mask and the +x rim transducer footprint
(sweep_runner.py:331-351), then calls ``forward_fused_labels``.

Here the same shots are prepared either on the host (``azimuth_shot``, the
reference semantics) or on the GPU (``GpuAzimuthPrep``, identical
results), and any number of shots of uneven grid size, dt and step count
advance together through ``ShotBatch``: one launch per half-step covers the
workgroups of every shot, per-shot step counters live on the device, and a
captured hipGraph of k steps is replayed until the longest shot is done.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi, host, materials, specimen

C_REF = specimen.C_REF

__all__ = ["disk_build", "rotated_grid", "rim_footprint", "azimuth_shot", "GpuAzimuthPrep",
           "ShotBatch"]


def disk_build(diameter, thickness, h, n_grains, seed, kappa=3.93, fabric_axis=(1.0, 0.0, 0.0)):
    """A Voronoi ice disk in the upstream build layout (specimen.py:178-277):
    ``labels`` (nx, ny, nz) int32 (x, y, z order, -1 outside), ``axes``
    (G, 3), ``seeds`` (G, 3) in metres about the disk centre, ``weights``
    (zero: a plain Voronoi diagram), ``h``."""
    rng = np.random.default_rng(seed)
    nx = int(np.ceil(diameter / h))
    nz = int(np.ceil(thickness / h))
    lab, seeds = specimen.voronoi_disk((nz, nx, nx), h, diameter, thickness, n_grains, rng)
    ax = np.asarray(fabric_axis, float)
    axes = specimen.sample_watson(n_grains, ax / np.linalg.norm(ax), kappa, rng)
    return dict(labels=np.ascontiguousarray(lab.transpose(2, 1, 0)), axes=axes, seeds=seeds,
                weights=np.zeros(len(seeds)), h=h)


def _resample_index(nx_s, nz_s, h_spec, h, rot_deg):
    """Nearest-neighbour source indices of the rotated re-grid
    (rotation_test.py:77-91): (gi, gj, ok) over (nd, nd) solver columns and
    gz over ndz planes."""
    D = nx_s * h_spec
    nd = int(round(D / h))
    ndz = int(round(nz_s * h_spec / h))
    a = np.radians(rot_deg)
    ca, sa = np.cos(a), np.sin(a)
    u = (np.arange(nd) + 0.5) * h - D / 2.0
    X, Y = np.meshgrid(u, u, indexing="ij")
    Xr = ca * X - sa * Y
    Yr = sa * X + ca * Y
    gi = np.rint(Xr / h_spec + (nx_s - 1) / 2.0).astype(int)
    gj = np.rint(Yr / h_spec + (nx_s - 1) / 2.0).astype(int)
    ok = (gi >= 0) & (gi < nx_s) & (gj >= 0) & (gj < nx_s)
    gi, gj = np.clip(gi, 0, nx_s - 1), np.clip(gj, 0, nx_s - 1)
    gz = np.clip((np.arange(ndz) + 0.5) * h / h_spec, 0, nz_s - 1).astype(int)
    return nd, ndz, gi, gj, ok, gz, ca, sa


def rotated_grid(build, h, sponge, rot_deg, single_raster=False):
    """Solver grid of the disk rotated by -rot about its axis, and the c-axes
    rotated likewise (rotation_test.py:60-137, frame 'rigid2').

    Returns (lab (nz, n, n) int32 with an m = sponge + 3 couplant pad, nd,
    m, axes @ R).  Default: nearest-neighbour resample of the build labels;
    ``single_raster``: the tessellation evaluated at the solver points by
    rotating the seeds (power metric with the build's weights).
    """
    labels, axes, h_spec = build["labels"], build["axes"], build["h"]
    nx_s, _, nz_s = labels.shape
    nd, ndz, gi, gj, ok, gz, ca, sa = _resample_index(nx_s, nz_s, h_spec, h, rot_deg)
    D = nx_s * h_spec
    if single_raster:
        from scipy.spatial import cKDTree
        Rm = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
        sd = np.asarray(build["seeds"], float) @ Rm
        wts = np.asarray(build.get("weights", np.zeros(len(sd))), float)
        u = (np.arange(nd) + 0.5) * h - D / 2.0
        zc = (np.arange(ndz) - (ndz - 1) / 2) * h
        X3, Y3, Z3 = np.meshgrid(u, u, zc, indexing="ij")
        ins = (X3 ** 2 + Y3 ** 2) <= (0.5 * D) ** 2
        raw = np.full(X3.shape, -1, np.int32)
        P = np.stack([X3[ins], Y3[ins], Z3[ins]], axis=1)
        k = min(12, len(sd))
        _, idx = cKDTree(sd).query(P, k=k, workers=-1)
        idx = np.atleast_2d(idx.T).T
        d2 = ((P[:, None, :] - sd[idx]) ** 2).sum(-1) - wts[idx]
        raw[ins] = idx[np.arange(len(P)), d2.argmin(1)]
        core = raw.transpose(2, 1, 0)
    else:
        core = np.where(ok[:, :, None], labels[gi, gj][:, :, gz], -1).astype(np.int32)
        core = core.transpose(2, 1, 0)
    m = sponge + 3
    nz, nx = ndz + 2 * m, nd + 2 * m
    lab = np.full((nz, nx, nx), -1, np.int32)
    lab[m:m + ndz, m:m + nd, m:m + nd] = core
    R = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    return lab, nd, m, np.asarray(axes) @ R


def rim_footprint(lab, nd, h, element_d_mm=6.35):
    """The +x rim transducer of sweep_runner.py:342-351: the first cell
    inside the outermost ice cell of the centre row, and the disk of radius
    er = max(int(d/2/h), 1) cells in its (z, y) plane restricted to ice."""
    nz, ny = lab.shape[0], lab.shape[1]
    cz, cy, cx = nz // 2, ny // 2, ny // 2
    Rc = nd // 2
    ixp = next(cx + k - 1 for k in range(Rc - 1, 0, -1) if lab[cz, cy, cx + k] >= 0)
    er = max(int(element_d_mm * 1e-3 / 2 / h), 1)
    # cells beyond the grid (an element wider than the disk) are left out
    pts = [(cz + dz, cy + dy, ixp) for dy in range(-er, er + 1) for dz in range(-er, er + 1)
           if dy * dy + dz * dz <= er * er and 0 <= cz + dz < nz and 0 <= cy + dy < ny
           and lab[cz + dz, cy + dy, ixp] >= 0]
    return pts, 1.0 / len(pts)


def azimuth_shot(build, h, az, coeffs, f0, sponge=10, fluid_damp=0.02, element_d_mm=6.35,
                 record_factor=2.7, single_raster=False, nt=None):
    """forward_fused_labels arguments of one azimuth (sweep_runner.py:331-365):
    dict(lab, axes, h, dt, nt, src_pts, wavelet, rec_groups, sponge_width,
    damp_mask, mat_tables, coeffs)."""
    lab, nd, m, axes_rot = rotated_grid(build, h, sponge, az, single_raster)
    Ct, rho_t = materials.material_tables(axes_rot)
    dt = host.safe_dt_labels(lab + 1, Ct, rho_t, h, coeffs, safety=0.5)
    if nt is None:
        nt = int(record_factor * nd * h / C_REF / dt)
    dm = host.fluid_damp_mask(lab, fluid_damp)
    pts, w = rim_footprint(lab, nd, h, element_d_mm)
    return dict(lab=lab, axes=axes_rot, h=h, dt=dt, nt=int(nt), src_pts=[(p, w) for p in pts],
                wavelet=host.ricker(f0, dt, int(nt)), rec_groups=[(pts, np.full(len(pts), w))],
                sponge_width=sponge, damp_mask=dm, mat_tables=(Ct, rho_t), coeffs=coeffs, nd=nd,
                m=m)


def _tables_job(args):
    axes, ca, sa = args
    R = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
    return materials.material_tables(np.asarray(axes) @ R)


def rotated_tables(axes, azimuths, workers=None):
    """material_tables(axes @ R(az)) for every azimuth, computed in parallel
    worker processes (the per-grain 4-tensor rotation is the host's largest
    per-azimuth cost); identical to computing them one by one."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import os
    jobs = []
    for az in azimuths:
        a = np.radians(az)
        jobs.append((np.asarray(axes, float), float(np.cos(a)), float(np.sin(a))))
    workers = workers or min(16, os.cpu_count() or 1, max(1, len(jobs)))
    if workers <= 1 or len(jobs) <= 1:
        return [_tables_job(j) for j in jobs]
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_tables_job, jobs))


class GpuAzimuthPrep:
    """``azimuth_shot`` computed on the device: the build labels are uploaded
    once; per azimuth the re-grid, couplant distance transform, damping
    volume and stable-dt reduction run as kernels (pe_prep_*), and only the
    rotated stiffness table (O(grains)), the wavelet and the rim footprint
    (one plane of labels) are formed on the host.  Gives the same labels,
    dt, footprint and damping volume as ``azimuth_shot`` (nearest-neighbour
    re-grid, or with ``single_raster`` the power diagram evaluated at the
    rotated solver points).  ``cpml_width`` > 0 absorbs with a CPML of that
    width instead of the sponge (the disk is padded by cpml_width + 3 couplant
    cells, as by sponge + 3 otherwise; profiles from the azimuth's dt and
    tables, R = 1e-4, kappa = 1, f0 the wavelet's).
    """

    def __init__(self, build, h, coeffs, f0, sponge=10, fluid_damp=0.02, element_d_mm=6.35,
                 record_factor=2.7, single_raster=False, device=None, cpml_width=0):
        import torch
        self.torch = torch
        self.dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.build = build
        self.h, self.coeffs, self.f0 = float(h), np.asarray(coeffs, float), float(f0)
        self.sponge, self.fluid_damp = int(sponge), float(fluid_damp)
        self.cpml_width = int(cpml_width)
        self.element_d_mm, self.record_factor = float(element_d_mm), float(record_factor)
        lab = np.ascontiguousarray(build["labels"], np.int32)
        self.src_shape = lab.shape
        self.src = torch.as_tensor(lab, device=self.dev)
        self.single_raster = bool(single_raster)
        if self.single_raster:
            self.w_dev = torch.as_tensor(np.ascontiguousarray(
                build.get("weights", np.zeros(len(build["seeds"]))), np.float64), device=self.dev)
        self.lib = _abi.load()
        self._scratch = None

    def _stream(self):
        return self.torch.cuda.current_stream(self.dev).cuda_stream

    def geometry(self, az):
        nx_s, ny_s, nz_s = self.src_shape
        h_spec = self.build["h"]
        D = nx_s * h_spec
        nd = int(round(D / self.h))
        ndz = int(round(nz_s * h_spec / self.h))
        a = np.radians(az)
        pad = (self.cpml_width if self.cpml_width > 0 else self.sponge) + 3
        return nd, ndz, pad, float(np.cos(a)), float(np.sin(a))

    def labels(self, az):
        """(lab device int32 (nz, n, n), nd, m, axes_rot)."""
        torch = self.torch
        nx_s, ny_s, nz_s = self.src_shape
        nd, ndz, m, ca, sa = self.geometry(az)
        n = nd + 2 * m
        lab = torch.empty((ndz + 2 * m, n, n), dtype=torch.int32, device=self.dev)
        R = np.array([[ca, -sa, 0.0], [sa, ca, 0.0], [0.0, 0.0, 1.0]])
        if self.single_raster:
            sd = torch.as_tensor(np.ascontiguousarray(np.asarray(self.build["seeds"], float) @ R),
                                 device=self.dev)
            D = nx_s * float(self.build["h"])
            _abi.check(self.lib.pe_prep_label_power(sd.data_ptr(), self.w_dev.data_ptr(),
                                                    int(sd.shape[0]), self.h, D, nd, ndz, m,
                                                    lab.data_ptr(), self._stream()),
                       "pe_prep_label_power")
        else:
            _abi.check(self.lib.pe_prep_rotate(self.src.data_ptr(), nx_s, ny_s, nz_s,
                                               float(self.build["h"]), self.h, ca, sa, nd, ndz,
                                               m, lab.data_ptr(), self._stream()),
                       "pe_prep_rotate")
        return lab, nd, m, np.asarray(self.build["axes"]) @ R

    def damping(self, lab):
        torch = self.torch
        nz, ny, nx = lab.shape
        ncell = lab.numel()
        if self._scratch is None or self._scratch[0].numel() < ncell:
            self._scratch = [torch.empty(ncell, dtype=torch.int32, device=self.dev)
                             for _ in range(2)]
        width = 0 if self.cpml_width > 0 else self.sponge
        pz, py, px = [torch.as_tensor(p, device=self.dev)
                      for p in host.sponge_axis_profiles((nz, ny, nx), width)]
        damp = torch.empty((nz, ny, nx), dtype=torch.float32, device=self.dev)
        _abi.check(self.lib.pe_prep_edt_damp(lab.data_ptr(), nz, ny, nx, self.fluid_damp,
                                             pz.data_ptr(), py.data_ptr(), px.data_ptr(),
                                             self._scratch[0].data_ptr(),
                                             self._scratch[1].data_ptr(), damp.data_ptr(),
                                             self._stream()), "pe_prep_edt_damp")
        return damp

    def stable_dt(self, mats, Ct, rho_t):
        torch = self.torch
        nz, ny, nx = mats.shape
        Ct_d = torch.as_tensor(np.ascontiguousarray(Ct), device=self.dev)
        rho_d = torch.as_tensor(np.ascontiguousarray(rho_t), device=self.dev)
        nb = 1024
        part = torch.empty(nb, dtype=torch.float64, device=self.dev)
        prec = _abi.PE_F32 if Ct_d.dtype == torch.float32 else _abi.PE_F64
        _abi.check(self.lib.pe_prep_vmax(mats.data_ptr(), nz, ny, nx, Ct_d.data_ptr(),
                                         rho_d.data_ptr(), prec, part.data_ptr(), nb,
                                         self._stream()), "pe_prep_vmax")
        vmax = float(np.sqrt(float(part.max())))
        return 0.5 * host.cfl_limit(self.coeffs) * self.h / vmax

    def shot(self, az, precision="f32", nt=None, kernel="tiled", tables=None):
        """(DeviceShot, device wavelet, nt, info dict) of azimuth ``az``;
        ``tables``: its (Ct, rho_t) if already computed (rotated_tables)."""
        from .solver import DeviceShot, _csr
        lab, nd, m, axes_rot = self.labels(az)
        Ct, rho_t = materials.material_tables(axes_rot) if tables is None else tables
        mats = lab + 1
        dt = self.stable_dt(mats, Ct, rho_t)
        if nt is None:
            nt = int(self.record_factor * nd * self.h / C_REF / dt)
        damp = self.damping(lab)
        nz, ny, _ = lab.shape
        cz, cy = nz // 2, ny // 2
        row = lab[cz, cy].cpu().numpy()
        cx, Rc = ny // 2, nd // 2
        ixp = next(cx + k - 1 for k in range(Rc - 1, 0, -1) if row[cx + k] >= 0)
        plane = lab[:, :, ixp].cpu().numpy()
        er = max(int(self.element_d_mm * 1e-3 / 2 / self.h), 1)
        pts = [(cz + dz, cy + dy, ixp) for dy in range(-er, er + 1) for dz in range(-er, er + 1)
               if dy * dy + dz * dz <= er * er and 0 <= cz + dz < nz and 0 <= cy + dy < ny
               and plane[cz + dz, cy + dy] >= 0]
        w = 1.0 / len(pts)
        shape = tuple(lab.shape)
        src_idx = np.array([np.ravel_multi_index(p, shape) for p in pts], np.int64)
        rec_ptr, rec_idx, rec_w = _csr([(pts, np.full(len(pts), w))], shape)
        cp = None
        if self.cpml_width > 0:
            from . import cpml
            cp = cpml.Cpml(shape, self.cpml_width, self.h, dt, cpml.vmax_table(Ct, rho_t),
                           self.f0)
        shot = DeviceShot(mats, Ct, rho_t, damp, self.coeffs, self.h, dt, src_idx,
                          np.full(len(pts), w), rec_ptr, rec_idx, rec_w, precision=precision,
                          kernel=kernel, cpml=cp)
        wav = shot.wavelet(host.ricker(self.f0, dt, int(nt)))
        return shot, wav, int(nt), dict(dt=dt, nd=nd, m=m, pts=pts, axes=axes_rot, lab=lab)


class ShotBatch:
    """Many DeviceShots advanced together (pe_batch).

    ``shots``: DeviceShot objects (one precision, tiled kernels; the tile
    variants of shots[0] are used for all).  ``wavelets``: per-shot device
    float64 tensors; ``nts``: per-shot step counts.  Traces are allocated
    per shot.  ``run(graph_steps)`` captures ``graph_steps`` batched steps in
    a hipGraph and replays it until every shot has run its steps.

    ``noalloc=True`` keeps the batch table and step counters in the
    library's own device memory (pe_batch_setup_noalloc) instead of tensors
    allocated here: batches set up that way on one device share that memory,
    so only the most recently set-up one may run.
    """

    def __init__(self, shots, wavelets, nts, record=True, noalloc=False):
        torch = shots[0].torch
        self.torch = torch
        self.shots = list(shots)
        self.nts = [int(n) for n in nts]
        self.wavelets = list(wavelets)
        if not (len(self.shots) == len(self.nts) == len(self.wavelets)) or not self.shots:
            raise ValueError("one wavelet and one step count per shot are needed")
        for i, (s, w, n) in enumerate(zip(self.shots, self.wavelets, self.nts)):
            # the batched source kernel reads wavelet[n] for every step n < nt
            if not torch.is_tensor(w) or w.device != s.device or w.dtype != torch.float64 \
                    or not w.is_contiguous():
                raise ValueError(f"shot {i}: the wavelet must be a contiguous float64 tensor "
                                 f"on the shot's device")
            if w.numel() < n or n < 0:
                raise ValueError(f"shot {i}: wavelet has {w.numel()} samples, nt = {n}")
        self.traces = [s.new_traces(n) for s, n in zip(self.shots, self.nts)] if record else None
        self.lib = _abi.load()
        n = len(self.shots)
        self._descs = (_abi.PeShot * n)(*[s.desc for s in self.shots])
        self._wav = (C.c_void_p * n)(*[w.data_ptr() for w in self.wavelets])
        self._tr = (C.c_void_p * n)(*[t.data_ptr() for t in self.traces]) if record else None
        self._nt = (C.c_int32 * n)(*self.nts)
        nbytes = self.lib.pe_batch_table_bytes(self._descs, n)
        if nbytes <= 0:
            raise ValueError("invalid batch: %s" % self.lib.pe_last_error().decode())
        dev = shots[0].device
        self.noalloc = bool(noalloc)
        if self.noalloc:
            self.table = self.step = None
        else:
            self.table = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
            self.step = torch.zeros(n, dtype=torch.int32, device=dev)
        b = _abi.PeBatch()
        b.n_shot = n
        b.shots = C.cast(self._descs, C.c_void_p)
        b.wavelet = C.cast(self._wav, C.c_void_p)
        b.traces = None if self._tr is None else C.cast(self._tr, C.c_void_p)
        b.nt = C.cast(self._nt, C.c_void_p)
        b.table = None if self.noalloc else self.table.data_ptr()
        b.step = None if self.noalloc else self.step.data_ptr()
        self.desc = b
        self._graph = None
        self._graph_steps = 0

    def _stream(self):
        return self.shots[0]._stream()

    def setup(self):
        """Upload the table and zero the step counters (blocking)."""
        if self.noalloc:
            _abi.check(self.lib.pe_batch_setup_noalloc(C.byref(self.desc), self._stream()),
                       "pe_batch_setup_noalloc")
        else:
            _abi.check(self.lib.pe_batch_setup(C.byref(self.desc), self._stream()),
                       "pe_batch_setup")

    def step(self, k=1):
        _abi.check(self.lib.pe_batch_step(C.byref(self.desc), int(k), self._stream()),
                   "pe_batch_step")

    def run(self, graph_steps=32, progress=None, progress_every=1000):
        """Run every shot to its step count; returns the per-shot traces.
        ``progress(done, total)`` is called on the host about every
        ``progress_every`` steps (after a stream synchronisation)."""
        torch = self.torch
        self.setup()
        total = max(self.nts) if self.nts else 0
        k = max(1, min(int(graph_steps), total))
        stream = torch.cuda.Stream(device=self.shots[0].device)
        cur = torch.cuda.current_stream(self.shots[0].device)
        stream.wait_stream(cur)
        if self._graph is None or self._graph_steps != k:
            self.close()
            g = C.c_void_p()
            _abi.check(self.lib.pe_batch_graph(C.byref(self.desc), k, stream.cuda_stream,
                                               C.byref(g)), "pe_batch_graph")
            self._graph, self._graph_steps = g, k
        done = 0
        next_report = progress_every
        while done < total:
            _abi.check(self.lib.pe_graph_launch(self._graph, stream.cuda_stream),
                       "pe_graph_launch")
            done += k
            if progress is not None and (done >= next_report or done >= total):
                stream.synchronize()
                progress(min(done, total), total)
                next_report += progress_every
        cur.wait_stream(stream)
        return self.traces

    def close(self):
        if self._graph is not None:
            self.lib.pe_graph_destroy(self._graph)
            self._graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
