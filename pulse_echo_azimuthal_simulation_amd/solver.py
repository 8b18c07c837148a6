"""forward_fused_labels on MI355X: the upstream call surface over libpefdtd.so.

``forward_fused_labels`` keeps the upstream signature, defaults, validation
messages and return value (sim/core/fdtd.py:1618-1816): a specimen label
volume (-1 = couplant) plus per-grain c-axes, or explicit (Ct, rho_t)
tables, go in; an (nt, n_rec) float32 array of receiver traces comes out.

The time loop runs in hand-written gfx950 kernels reached through the C ABI
(include/pefdtd.h).  PyTorch only allocates the device buffers and supplies
the stream.  Labels are int32 (the upstream uint8 limit of 254 grains is
lifted; ``compat_uint8=True`` restores the upstream refusal).
"""
from __future__ import annotations

import ctypes as C
import ctypes as _ctypes

import numpy as np

from . import _abi, host, materials

__all__ = ["DeviceShot", "forward_fused_labels", "forward_fused", "prepare_shot"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible: the solver runs on the GPU only")
    return torch


class DeviceShot:
    """Device-resident state of one shot, plus its pe_shot descriptor.

    All buffers are torch tensors on the current HIP device; the descriptor
    holds their addresses.  Fields start at zero.  ``mats`` and ``damp`` may
    be given as device tensors (prepared on the GPU), else host arrays.
    ``cpml``: a ``cpml.Cpml`` for a CPML boundary (profiles on the device,
    memory variables allocated here, zero; the tiled kernels also get a
    6 x ncell buffer for the band corrections), or None.
    """

    def __init__(self, mats, Ct, rho_t, damp, coeffs, h, dt, src_idx, src_w,
                 rec_ptr, rec_idx, rec_w, record="pressure", precision="f32",
                 kernel="tiled", act_mask=None, device=None, tile=(0, 0), zchunk=(0, 0),
                 cpml=None):
        torch = _torch()
        self.torch = torch
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        self.device = dev
        self.real = torch.float32 if precision == "f32" else torch.float64
        self.np_real = np.float32 if precision == "f32" else np.float64
        mats_t = mats if torch.is_tensor(mats) else None
        if mats_t is None:
            mats = np.ascontiguousarray(mats, np.int32)
        self.shape = tuple(int(s) for s in mats.shape)
        nz, ny, nx = self.shape
        ncell = nz * ny * nx
        Ct = np.asarray(Ct).reshape(-1, 36)
        if act_mask is None:
            act_mask = materials.active_mask(Ct)
        coeffs = np.asarray(coeffs, float)
        if len(coeffs) > 4:
            raise ValueError("at most 4 stencil coefficients (order 8) are supported")

        def dev_real(a):
            if torch.is_tensor(a):
                return a.to(device=dev, dtype=self.real).contiguous()
            return torch.as_tensor(np.ascontiguousarray(a, self.np_real), device=dev)

        self.mats = (mats_t.to(device=dev, dtype=torch.int32).contiguous() if mats_t is not None
                     else torch.as_tensor(mats, device=dev))
        self.Ct = dev_real(Ct)
        self.rho_t = dev_real(rho_t)
        self.damp = None
        if damp is not None:
            self.damp = (dev_real(damp).reshape(self.shape) if torch.is_tensor(damp)
                         else dev_real(np.asarray(damp).reshape(self.shape)))
        self.v = [torch.zeros(self.shape, dtype=self.real, device=dev) for _ in range(3)]
        self.s = [torch.zeros(self.shape, dtype=self.real, device=dev) for _ in range(6)]
        self.g = (torch.zeros((6,) + self.shape, dtype=self.real, device=dev)
                  if kernel == "direct" else None)
        self.src_idx = torch.as_tensor(np.ascontiguousarray(src_idx, np.int64), device=dev)
        self.src_w = dev_real(src_w)
        self.rec_ptr = torch.as_tensor(np.ascontiguousarray(rec_ptr, np.int32), device=dev)
        self.rec_idx = torch.as_tensor(np.ascontiguousarray(rec_idx, np.int64), device=dev)
        self.rec_w = dev_real(rec_w)
        self.n_rec = len(rec_ptr) - 1

        d = _abi.PeShot()
        d.nz, d.ny, d.nx = nz, ny, nx
        d.nc = len(coeffs)
        for j in range(4):
            d.coeffs[j] = float(coeffs[j]) if j < len(coeffs) else 0.0
        d.dt = float(dt)
        d.inv_h = 1.0 / float(h)
        d.dt_h = float(dt) / float(h)
        d.n_mat = int(self.rho_t.numel())
        d.precision = _abi.PE_F32 if precision == "f32" else _abi.PE_F64
        d.act_mask = int(act_mask)
        d.mats = self.mats.data_ptr()
        d.Ct = self.Ct.data_ptr()
        d.rho_t = self.rho_t.data_ptr()
        d.damp = None if self.damp is None else self.damp.data_ptr()
        for k in range(3):
            d.v[k] = self.v[k].data_ptr()
        for k in range(6):
            d.s[k] = self.s[k].data_ptr()
        d.g = None if self.g is None else self.g.data_ptr()
        d.n_src = int(self.src_idx.numel())
        d.src_idx = self.src_idx.data_ptr() if d.n_src else None
        d.src_w = self.src_w.data_ptr() if d.n_src else None
        d.n_rec = self.n_rec
        d.rec_ptr = self.rec_ptr.data_ptr()
        d.rec_idx = self.rec_idx.data_ptr() if self.rec_idx.numel() else None
        d.rec_w = self.rec_w.data_ptr() if self.rec_w.numel() else None
        d.record_mode = _abi.RECORD_MODES[record]
        d.kernel = _abi.KERNELS[kernel]
        d.tile[0], d.tile[1] = int(tile[0]), int(tile[1])
        d.zchunk[0], d.zchunk[1] = int(zchunk[0]), int(zchunk[1])
        self.cpml = cpml
        self.cpml_prof, self.psi = [], []
        if cpml is not None:
            if tuple(cpml.shape) != self.shape:
                raise ValueError(f"CPML built for {cpml.shape}, grid is {self.shape}")
            d.cpml_width = int(cpml.width)
            for a in range(3):
                self.cpml_prof.append(dev_real(cpml.prof[a]))
                self.psi.append(torch.zeros((6, int(cpml.vol[a])), dtype=self.real, device=dev))
                d.cpml_prof[a] = self.cpml_prof[a].data_ptr()
                d.cpml_psi[a] = self.psi[a].data_ptr()
            if kernel == "tiled":
                # band corrections the tiled kernels' pre-pass writes (band
                # cells only are written and read)
                self.cpml_corr = torch.empty((6,) + self.shape, dtype=self.real, device=dev)
                d.cpml_corr = self.cpml_corr.data_ptr()
        self.desc = d
        self.lib = _abi.load()
        _abi.check(self.lib.pe_check_shot(C.byref(d)), "pe_check_shot")

    # ---------------------------------------------------------------- steps
    def _stream(self):
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def reset(self):
        """Zero the fields (and the CPML memory variables)."""
        for f in self.v + self.s + self.psi:
            f.zero_()
        if self.g is not None:
            self.g.zero_()

    def stress(self):
        _abi.check(self.lib.pe_stress(C.byref(self.desc), self._stream()), "pe_stress")

    def velocity(self):
        _abi.check(self.lib.pe_velocity(C.byref(self.desc), self._stream()), "pe_velocity")

    def source(self, wavelet_dev, n):
        self._check_run(wavelet_dev, n, int(n) + 1, None)
        _abi.check(self.lib.pe_source(C.byref(self.desc), wavelet_dev.data_ptr(), int(n),
                                      self._stream()), "pe_source")

    def record(self, traces_dev, n):
        if not (0 <= int(n) < traces_dev.shape[0]) or traces_dev.dtype != self.real \
                or traces_dev.dim() != 2 or traces_dev.shape[1] != self.n_rec \
                or not traces_dev.is_contiguous():
            raise ValueError(f"record: row {n} outside the (nt, n_rec) traces buffer")
        _abi.check(self.lib.pe_record(C.byref(self.desc), traces_dev.data_ptr(), int(n),
                                      self._stream()), "pe_record")

    def _check_run(self, wavelet_dev, n0, n1, traces_dev):
        """The kernels read wavelet[n] and write traces row n for every step
        n0 <= n < n1 through raw pointers: refuse buffers that are too short."""
        torch = self.torch
        if not (0 <= int(n0) <= int(n1)):
            raise ValueError(f"bad step range [{n0}, {n1})")
        if not torch.is_tensor(wavelet_dev) or wavelet_dev.device != self.device \
                or wavelet_dev.dtype != torch.float64 or not wavelet_dev.is_contiguous():
            raise ValueError("wavelet must be a contiguous float64 tensor on the shot's device "
                             "(DeviceShot.wavelet)")
        if wavelet_dev.numel() < int(n1):
            raise ValueError(f"wavelet has {wavelet_dev.numel()} samples, steps run to n = {n1}")
        if traces_dev is not None:
            if traces_dev.device != self.device or traces_dev.dtype != self.real \
                    or not traces_dev.is_contiguous() or traces_dev.dim() != 2 \
                    or traces_dev.shape[1] != self.n_rec:
                raise ValueError("traces must be a contiguous (nt, n_rec) tensor of the shot's "
                                 "real type on its device (DeviceShot.new_traces)")
            if traces_dev.shape[0] < int(n1):
                raise ValueError(f"traces hold {traces_dev.shape[0]} rows, steps run to n = {n1}")

    def run(self, wavelet_dev, n0, n1, traces_dev=None):
        """Advance steps n0..n1-1 (asynchronous on the current stream)."""
        self._check_run(wavelet_dev, n0, n1, traces_dev)
        _abi.check(self.lib.pe_run(C.byref(self.desc), int(n0), int(n1),
                                   wavelet_dev.data_ptr(),
                                   None if traces_dev is None else traces_dev.data_ptr(),
                                   self._stream()), "pe_run")

    def run_timed(self, wavelet_dev, n0, n1, traces_dev=None):
        """Advance steps n0..n1-1 with per-kernel hipEvent timing; blocks.
        Returns milliseconds summed over the run: (stress, velocity, source/record)."""
        self._check_run(wavelet_dev, n0, n1, traces_dev)
        ms = (C.c_double * 3)()
        _abi.check(self.lib.pe_run_timed(C.byref(self.desc), int(n0), int(n1),
                                         wavelet_dev.data_ptr(),
                                         None if traces_dev is None else traces_dev.data_ptr(),
                                         self._stream(), ms), "pe_run_timed")
        return float(ms[0]), float(ms[1]), float(ms[2])

    def new_traces(self, nt):
        return self.torch.zeros((int(nt), self.n_rec), dtype=self.real, device=self.device)

    def wavelet(self, wavelet):
        return self.torch.as_tensor(np.ascontiguousarray(wavelet, np.float64), device=self.device)

    def fields(self):
        """(v, s) copied to host numpy arrays."""
        self.torch.cuda.synchronize(self.device)
        return ([a.cpu().numpy() for a in self.v], [a.cpu().numpy() for a in self.s])


def _csr(groups, shape):
    ptr, idx, w = [0], [], []
    for pts, wts in groups:
        pts = [tuple(int(a) for a in p) for p in pts]
        wts = np.asarray(wts, float).ravel()
        idx.extend(np.ravel_multi_index(tuple(np.array(pts).T), shape).tolist() if pts else [])
        w.extend(wts.tolist())
        ptr.append(len(idx))
    return np.array(ptr, np.int32), np.array(idx, np.int64), np.array(w, float)


def prepare_shot(lab, axes, h, dt, src_pts, rec_groups, order=8, coeffs=None,
                 sponge_width=0, sponge_strength=0.09, damp_mask=None,
                 c_couplant=1500.0, rho_couplant=300.0, mat_tables=None,
                 compat_uint8=False):
    """Host-side setup and validation of forward_fused_labels
    (sim/core/fdtd.py:1631-1726).  Returns a dict of host arrays."""
    if coeffs is None:
        coeffs = host.optimised_coeffs(order)
    coeffs = np.asarray(coeffs, float)
    if len(coeffs) > 4:
        raise ValueError(
            f"forward_fused_labels supports order <= 8 ({len(coeffs)} coefficients given): "
            f"the z-marching kernels hold only +/-4 z-planes. Use order <= 8.")
    lab = np.asarray(lab)
    if lab.ndim != 3:
        raise ValueError("lab must be a 3-D (nz, ny, nx) label volume")
    lmin, lmax = int(lab.min()), int(lab.max())
    hi = 254 if compat_uint8 else np.iinfo(np.int32).max - 1
    if lmin < -1 or lmax > hi:
        raise ValueError(
            f"labels must lie in [-1, {hi}]; got [{lmin}, {lmax}]"
            + (": the uint8 material index would wrap silently (>=255 grains needs a "
               "wider label type)." if compat_uint8 else "."))
    if mat_tables is None and lmax >= len(axes):
        raise ValueError(f"label {lmax} has no c-axis (len(axes) = {len(axes)}).")
    shape = lab.shape
    mats = (lab.astype(np.int64) + 1).astype(np.int32)
    if mat_tables is None:
        Ct, rho_t = materials.material_tables(axes, c_couplant, rho_couplant)
    else:
        Ct, rho_t = mat_tables
        rho_t = np.ascontiguousarray(rho_t, np.float32)
        Ct = np.ascontiguousarray(Ct, np.float32).reshape(len(rho_t), 36)
        if lmax + 1 >= len(rho_t):
            raise ValueError(
                f"label {lmax} has no row in mat_tables (n_mat = {len(rho_t)}, "
                f"need >= {lmax + 2}).")
    damp = None
    if sponge_width > 0:
        damp = host.sponge_profile(shape, sponge_width, sponge_strength)
    if damp_mask is not None:
        dm = np.asarray(damp_mask, np.float32)
        damp = dm if damp is None else damp * dm
    src_idx = np.array([np.ravel_multi_index(tuple(int(a) for a in p), shape)
                        for p, _ in src_pts], np.int64)
    src_w = np.array([w for _, w in src_pts], float)
    if len(np.unique(src_idx)) != len(src_idx):
        raise ValueError("source footprint cells must be unique")
    rec_ptr, rec_idx, rec_w = _csr(rec_groups, shape)
    if damp is not None:
        dflat = damp.reshape(-1)
        bad = [int(q) for q in src_idx if dflat[q] < 0.999]
        bad += [int(q) for q in rec_idx if dflat[q] < 0.999]
        if bad:
            raise ValueError(
                f"{len(bad)} source/receiver cells sit where damp < 0.999 (e.g. flat index "
                f"{bad[0]}): the deferred-damping scheme records pre-damp stress there. "
                f"Move the footprint out of the sponge/fluid-damped region.")
    return dict(mats=mats, Ct=Ct, rho_t=rho_t, damp=damp, coeffs=coeffs, h=h, dt=dt,
                src_idx=src_idx, src_w=src_w, rec_ptr=rec_ptr, rec_idx=rec_idx, rec_w=rec_w)


def forward_fused_labels(lab, axes, h, dt, nt, src_pts, wavelet, rec_groups,
                         order=8, coeffs=None, sponge_width=0,
                         sponge_strength=0.09, damp_mask=None,
                         c_couplant=1500.0, rho_couplant=300.0,
                         record="pressure", slab=32, stress_smem=False,
                         progress=None, mat_tables=None, *, precision="f32",
                         kernel="tiled", compat_uint8=False, progress_every=200, tune=False,
                         boundary="sponge", cpml_width=20, cpml_R=1e-4, cpml_f0=None,
                         cpml_kappa=1.0):
    """Label-indexed anisotropic elastic FDTD, drop-in for the upstream
    ``fdtd.forward_fused_labels`` (sim/core/fdtd.py:1618-1816).

    Returns the (nt, n_rec) traces as a float32 numpy array (float64 with
    ``precision="f64"``).  ``slab`` and ``stress_smem`` are accepted for
    signature compatibility; the gfx950 kernels march z without slabs.
    ``progress(frac)`` is called on the host every ``progress_every`` steps
    (at n = 0, 200, ...) and may raise to abort, as upstream.  ``tune=True``
    times the tiled kernels' launch shapes on this grid first (see autotune).

    ``boundary="cpml"`` (no upstream counterpart) replaces the sponge by a
    convolutional PML of ``cpml_width`` cells (see cpml.py): reflection
    coefficient ``cpml_R``, frequency shift pi * ``cpml_f0`` (default: the
    wavelet's spectral peak), ``cpml_kappa`` the largest stretch.  The sponge
    must then be off (sponge_width 0); ``damp_mask`` still applies.
    """
    if record not in _abi.RECORD_MODES:
        raise ValueError(f"record must be one of {sorted(_abi.RECORD_MODES)}")
    if boundary not in ("sponge", "cpml"):
        raise ValueError("boundary must be 'sponge' or 'cpml'")
    if boundary == "cpml" and sponge_width > 0:
        raise ValueError("boundary='cpml' replaces the sponge: pass sponge_width=0")
    prep = prepare_shot(lab, axes, h, dt, src_pts, rec_groups, order, coeffs, sponge_width,
                        sponge_strength, damp_mask, c_couplant, rho_couplant, mat_tables,
                        compat_uint8)
    nt = int(nt)
    wavelet = np.asarray(wavelet, float)
    if len(wavelet) < nt:
        raise ValueError(f"wavelet has {len(wavelet)} samples, nt = {nt}")
    cp = None
    if boundary == "cpml":
        from . import cpml as _cpml
        f0 = _cpml.dominant_frequency(wavelet[:nt], dt) if cpml_f0 is None else cpml_f0
        cp = _cpml.Cpml(prep["mats"].shape, cpml_width, h, dt,
                        _cpml.vmax_table(prep["Ct"], prep["rho_t"]), f0, cpml_R, cpml_kappa)
    shot = DeviceShot(prep["mats"], prep["Ct"], prep["rho_t"], prep["damp"], prep["coeffs"],
                      h, dt, prep["src_idx"], prep["src_w"], prep["rec_ptr"],
                      prep["rec_idx"], prep["rec_w"], record=record, precision=precision,
                      kernel=kernel, cpml=cp)
    if tune and kernel == "tiled":
        from . import autotune
        autotune.tune(shot)
    wav = shot.wavelet(wavelet[:nt])
    traces = shot.new_traces(nt)
    # upstream calls progress(n / nt) after step n whenever n % 200 == 0
    n = 0
    while n < nt:
        n1 = min(nt, n + 1 if n % progress_every == 0 else (n // progress_every + 1) * progress_every)
        shot.run(wav, n, n1, traces)
        if progress is not None and n % progress_every == 0:
            shot.torch.cuda.synchronize(shot.device)
            progress(n / nt)
        n = n1
    return traces.cpu().numpy()


def forward_fused(C, rho, h, dt, nt, src_pts, wavelet, rec_groups, order=8, coeffs=None,
                  sponge_width=0, sponge_strength=0.09, damp_mask=None, record="pressure",
                  progress=None, *, precision="f32", progress_every=200):
    """Dense-table twin of ``forward_fused_labels`` (the upstream
    ``fdtd.forward_fused``, sim/core/fdtd.py:831-921): stiffness as per-cell
    volumes ``C`` (a dict "Cij", i <= j, of volumes or scalars; C_ji = C_ij),
    density as a volume ``rho``; the sponge and ``damp_mask`` damp v and s
    after the velocity update and the traces are recorded after the damping.
    Returns the (nt, n_rec) float32 traces (float64 with precision="f64").

    It runs separate kernels (pe_dense_run) with the label kernels'
    statement order, so on a label volume's dense expansion its traces equal
    forward_fused_labels' bit for bit wherever the receivers sit at damp ==
    1 -- the upstream's bit-equivalence gate (sim/core/validate_labels.py:
    66-100).
    """
    torch = _torch()
    if record not in _abi.RECORD_MODES:
        raise ValueError(f"record must be one of {sorted(_abi.RECORD_MODES)}")
    if coeffs is None:
        coeffs = host.optimised_coeffs(order)
    coeffs = np.asarray(coeffs, float)
    if len(coeffs) > 4:
        raise ValueError("forward_fused supports order <= 8 (at most 4 coefficients)")
    rho_h = np.asarray(rho, float)
    if rho_h.ndim != 3:
        raise ValueError("rho must be a (nz, ny, nx) volume")
    shape = rho_h.shape
    nz, ny, nx = shape
    nt = int(nt)
    wavelet = np.asarray(wavelet, float)
    if len(wavelet) < nt:
        raise ValueError(f"wavelet has {len(wavelet)} samples, nt = {nt}")
    real = np.float32 if precision == "f32" else np.float64
    # the planes hold float32 values in both precisions, as the label path's
    # tables do: precision="f64" is float64 arithmetic on them
    Cflat = np.zeros((36,) + shape, np.float32)
    act = 0
    for I in range(1, 7):
        for J in range(1, 7):
            a = np.asarray(C["C%d%d" % (min(I, J), max(I, J))], float)
            if not np.any(a):
                continue
            Cflat[(I - 1) * 6 + (J - 1)] = a if a.ndim == 3 else float(a)
            act |= 1 << ((I - 1) * 6 + (J - 1))
    rv = [host.face_density(rho_h, ax).astype(np.float32).astype(real) for ax in (2, 1, 0)]
    damp = host.sponge_profile(shape, sponge_width, sponge_strength) if sponge_width > 0 else None
    if damp_mask is not None:
        dm = np.asarray(damp_mask, np.float32)
        damp = dm if damp is None else damp * dm
    src_idx = np.array([np.ravel_multi_index(tuple(int(a) for a in p), shape)
                        for p, _ in src_pts], np.int64)
    src_w = np.array([w for _, w in src_pts], float)
    if len(np.unique(src_idx)) != len(src_idx):
        raise ValueError("source footprint cells must be unique")
    rec_ptr, rec_idx, rec_w = _csr(rec_groups, shape)
    # the shot's label and table are not read by the dense kernels
    sh = DeviceShot(np.zeros(shape, np.int32), np.zeros((1, 36)), np.ones(1), damp, coeffs, h,
                    dt, src_idx, src_w, rec_ptr, rec_idx, rec_w, record=record,
                    precision=precision, kernel="direct", act_mask=act)
    dev = sh.device
    Cd = torch.as_tensor(Cflat, device=dev).to(sh.real)
    del Cflat
    rvd = [torch.as_tensor(np.ascontiguousarray(r), device=dev) for r in rv]
    d = _abi.PeDense()
    d.Cflat = Cd.data_ptr()
    d.act_mask = int(act)
    for k in range(3):
        d.rv[k] = rvd[k].data_ptr()
    wav = sh.wavelet(wavelet[:nt])
    traces = sh.new_traces(nt)
    n = 0
    while n < nt:
        n1 = min(nt, n + 1 if n % progress_every == 0 else (n // progress_every + 1) * progress_every)
        _abi.check(sh.lib.pe_dense_run(_ctypes.byref(sh.desc), _ctypes.byref(d), int(n), int(n1),
                                       wav.data_ptr(), traces.data_ptr(), sh._stream()),
                   "pe_dense_run")
        if progress is not None and n % progress_every == 0:
            torch.cuda.synchronize(dev)
            progress(n / nt)
        n = n1
    return traces.cpu().numpy()
