"""Misfit gradients through the exact discrete adjoint of the GPU step.

The upstream repository differentiates only its CPU array-path solver
(vendor/ringfwi/elastic3d.py:276-463, chained to per-voxel c-axis angles
by vendor/ringfwi/axisfield.py:37-161).  Here the label-indexed GPU step of
``forward_fused_labels`` is differentiated by hand-written reverse kernels
(``pe_adjoint_run``), with

* the gradient reduced per material (grain) in float64, deterministically
  (per-column material runs, see ``run_tables``);
* a CPML boundary: the adjoints of the memory variables are carried through
  the reverse pass, checkpoints hold the memory variables, and the gradient
  reads the strain history;
* wavefield checkpointing: the forward pass keeps the state every
  ``checkpoint_every`` steps; each segment is recomputed from its checkpoint
  while a per-step history is stored -- the strains the stress half-step
  computes from v (6 fields; the reverse pass then forms only their moves)
  or v itself (3 fields, when memory is short) -- then the reverse steps of
  that segment run; segments whose history fits in ``history_bytes`` are
  stored during the first pass and not recomputed;
* a ``torch.autograd.Function`` (``traces_from_table``) whose backward is that
  adjoint, so a misfit written in torch differentiates down to the stiffness
  table and, through ``fabric.stiffness_table`` (Bond rotation in torch), to
  per-grain c-axis angles.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi

__all__ = ["run_tables", "DeviceAdjoint", "misfit_gradient", "traces_from_table"]


def run_tables(mats, zbreak=0):
    """Per-column material runs of a (nz, ny, nx) material volume.

    A run is a maximal z-interval of one (y, x) column with one material,
    also cut at z % zbreak == 0 when zbreak > 0.  Runs are numbered column
    by column (flat y*nx + x), in z order inside a column.  With nchunk =
    ceil(nz / zbreak) (1 when zbreak = 0), run_ptr[c*nchunk + k] is the
    first run of column c at or after z = k*zbreak.  Returns (run_ptr
    (ncol*nchunk+1,), mat_ptr (n_mat+1,), mat_runs (n_runs,), run_mat), all
    int32, where mat_runs lists the runs of each material ascending.
    """
    m = np.asarray(mats)
    nz = m.shape[0]
    cols = m.reshape(nz, -1).T                           # (ncol, nz)
    ncol = cols.shape[0]
    start = np.ones(cols.shape, bool)
    start[:, 1:] = cols[:, 1:] != cols[:, :-1]
    zb = zbreak if 0 < zbreak < nz else nz
    if zbreak > 0:
        start[:, ::zb] = True
    nchunk = -(-nz // zb)
    first = np.cumsum(start.ravel()).reshape(ncol, nz) - 1   # run index of every cell
    run_ptr = np.empty(ncol * nchunk + 1, np.int64)
    run_ptr[:-1] = first[:, ::zb].ravel()
    run_ptr[-1] = int(start.sum())
    run_mat = cols[start].astype(np.int64)               # row-major: column, then z
    n_mat = int(m.max()) + 1 if m.size else 1
    order = np.argsort(run_mat, kind="stable")
    mat_ptr = np.zeros(n_mat + 1, np.int64)
    np.cumsum(np.bincount(run_mat, minlength=n_mat), out=mat_ptr[1:])
    if run_ptr[-1] >= 2 ** 31:
        raise ValueError("more than 2^31-1 material runs")
    return (run_ptr.astype(np.int32), mat_ptr.astype(np.int32), order.astype(np.int32),
            run_mat.astype(np.int32))


class DeviceAdjoint:
    """Adjoint fields, scratch and gradient accumulators of one DeviceShot.

    n_proj = 0 accumulates dJ/dCt for all 36 table entries of every
    material; 1 <= n_proj <= 4 accumulates projections sum_k proj[m, p, k] *
    dJ/dCt[m, k] instead (``proj`` of shape (n_mat, n_proj, 36)).
    """

    def __init__(self, shot, n_proj=0, proj=None, zbreak=None, gradient=True, kernel=None,
                 hist="velocity"):
        torch = shot.torch
        self.shot = shot
        self.torch = torch
        dev, real = shot.device, shot.real
        n_mat = int(shot.rho_t.numel())
        if n_proj and proj is None:
            raise ValueError("n_proj > 0 needs a projection table")
        self.K = n_proj if n_proj else 36
        if kernel is None:
            kernel = "tiled" if shot.desc.kernel == _abi.PE_KERNEL_TILED else "direct"
        if zbreak is None:
            # z-chunk of the tiled gradient kernel: enough workgroups to fill
            # the device for the grid size, at least 16 planes (80x160x160,
            # 4000 steps: gradient 1.43 s at 16, 1.48 at 8, 1.50 at 4)
            nz, ny, nx = shot.shape
            tiles = -(-nx // 32) * -(-ny // 8)
            zbreak = 0 if kernel == "direct" else max(16, min(nz, -(-nz * tiles // (256 * 8))))
        self.kernel = kernel
        self.zbreak = int(zbreak)
        if hist not in ("velocity", "strain"):
            raise ValueError("hist must be 'velocity' or 'strain'")
        self.cpml = shot.cpml is not None
        if self.cpml:
            # the memory variables make the strains of a step depend on more
            # than v, so the gradient reads the strains the forward stored
            if gradient and hist != "strain":
                raise ValueError("the gradient of a CPML shot needs hist='strain'")
        self.hist = hist
        self.hist_fields = 6 if hist == "strain" else 3
        self.lv = [torch.zeros(shot.shape, dtype=real, device=dev) for _ in range(3)]
        self.ls = [torch.zeros(shot.shape, dtype=real, device=dev) for _ in range(6)]
        self.work = torch.zeros((6,) + shot.shape, dtype=real, device=dev)
        self.wface = (torch.zeros((3,) + shot.shape, dtype=real, device=dev)
                      if kernel == "tiled" else None)
        self.proj = None
        if n_proj:
            p = np.asarray(proj, float).reshape(n_mat, n_proj, 36)
            self.proj = torch.as_tensor(np.ascontiguousarray(p, shot.np_real), device=dev)
        mats = shot.mats.cpu().numpy()
        if mats.max() >= n_mat:
            raise ValueError("material index beyond the table")
        run_ptr, mat_ptr, mat_runs, _ = run_tables(mats, self.zbreak)
        mat_ptr = np.concatenate([mat_ptr, np.full(n_mat + 1 - len(mat_ptr), mat_ptr[-1],
                                                   np.int32)])
        self.n_runs = int(run_ptr[-1])
        self.run_ptr = torch.as_tensor(run_ptr, device=dev)
        self.mat_ptr = torch.as_tensor(mat_ptr, device=dev)
        self.mat_runs = torch.as_tensor(mat_runs, device=dev)
        self.run_acc = (torch.zeros((self.n_runs, self.K), dtype=torch.float64, device=dev)
                        if gradient else None)
        # CPML: adjoints of the memory variables (laid out like shot.psi) and
        # the band cells' derivative adjoints
        self.psib = [torch.zeros_like(p) for p in shot.psi]
        self.cpml_work = [torch.zeros((3, p.shape[1]), dtype=real, device=dev) for p in shot.psi]
        d = _abi.PeAdjoint()
        for a in range(len(self.psib)):
            d.cpml_psib[a] = self.psib[a].data_ptr()
            d.cpml_work[a] = self.cpml_work[a].data_ptr()
        for k in range(3):
            d.lv[k] = self.lv[k].data_ptr()
        for k in range(6):
            d.ls[k] = self.ls[k].data_ptr()
        d.work = self.work.data_ptr()
        d.wface = None if self.wface is None else self.wface.data_ptr()
        d.kernel = _abi.KERNELS[kernel]
        d.hist_kind = _abi.PE_HIST_STRAIN if hist == "strain" else _abi.PE_HIST_VELOCITY
        d.n_proj = int(n_proj)
        d.proj = None if self.proj is None else self.proj.data_ptr()
        d.zbreak = self.zbreak
        d.n_runs = self.n_runs
        d.run_ptr = self.run_ptr.data_ptr()
        d.mat_ptr = self.mat_ptr.data_ptr()
        d.mat_runs = self.mat_runs.data_ptr()
        d.run_acc = None if self.run_acc is None else self.run_acc.data_ptr()
        self.desc = d
        self.lib = shot.lib
        _abi.check(self.lib.pe_check_adjoint(C.byref(shot.desc), C.byref(d)), "pe_check_adjoint")

    def reset(self):
        _abi.check(self.lib.pe_adjoint_reset(C.byref(self.shot.desc), C.byref(self.desc),
                                             self.shot._stream()), "pe_adjoint_reset")

    def run(self, n_hi, n_lo, vhist=None, residual=None, grad_wavelet=None):
        """Reverse steps n_hi-1 .. n_lo (asynchronous)."""
        shot = self.shot
        if not (0 <= int(n_lo) <= int(n_hi)):
            raise ValueError(f"bad reverse step range [{n_lo}, {n_hi})")
        if residual is not None and (residual.dim() != 2 or residual.shape[0] < int(n_hi)
                                     or residual.shape[1] != shot.n_rec
                                     or residual.dtype != shot.real
                                     or not residual.is_contiguous()):
            raise ValueError("residual must be a contiguous (nt >= n_hi, n_rec) tensor of the "
                             "shot's real type")
        if grad_wavelet is not None and (grad_wavelet.numel() < int(n_hi)
                                         or grad_wavelet.dtype != self.torch.float64):
            raise ValueError("grad_wavelet must be a float64 tensor of at least n_hi samples")
        if vhist is not None and self.run_acc is not None and (
                vhist.shape[0] < int(n_hi) - int(n_lo) or vhist.dtype != shot.real
                or vhist.shape[1] != self.hist_fields or not vhist.is_contiguous()):
            raise ValueError("history must hold %d slots of %d fields"
                             % (int(n_hi) - int(n_lo), self.hist_fields))
        _abi.check(self.lib.pe_adjoint_run(
            C.byref(self.shot.desc), C.byref(self.desc), int(n_hi), int(n_lo),
            None if vhist is None else vhist.data_ptr(),
            None if residual is None else residual.data_ptr(),
            None if grad_wavelet is None else grad_wavelet.data_ptr(),
            self.shot._stream()), "pe_adjoint_run")

    def gradient(self):
        """(n_mat, K) float64 device tensor of the accumulated gradient."""
        out = self.torch.empty((int(self.shot.rho_t.numel()), self.K), dtype=self.torch.float64,
                               device=self.shot.device)
        _abi.check(self.lib.pe_adjoint_gradient(C.byref(self.shot.desc), C.byref(self.desc),
                                                out.data_ptr(), self.shot._stream()),
                   "pe_adjoint_gradient")
        return out


def _segments(nt, every):
    every = max(1, int(every))
    return [(a, min(a + every, nt)) for a in range(0, nt, every)]


def forward_checkpointed(shot, wav, nt, traces, checkpoint_every=200, history=None,
                         hist_fields=3):
    """Forward run from zero fields.  Returns (checkpoints, stored) where
    checkpoints[a] is the state at the start of segment [a, b) -- (v, s), and
    with a CPML the memory variables too -- and stored[a] the (b-a,
    hist_fields, nz, ny, nx) history of a segment kept during this pass
    (segments listed in ``history``): v (3 fields) or the strains of v (6
    fields)."""
    torch = shot.torch
    shot.reset()
    ckpt, stored = {}, {}
    history = history or ()
    for a, b in _segments(nt, checkpoint_every):
        if a not in history:
            ckpt[a] = (torch.stack(shot.v + shot.s), [p.clone() for p in shot.psi])
            shot.run(wav, a, b, traces)
        else:
            h = torch.empty((b - a, hist_fields) + shot.shape, dtype=shot.real,
                            device=shot.device)
            _run_storing(shot, wav, a, b, traces, h)
            stored[a] = h
    return ckpt, stored


def _run_storing(shot, wav, a, b, traces, hist):
    shot._check_run(wav, a, b, traces)
    if hist.shape[0] < b - a or hist.shape[2:] != shot.shape or hist.dtype != shot.real \
            or not hist.is_contiguous():
        raise ValueError("history buffer too small for steps [%d, %d)" % (a, b))
    fn = shot.lib.pe_run_store_strains if hist.shape[1] == 6 else shot.lib.pe_run_store
    _abi.check(fn(C.byref(shot.desc), int(a), int(b), wav.data_ptr(),
                  None if traces is None else traces.data_ptr(), hist.data_ptr(),
                  shot._stream()), "pe_run_store")


def _restore(shot, state):
    fields, psi = state
    for k in range(3):
        shot.v[k].copy_(fields[k])
    for k in range(6):
        shot.s[k].copy_(fields[3 + k])
    for p, q in zip(shot.psi, psi):
        p.copy_(q)


def reverse_pass(shot, adj, wav, nt, residual, ckpt, stored, checkpoint_every=200,
                 grad_wavelet=None):
    """Reverse steps nt-1 .. 0 over the checkpointed segments."""
    torch = shot.torch
    adj.reset()
    hist = None
    for a, b in reversed(_segments(nt, checkpoint_every)):
        if a in stored:
            h = stored.pop(a)
        else:
            if hist is None or hist.shape[0] < b - a:
                hist = torch.empty((b - a, adj.hist_fields) + shot.shape, dtype=shot.real,
                                   device=shot.device)
            _restore(shot, ckpt[a])
            _run_storing(shot, wav, a, b, None, hist)
            h = hist
        adj.run(b, a, h, residual, grad_wavelet)


def history_segments(nt, checkpoint_every, shape, real_bytes, budget_bytes, fields=3):
    """Start steps of the last segments whose history fits the budget."""
    per_step = fields * int(np.prod(shape)) * real_bytes
    keep, used = [], 0
    for a, b in reversed(_segments(nt, checkpoint_every)):
        need = (b - a) * per_step
        if used + need > budget_bytes:
            break
        keep.append(a)
        used += need
    return keep


def choose_history(shot, checkpoint_every, budget_fraction=0.3):
    """'strain' (6 fields per step, no strain recompute in the reverse pass)
    when a segment of it fits in budget_fraction of the free device memory,
    else 'velocity' (3 fields).  A CPML shot always takes 'strain' (its
    strains depend on the memory variables as well as on v)."""
    torch = shot.torch
    if shot.cpml is not None:
        return "strain"
    real_bytes = 4 if shot.real == torch.float32 else 8
    need = int(checkpoint_every) * 6 * int(np.prod(shot.shape)) * real_bytes
    free = torch.cuda.mem_get_info(shot.device)[0]
    return "strain" if need <= budget_fraction * free else "velocity"


def misfit_gradient(shot, wavelet, nt, residual_fn, checkpoint_every=200, n_proj=0, proj=None,
                    history_bytes=0, with_wavelet=False, adjoint=None, hist=None):
    """Misfit and its gradient for one DeviceShot.

    ``residual_fn(traces)`` receives the (nt, n_rec) device traces and returns
    (J, dJ/dtraces) with the second a device tensor of the traces' dtype.
    ``hist``: 'strain' or 'velocity' history (default: ``choose_history``;
    an explicit ``adjoint`` brings its own).  Returns (J, grad (n_mat, K)
    float64 device tensor, traces[, dJ/dwavelet]).
    """
    torch = shot.torch
    wav = shot.wavelet(wavelet) if not torch.is_tensor(wavelet) else wavelet
    traces = shot.new_traces(nt)
    if adjoint is None:
        hist = hist or choose_history(shot, checkpoint_every)
        adjoint = DeviceAdjoint(shot, n_proj=n_proj, proj=proj, hist=hist)
    adj = adjoint
    real_bytes = 4 if shot.real == torch.float32 else 8
    keep = history_segments(nt, checkpoint_every, shot.shape, real_bytes, history_bytes,
                            adj.hist_fields)
    ckpt, stored = forward_checkpointed(shot, wav, nt, traces, checkpoint_every, keep,
                                        adj.hist_fields)
    J, res = residual_fn(traces)
    res = res.to(shot.real).contiguous()
    gw = torch.zeros(nt, dtype=torch.float64, device=shot.device) if with_wavelet else None
    reverse_pass(shot, adj, wav, nt, res, ckpt, stored, checkpoint_every, gw)
    grad = adj.gradient()
    out = (J, grad, traces)
    return out + (gw,) if with_wavelet else out


def _autograd_function():
    import torch

    class _Traces(torch.autograd.Function):
        """traces = F(Ct); backward = exact discrete adjoint (pe_adjoint_run)."""

        @staticmethod
        def forward(ctx, Ct, shot, wav, nt, checkpoint_every, history_bytes):
            table = Ct.detach().reshape(shot.Ct.shape).to(device=shot.device,
                                                          dtype=shot.real).clone()
            shot.Ct.copy_(table)
            traces = shot.new_traces(nt)
            real_bytes = 4 if shot.real == torch.float32 else 8
            hist = choose_history(shot, checkpoint_every)
            fields = 6 if hist == "strain" else 3
            keep = history_segments(nt, checkpoint_every, shot.shape, real_bytes, history_bytes,
                                    fields)
            ckpt, stored = forward_checkpointed(shot, wav, nt, traces, checkpoint_every, keep,
                                                fields)
            ctx.hist = hist
            # the shot's table may be overwritten by another forward before this
            # backward runs (a line search): the reverse pass and the segment
            # recomputes must see the table of this forward
            ctx.table = table
            ctx.state = (shot, wav, nt, checkpoint_every, ckpt, stored)
            ctx.ct_shape, ctx.ct_dtype, ctx.ct_device = Ct.shape, Ct.dtype, Ct.device
            return traces

        @staticmethod
        def backward(ctx, grad_traces):
            if ctx.state is None:
                raise RuntimeError("backward through the same traces twice is not supported")
            shot, wav, nt, every, ckpt, stored = ctx.state
            shot.Ct.copy_(ctx.table)
            adj = DeviceAdjoint(shot, hist=ctx.hist)
            res = grad_traces.to(shot.real).contiguous()
            reverse_pass(shot, adj, wav, nt, res, ckpt, stored, every)
            g = adj.gradient()
            ctx.state = None
            ctx.table = None
            return (g.to(ctx.ct_dtype).to(ctx.ct_device).reshape(ctx.ct_shape),
                    None, None, None, None, None)

    return _Traces


_TRACES = None


def traces_from_table(Ct, shot, wavelet, nt, checkpoint_every=200, history_bytes=0):
    """Differentiable traces of ``shot`` as a function of its stiffness table.

    ``Ct`` is a (n_mat, 36) torch tensor (any float dtype, any device); the
    forward copies it into the shot's table and runs the GPU solver; the
    backward runs the checkpointed adjoint and returns dJ/dCt (float64 sums).
    """
    global _TRACES
    if _TRACES is None:
        _TRACES = _autograd_function()
    torch = shot.torch
    wav = wavelet if torch.is_tensor(wavelet) else shot.wavelet(wavelet)
    return _TRACES.apply(Ct, shot, wav, int(nt), int(checkpoint_every), int(history_bytes))
