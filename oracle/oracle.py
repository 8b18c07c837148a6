"""CPU ORACLE -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker: the product
path (``pulse_echo_azimuthal_simulation_amd``) never loads it.

ctypes front end to ``liboracle_fdtd.so`` (fdtd_oracle.c), the C restatement
of one step of the upstream label-indexed solver
(sim/core/fdtd.py::forward_fused_labels, lines 1618-1816).

Parity pins (each checked in tests/):

* whole-step semantics, FP32: the upstream float32 solver archives
  sim/results/rsg_green_ssg_{iso,ani}_ppw{4,6}.npz, rsg_lattice_ssg_ppw6.npz
  and rsg_twoway_ssg_ppw6.npz (copied under tests/golden/reference/),
  rel-L2 <= 1e-5 per trace set;
* FP64: against the FP32 path (same semantics) and against closed forms --
  a homogeneous medium stays spatially uniform under a uniform initial
  stress, stress symmetry under mirror of the grid, energy decay with damping;
* a brute-force per-cell Python step on tiny grids (tests/test_oracle_small.py)
  that restates the stencils from their definitions independently of this C.

The CPML boundary (no upstream counterpart) is pinned by its reduction to
the plain scheme when every profile is (b, a, 1/kappa) = (1, 0, 1), by an
independent per-cell Python restatement of the band path on tiny grids, and
by a wide-domain reflection test (tests/test_oracle_cpml.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_fdtd.so")


class _Shot(C.Structure):
    _fields_ = [
        ("nz", C.c_int32), ("ny", C.c_int32), ("nx", C.c_int32),
        ("nc", C.c_int32),
        ("coeffs", C.c_double * 4),
        ("dt", C.c_double), ("inv_h", C.c_double), ("dt_h", C.c_double),
        ("n_mat", C.c_int32),
        ("act_mask", C.c_uint64),
        ("mats", C.c_void_p), ("Ct", C.c_void_p), ("rho_t", C.c_void_p),
        ("damp", C.c_void_p),
        ("v", C.c_void_p * 3), ("s", C.c_void_p * 6), ("g", C.c_void_p),
        ("n_src", C.c_int32), ("src_idx", C.c_void_p), ("src_w", C.c_void_p),
        ("n_rec", C.c_int32), ("rec_ptr", C.c_void_p), ("rec_idx", C.c_void_p),
        ("rec_w", C.c_void_p),
        ("record_mode", C.c_int32),
        ("cpml_width", C.c_int32), ("cpml_unused", C.c_int32),
        ("cpml_prof", C.c_void_p * 3), ("cpml_psi", C.c_void_p * 3),
    ]


def build():
    """Compile liboracle_fdtd.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        lib = C.CDLL(_LIB)
        for suf in ("f32", "f64"):
            getattr(lib, "orc_run_" + suf).argtypes = [C.POINTER(_Shot), C.c_int, C.c_int,
                                                      C.c_void_p, C.c_void_p]
            getattr(lib, "orc_stress_" + suf).argtypes = [C.POINTER(_Shot)]
            getattr(lib, "orc_velocity_" + suf).argtypes = [C.POINTER(_Shot)]
            getattr(lib, "orc_source_" + suf).argtypes = [C.POINTER(_Shot), C.c_double]
            getattr(lib, "orc_record_" + suf).argtypes = [C.POINTER(_Shot), C.c_void_p]
            getattr(lib, "orc_adjoint_step_" + suf).argtypes = [C.POINTER(_Shot)] + [C.c_void_p] * 8
        _lib = lib
    return _lib


RECORD_MODES = {"pressure": 0, "vx": 1, "vy": 2, "vz": 3}


class OracleShot:
    """Host arrays and scalars of one shot, held alive for the C calls.

    mats      int32 (nz, ny, nx) material index (label + 1)
    Ct        (n_mat, 36) stiffness table, rho_t (n_mat,) densities
    damp      (nz, ny, nx) or None
    src       (flat int64 indices, weights); rec = (ptr int32, idx int64, w)
    """

    def __init__(self, mats, Ct, rho_t, damp, coeffs, h, dt, src, rec,
                 record="pressure", precision="f32", act_mask=None, cpml=None):
        self.real = np.float32 if precision == "f32" else np.float64
        self.suf = precision
        r = self.real
        self.mats = np.ascontiguousarray(mats, np.int32)
        self.shape = self.mats.shape
        self.Ct = np.ascontiguousarray(np.asarray(Ct).reshape(-1, 36), r)
        self.rho_t = np.ascontiguousarray(rho_t, r)
        self.damp = None if damp is None else np.ascontiguousarray(damp, r)
        co = np.zeros(4)
        co[:len(coeffs)] = np.asarray(coeffs, float)
        self.v = [np.zeros(self.shape, r) for _ in range(3)]
        self.s = [np.zeros(self.shape, r) for _ in range(6)]
        self.g = np.zeros((6,) + self.shape, r)
        self.src_idx = np.ascontiguousarray(src[0], np.int64)
        self.src_w = np.ascontiguousarray(src[1], r)
        self.rec_ptr = np.ascontiguousarray(rec[0], np.int32)
        self.rec_idx = np.ascontiguousarray(rec[1], np.int64)
        self.rec_w = np.ascontiguousarray(rec[2], r)
        if act_mask is None:
            flags = (np.abs(self.Ct) > 0).any(axis=0)
            act_mask = sum(1 << k for k in range(36) if flags[k])
        nz, ny, nx = self.shape
        st = _Shot()
        st.nz, st.ny, st.nx = nz, ny, nx
        st.nc = len(coeffs)
        for j in range(4):
            st.coeffs[j] = co[j]
        st.dt = float(dt)
        st.inv_h = 1.0 / h
        st.dt_h = dt / h
        st.n_mat = len(self.rho_t)
        st.act_mask = int(act_mask)
        st.mats = self.mats.ctypes.data
        st.Ct = self.Ct.ctypes.data
        st.rho_t = self.rho_t.ctypes.data
        st.damp = None if self.damp is None else self.damp.ctypes.data
        for k in range(3):
            st.v[k] = self.v[k].ctypes.data
        for k in range(6):
            st.s[k] = self.s[k].ctypes.data
        st.g = self.g.ctypes.data
        st.n_src = len(self.src_idx)
        st.src_idx = self.src_idx.ctypes.data
        st.src_w = self.src_w.ctypes.data
        st.n_rec = len(self.rec_ptr) - 1
        st.rec_ptr = self.rec_ptr.ctypes.data
        st.rec_idx = self.rec_idx.ctypes.data
        st.rec_w = self.rec_w.ctypes.data
        # CPML: (width, [prof_z, prof_y, prof_x]) with prof real (2, 3, n);
        # memory variables (6 components, band volume) per axis, advanced in
        # place
        self.psi = []
        self.prof = []
        if cpml is not None:
            width, prof = cpml
            W = 2 * int(width) + 2
            vols = (W * ny * nx, nz * W * nx, nz * ny * W)
            st.cpml_width = int(width)
            for a in range(3):
                self.prof.append(np.ascontiguousarray(prof[a], r))
                self.psi.append(np.zeros((6, vols[a]), r))
                st.cpml_prof[a] = self.prof[a].ctypes.data
                st.cpml_psi[a] = self.psi[a].ctypes.data
        st.record_mode = RECORD_MODES[record]
        self._st = st

    def run(self, wavelet, n0, n1, traces=None):
        """Advance steps n0..n1-1; returns the (n1, n_rec) traces array."""
        lib = _load()
        wav = np.ascontiguousarray(wavelet, np.float64)
        n_rec = self._st.n_rec
        if traces is None:
            traces = np.zeros((n1, n_rec), self.real)
        getattr(lib, "orc_run_" + self.suf)(C.byref(self._st), int(n0), int(n1),
                                            wav.ctypes.data, traces.ctypes.data)
        return traces

    def stress(self):
        getattr(_load(), "orc_stress_" + self.suf)(C.byref(self._st))

    def velocity(self):
        getattr(_load(), "orc_velocity_" + self.suf)(C.byref(self._st))

    def source(self, wn):
        getattr(_load(), "orc_source_" + self.suf)(C.byref(self._st), float(wn))

    def record(self):
        row = np.zeros(self._st.n_rec, self.real)
        getattr(_load(), "orc_record_" + self.suf)(C.byref(self._st), row.ctypes.data)
        return row


def set_adjoint_plant(k):
    """Planted mistakes in the adjoint (0 = none; see fdtd_oracle.c)."""
    C.c_int.in_dll(_load(), "orc_adj_plant").value = int(k)


class OracleAdjoint:
    """Adjoint state (lv, ls) of one OracleShot and the reverse step.

    ``step(vn, res)`` performs reverse step n given v^n (the forward state at
    the start of step n, three (nz, ny, nx) arrays) and the trace adjoint row
    dJ/dy_n (or None); it accumulates dJ/dCt into ``gC`` (n_mat, 36) float64
    and returns dJ/dwavelet[n].  With a CPML the adjoint state also holds
    ``psib``, the adjoints of the memory variables (laid out like
    ``shot.psi``), and the gradient needs ``gn``, the strains of step n
    (6, nz, ny, nx) as the forward computed them (they depend on the forward
    memory variables).
    """

    def __init__(self, shot):
        self.shot = shot
        r = shot.real
        self.lv = [np.zeros(shot.shape, r) for _ in range(3)]
        self.ls = np.zeros((6,) + shot.shape, r)
        self.gC = np.zeros((len(shot.rho_t), 36))
        self.gbar = np.zeros((6,) + shot.shape, r)
        self._vn = [np.zeros(shot.shape, r) for _ in range(3)]
        self._gn = np.zeros((6,) + shot.shape, r)
        self.psib = [np.zeros_like(p) for p in shot.psi]

    def step(self, vn, res=None, grad=True, gn=None):
        sh = self.shot
        for k in range(3):
            self._vn[k][...] = vn[k]
        st = _Shot.from_buffer_copy(sh._st)
        for k in range(3):
            st.v[k] = self._vn[k].ctypes.data
        if self.psib:
            for a in range(3):
                st.cpml_psi[a] = self.psib[a].ctypes.data
            if grad:
                if gn is None:
                    raise ValueError("a CPML shot's gradient needs the strains of step n (gn)")
                self._gn[...] = gn
                st.g = self._gn.ctypes.data
        gw = C.c_double(0.0)
        resp = None
        if res is not None:
            res = np.ascontiguousarray(res, sh.real)
            resp = res.ctypes.data
        getattr(_load(), "orc_adjoint_step_" + sh.suf)(
            C.byref(st), self.lv[0].ctypes.data, self.lv[1].ctypes.data, self.lv[2].ctypes.data,
            self.ls.ctypes.data, resp, self.gC.ctypes.data if grad else None,
            C.addressof(gw), self.gbar.ctypes.data)
        return gw.value


def forward_history(shot, wavelet, nt, strains=False):
    """Run nt steps from zero fields; returns (traces, v history (nt, 3, ...))
    where history[n] is v at the start of step n, and with ``strains`` also
    the strains each step computed, (nt, 6, ...)."""
    hist = np.zeros((nt, 3) + shot.shape, shot.real)
    ghist = np.zeros((nt, 6) + shot.shape, shot.real) if strains else None
    traces = np.zeros((nt, shot._st.n_rec), shot.real)
    for n in range(nt):
        for k in range(3):
            hist[n, k] = shot.v[k]
        shot.run(wavelet, n, n + 1, traces)
        if strains:
            ghist[n] = shot.g
    if strains:
        return traces, hist, ghist
    return traces, hist


def adjoint_gradient(shot, wavelet, nt, residual_fn):
    """Misfit gradient by forward (storing v, and with a CPML the strains)
    then the reverse loop.

    ``residual_fn(traces)`` returns (J, dJ/dtraces).  Returns (J, gC (n_mat,
    36), gw (nt,), traces)."""
    cp = bool(shot.psi)
    out = forward_history(shot, wavelet, nt, strains=cp)
    traces, hist = out[0], out[1]
    J, res = residual_fn(traces)
    adj = OracleAdjoint(shot)
    gw = np.zeros(nt)
    for n in range(nt - 1, -1, -1):
        gw[n] = adj.step(hist[n], res[n], gn=out[2][n] if cp else None)
    return J, adj.gC, gw, traces


def run_forward(mats, Ct, rho_t, damp, coeffs, h, dt, nt, src, rec, wavelet,
                record="pressure", precision="f32"):
    """Whole-shot convenience: fresh fields, nt steps, (nt, n_rec) traces."""
    sh = OracleShot(mats, Ct, rho_t, damp, coeffs, h, dt, src, rec, record, precision)
    return sh.run(wavelet, 0, int(nt))
