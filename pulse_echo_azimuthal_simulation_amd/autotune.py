"""Launch-shape autotuner for the tiled kernels.

Each half-step kernel (stress, velocity) comes in several tile shapes
(pe_tile_info) and can split z into chunks of any length.  The best choice
depends on the grid (partial edge tiles, how many workgroups the chip sees)
and on the precision, so it is measured: every (tile, z-chunk) candidate is
timed over a few launches with HIP events on the launch stream, and the
fastest is kept per (grid shape, precision) for the life of the process.

Timing runs on the shot's own fields while they are still zero and before
any source is applied: both half-steps (CPML band pre-passes included) map
zero fields and memory variables to zero, so tuning leaves the shot's state
untouched.
"""
from __future__ import annotations

import ctypes as C

from . import _abi

_CACHE: dict = {}
# z planes per workgroup (0: the library's default, at least 16).  Short
# chunks give small grids enough workgroups (on the 96x96x48 disk + 23-cell
# pads the stress half-step picks 12 planes).  Beyond that the time depends
# on the chunk count, sharply and not monotonically (workgroup rounds per
# launch against the halo planes each chunk re-reads), and an uneven last
# chunk costs: at 400x400x200 the plain stress kernel runs 0.917 ms in
# 25-plane chunks (8 even ones), 0.978 in 24 (9, the last of 8 planes),
# 0.951 in 23 (9, the last of 16) and 0.902 in 29 (7, the last of 26).  So the
# candidates are, besides the default and the short chunks, ceil(nz / k) for
# k = 1 .. 16 chunks: for each count the most even split.
SHORT_ZCHUNKS = (0, 6, 8, 12)


def zchunk_candidates(nz, kmax=16):
    """Z-chunk lengths the autotuner times on an nz-plane grid."""
    c = set(SHORT_ZCHUNKS)
    for k in range(1, kmax + 1):
        c.add(-(-nz // k))
    return tuple(sorted(z for z in c if z <= nz))


def _time_half(shot, half, reps):
    torch = shot.torch
    fn = shot.stress if half == 0 else shot.velocity
    fn()                                     # first launch: code object load
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def tune(shot, reps=5, zchunks=None, verbose=False, keep=3):
    """Pick tile variants and z-chunk lengths for ``shot`` (a DeviceShot with
    zero fields); sets them in its descriptor and returns {half: (variant,
    zchunk, ms)} plus the full timing table.  Two rounds per half-step: every
    tile variant at the library's default chunk, then the ``keep`` fastest
    of them at every chunk of ``zchunks`` (default zchunk_candidates of the
    grid's nz), which costs a few times less than timing every pair."""
    if zchunks is None:
        zchunks = zchunk_candidates(int(shot.shape[0]))
    key = (shot.shape, shot.desc.precision, shot.desc.cpml_width)
    if key in _CACHE:
        best = _CACHE[key]
        for h in (0, 1):
            shot.desc.tile[h], shot.desc.zchunk[h] = best[h][0], best[h][1]
        return best, None
    prec = "f32" if shot.desc.precision == _abi.PE_F32 else "f64"
    table = {}
    best = {}

    def timed(h, v, zc):
        if (h, v, zc) not in table:
            shot.desc.tile[h] = v
            shot.desc.zchunk[h] = zc
            _abi.check(shot.lib.pe_check_shot(C.byref(shot.desc)), "pe_check_shot")
            table[(h, v, zc)] = _time_half(shot, h, reps)
            if verbose:
                print("tune half %d tile %s zchunk %d: %.4f ms"
                      % (h, _abi.tile_variants(prec, h)[v], zc, table[(h, v, zc)]), flush=True)
        ms = table[(h, v, zc)]
        if h not in best or ms < best[h][2]:
            best[h] = (v, zc, ms)
        return ms

    for h in (0, 1):
        nv = len(_abi.tile_variants(prec, h))
        first = sorted(range(nv), key=lambda v: timed(h, v, 0))
        for v in first[:keep]:
            for zc in zchunks:
                timed(h, v, zc)
        shot.desc.tile[h], shot.desc.zchunk[h] = best[h][0], best[h][1]
    _CACHE[key] = best
    return best, table
