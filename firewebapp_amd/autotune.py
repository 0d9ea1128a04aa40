"""Autotuner of the step's launch shapes (no counterpart in the reference,
whose step is interpreted Python).

For a solver's grid it times the candidates of each tunable launch shape on
the solver's current state and applies the fastest (the default unless
another is faster by more than MARGIN, 5 %):

* the projection's inner PCG (``fw_proj_set_tiling``): the one-pass y-z
  preconditioner kernel or the column transforms, the operator's workgroup
  count and the x-transform tile height, timed as solves of the step's own
  projection system to a 1e-8 residual from x0 = 0 on a separate handle;
* the DOM sweep's tile side (``fw_dom_set_tiling``: 2, 4 or 8 rows square),
  timed per source iteration of solves on the step's fields with a separate
  handle (its own intensities, so the run's warm start is untouched);
* the cells per thread of the species kernel (``fw_tune_species_cz``: 1, 2
  or 4), timed on the step's fields into scratch outputs.

The DOM tile side and the species CZ leave every value bit-identical; the
projection tiling changes the float32 preconditioner's rounding only (every
solve still meets its residual tolerance).  Results are kept per device,
shape, dtype and quadrature in a JSON file (``$FIREWAVE_TUNE_CACHE``,
default ~/.cache/firewave/autotune.json), so a shape is timed once.

    from firewebapp_amd.autotune import tune
    choice = tune(solver)          # {"proj": [1, 1024, 16], "dom_tile": 4, ...}
"""
from __future__ import annotations

import ctypes as C
import json
import os
import threading
import time
from pathlib import Path
from typing import Optional

import numpy as np
import torch

from . import _lib

# (yz_plane, matvec_blocks, dct_rows); the first of each list is the default
PROJ_CANDIDATES = ((1, 1024, 16), (1, 512, 16), (1, 2048, 16), (1, 1024, 8), (0, 512, 16))
DOM_CANDIDATES = (4, 2, 8)
CZ_CANDIDATES = (2, 1, 4)

# a candidate replaces the default only when it is faster by more than this
# fraction, so shapes within the timing noise of each other do not flip from
# run to run
MARGIN = 0.05

_lock = threading.Lock()
_memo: dict = {}     # in-process results by key (batch cases of one shape)


def cache_path() -> Path:
    p = os.environ.get("FIREWAVE_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "firewave" / "autotune.json"


def shape_key(sv) -> str:
    g = sv.grid
    dev = torch.cuda.get_device_name(sv.dev)
    return f"{dev}|{g.Nx}x{g.Ny}x{g.Nz}|{str(sv.dtype).replace('torch.', '')}|M{sv.dom.M}"


def _load(path: Path) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path: Path, data: dict):
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass    # a read-only home: the result still applies to this process


def _timed(fn, reps):
    """Median wall time in ms of fn() over reps calls (stream-synchronised)."""
    ms = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ms.append(1e3 * (time.perf_counter() - t0))
    return float(np.median(ms))


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _tune_projection(sv, reps):
    from .physics_3d.fft_poisson_3d import create_handle
    b = torch.empty(sv.shape, dtype=torch.float64, device=sv.dev)
    _lib.call("fw_proj_rhs", sv.proj, C.c_void_p(b.data_ptr()), _stream())
    if not torch.any(b != 0):
        b.normal_(generator=torch.Generator(device=sv.dev).manual_seed(7))
    rho = sv.f["rho"].clone()
    h = create_handle(sv.dg, 1.0e-6, 1.0e-8, maxiter=5000)
    out = {}
    try:
        for cand in PROJ_CANDIDATES:
            act = C.c_int()
            _lib.call("fw_proj_set_tiling", h, *cand, C.byref(act))
            if cand[0] and not act.value:
                continue          # the one-pass kernel does not cover this grid
            x = torch.zeros_like(b)
            it, rel = C.c_int(), C.c_double()

            def solve():
                x.zero_()
                _lib.call("fw_proj_solve", h, C.c_void_p(rho.data_ptr()),
                          C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), 1.0e-8,
                          C.byref(it), C.byref(rel), _stream())
            solve()   # first call: warm-up of the launch shapes
            out[cand] = _timed(solve, reps)
    finally:
        _lib.lib().fw_proj_destroy(h)
    return out


def _tune_dom(sv, reps):
    from .physics_3d import dom_3d
    g, f = sv.grid, sv.f
    d = dom_3d.DOMRadiationSolver(g.Nz, g.Ny, g.Nx, g.dy, g.dx, g.dz_arr, g.d_face_above,
                                  g.d_face_below, "periodic", sv.cfg.dom_quadrature)
    h = d._handle(sv.dtype)
    qs, qg = torch.empty_like(f["q_rad"]), torch.empty_like(f["q_rad_gas"])
    qsoil = torch.empty_like(sv.q_in_soil)
    mloc = None if sv.eu is not None else f["M_local"]
    p = sv.params
    out = {}
    for tile in DOM_CANDIDATES:
        act = C.c_int()
        _lib.call("fw_dom_set_tiling", h, tile, C.byref(act))
        if act.value != tile:
            continue
        iters = C.c_int()
        total = [0]

        def solve():
            _lib.call("fw_dom_solve", h, *(C.c_void_p(f[k].data_ptr()) for k in
                                           ("T_s", "T_g", "alpha_s", "omega")),
                      float(p.sigma_sav), float(p.T_amb), C.c_void_p(qs.data_ptr()),
                      C.c_void_p(qg.data_ptr()), C.c_void_p(sv.T_soil.data_ptr()),
                      C.c_void_p(qsoil.data_ptr()), C.c_void_p(f["Y_H2O"].data_ptr()),
                      C.c_void_p(f["rho"].data_ptr()),
                      C.c_void_p(mloc.data_ptr()) if mloc is not None else None, _stream())
            _lib.call("fw_dom_stats", h, C.byref(iters))
            total[0] += iters.value
        solve()
        total[0] = 0
        ms = _timed(solve, reps)
        out[tile] = ms * reps / max(total[0], 1)     # per source iteration
    torch.cuda.synchronize()
    del d          # releases the tuning handle
    return out


def _tune_species(sv, reps):
    f = sv.f
    Y = (C.c_void_p * 3)(*(f[k].data_ptr() for k in ("Y_fuel", "Y_O2", "Y_H2O")))
    scratch = [torch.empty_like(f["Y_fuel"]) for _ in range(3)]
    Yo = (C.c_void_p * 3)(*(t.data_ptr() for t in scratch))
    y_in = (C.c_double * 3)(0.0, 0.232, 0.0)
    y_max = (C.c_double * 3)(1.0, 0.232, 1.0)
    out = {}
    for cz in CZ_CANDIDATES:
        _lib.call("fw_tune_species_cz", sv.dg.ref, cz)

        def run():
            for _ in range(10):
                _lib.call("fw_species_transport3", sv.dg.ref, Y, Yo,
                          *(C.c_void_p(f[k].data_ptr()) for k in ("rho", "u", "v", "w", "Sp",
                                                                    "Sd")),
                          1.0e-4, 1.0e-5, y_in, y_max, _stream())
        run()
        out[cz] = _timed(run, reps) / 10
    return out


def _pick(times: dict, default):
    """The fastest candidate, or the default when none beats it by MARGIN."""
    best = min(times, key=times.get)
    if default in times and times[best] > (1.0 - MARGIN) * times[default]:
        return default
    return best


def apply(sv, choice: dict):
    """Set a tuning result on the solver's handles."""
    _lib.call("fw_proj_set_tiling", sv.proj, *choice["proj"], None)
    _lib.call("fw_dom_set_tiling", sv.dom_h, int(choice["dom_tile"]), None)
    _lib.call("fw_tune_species_cz", sv.dg.ref, int(choice["species_cz"]))


def tune(sv, path: Optional[Path] = None, reps: int = 3, retune: bool = False) -> dict:
    """Pick and apply the launch shapes for sv's grid (timed on its current
    state unless a result for the shape is cached); returns the choice with
    the timings behind it (ms per solve / per source iteration / per launch)."""
    key = shape_key(sv)
    path = path or cache_path()
    with _lock:
        if not retune:
            hit = _memo.get(key) or _load(path).get(key)
            if hit is not None:
                _memo[key] = hit
                apply(sv, hit)
                return dict(hit, cached=True)
        t0 = time.perf_counter()
        proj = _tune_projection(sv, reps)
        dom = _tune_dom(sv, reps)
        cz = _tune_species(sv, reps)
        choice = {
            "proj": list(_pick(proj, PROJ_CANDIDATES[0])),
            "dom_tile": int(_pick(dom, DOM_CANDIDATES[0])),
            "species_cz": int(_pick(cz, CZ_CANDIDATES[0])),
            "ms": {"proj_solve": {"%d/%d/%d" % k: round(v, 4) for k, v in proj.items()},
                   "dom_per_iteration": {str(k): round(v, 4) for k, v in dom.items()},
                   "species3_launch": {str(k): round(v, 5) for k, v in cz.items()}},
            "tuning_s": round(time.perf_counter() - t0, 2),
        }
        _memo[key] = choice
        data = _load(path)
        data[key] = choice
        _save(path, data)
        apply(sv, choice)
        return dict(choice, cached=False)
