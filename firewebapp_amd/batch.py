"""Deck-driven batched sweeps: the ``batch.*`` deck section -> the batch
table -> the cases run on one GPU, several at a time.

Deck syntax (values of ``batch.<key>``; ``<key>`` is any ``outdoor.*`` key
of the deck, or ``wind_speed_m_s``, the run argument of run_3d_spread):

* ``v1, v2, ...``   a list: case i takes entry i mod len (``cycle`` mode) or
                    one factor of the Cartesian product (``product`` mode);
* ``lo:hi:step``    a grid, the list lo, lo + step, ... up to hi inclusive;
* ``lo~hi``         a uniform draw per case from PCG64(batch.seed) in the
                    order of the table's keys;
* anything else     one value for every case.

Control keys: ``batch.n_cases`` (cycle mode; product mode takes the size of
the product), ``batch.mode`` (``cycle`` or ``product``), ``batch.seed``,
``batch.concurrent`` (cases in flight at once).

``run_batch`` runs the table with one solver per case.  Each case in flight
runs on its own HIP stream from its own host thread: the native step holds
no interpreter lock while it issues kernels, so the small cases of a sweep
(a few tens of thousands of cells, launch-bound on a 256-CU part) overlap
their kernels instead of queueing behind each other.  Every case computes
exactly what it computes alone: the kernels are deterministic and share no
state.
"""
from __future__ import annotations

import copy
import itertools
import threading
import time
from dataclasses import dataclass, field
from typing import Any, Dict, List

import numpy as np

from .deck import RomInputs, load_text_input

_CONTROL = ("n_cases", "mode", "seed", "concurrent")


@dataclass
class BatchCase:
    index: int
    wind_speed_m_s: float
    overrides: Dict[str, Any] = field(default_factory=dict)

    def inputs(self, base: RomInputs) -> RomInputs:
        """The case's deck: the base deck with its outdoor overrides applied."""
        ri = copy.deepcopy(base)
        ri.batch_overrides = {}
        ri.outdoor_overrides.update(self.overrides)
        return ri


def _num(text):
    t = text.strip()
    try:
        v = float(t)
    except ValueError:
        return t
    return int(v) if v == int(v) and "." not in t and "e" not in t.lower() else v


def _parse(value: str):
    """('list', [...]) | ('range', (lo, hi)) | ('fixed', v)."""
    v = value.strip()
    if "~" in v:
        lo, hi = (float(x) for x in v.split("~", 1))
        return "range", (lo, hi)
    if ":" in v:
        parts = [float(x) for x in v.split(":")]
        if len(parts) != 3 or parts[2] <= 0.0:
            raise ValueError(f"grid {value!r} must be lo:hi:step with step > 0")
        lo, hi, st = parts
        n = int(np.floor((hi - lo) / st + 1e-9)) + 1
        return "list", [lo + k * st for k in range(n)]
    if "," in v:
        return "list", [_num(x) for x in v.split(",") if x.strip()]
    return "fixed", _num(v)


def batch_table(ri: RomInputs) -> List[BatchCase]:
    """Expand the deck's batch section into its cases (empty without one)."""
    b = dict(ri.batch_overrides)
    if not b:
        return []
    mode = str(b.get("mode", "cycle")).lower()
    if mode not in ("cycle", "product"):
        raise ValueError(f"batch.mode must be cycle or product (got {mode!r})")
    seed = int(float(b.get("seed", 0)))
    keys = sorted(k for k in b if k not in _CONTROL)
    spec = {k: _parse(b[k]) for k in keys}
    lists = [k for k in keys if spec[k][0] == "list"]
    if mode == "product":
        combos = list(itertools.product(*[spec[k][1] for k in lists]))
        n = len(combos)
    else:
        n = int(float(b.get("n_cases", 0)))
        if n <= 0:
            raise ValueError("batch.n_cases must be a positive integer in cycle mode")
        combos = [tuple(spec[k][1][i % len(spec[k][1])] for k in lists) for i in range(n)]
    rng = np.random.default_rng(seed)
    ranges = [k for k in keys if spec[k][0] == "range"]
    draws = {k: rng.uniform(*spec[k][1], n) for k in ranges}
    deck_wind = ri.outdoor_overrides.get("wind_speed_m_s", 0.0)
    out = []
    for i in range(n):
        ov: Dict[str, Any] = {}
        for k in keys:
            kind, val = spec[k]
            if kind == "fixed":
                ov[k] = val
            elif kind == "range":
                ov[k] = float(draws[k][i])
        for k, v in zip(lists, combos[i]):
            ov[k] = v
        wind = float(ov.pop("wind_speed_m_s", deck_wind or 0.0))
        out.append(BatchCase(index=i, wind_speed_m_s=wind, overrides=ov))
    return out


def _run_case(base: RomInputs, case: BatchCase, snapshot_dir=None, tune=False):
    import torch
    from .spread_3d import Spread3DConfig, Spread3DSolver
    cfg = Spread3DConfig.resolve(case.inputs(base), wind_speed_m_s=case.wind_speed_m_s)
    sv = Spread3DSolver(cfg)
    try:
        if tune:
            # the launch shapes of the case's grid (timed once per shape per
            # process and cached; firewebapp_amd.autotune)
            from .autotune import tune as _tune
            _tune(sv)
        torch.cuda.current_stream().synchronize()
        t0 = time.perf_counter()
        sv.run(snapshot_dir=snapshot_dir)
        torch.cuda.current_stream().synchronize()
        wall = time.perf_counter() - t0
        g = sv.grid
        cells = g.Nx * g.Ny * g.Nz
        return dict(index=case.index, Nx=g.Nx, Ny=g.Ny, Nz=g.Nz, cells=cells,
                    wind_speed_m_s=case.wind_speed_m_s, overrides=dict(case.overrides),
                    steps=sv.step_no, t_end=sv.t, wall_s=wall,
                    cell_steps_per_s=cells * sv.step_no / wall if wall > 0 else 0.0,
                    ros_m_s=sv.ros(), front_x=sv.front_x[-1], stop_reason=sv.stop_reason,
                    result=sv.result())
    finally:
        sv.close()


def run_batch(deck, concurrent=None, snapshot_root=None, on_done=None, tune=False):
    """Run every case of the deck's batch table; returns one record per case
    (table order) with its size, steps, wall time, cell-steps/s, ROS, stop
    reason and Spread3DResult.  ``concurrent`` cases run at once (default
    ``batch.concurrent``, else 4), each on its own stream; the largest cases
    start first so the tail of the sweep is short.  tune: pick each grid
    shape's launch shapes with the autotuner before its first case runs."""
    import torch
    base = load_text_input(deck) if not isinstance(deck, RomInputs) else deck
    cases = batch_table(base)
    if not cases:
        raise ValueError("the deck has no batch.* section")
    if concurrent is None:
        concurrent = int(float(base.batch_overrides.get("concurrent", 4)))
    concurrent = max(1, int(concurrent))

    def size(c):
        ov = c.inputs(base).outdoor_overrides
        return float(ov.get("lx", 1.0)) * float(ov.get("ly", 1.0)) * float(ov.get("lz", 1.0))

    order = sorted(cases, key=size, reverse=True)
    records: Dict[int, dict] = {}
    errors: List[BaseException] = []
    lock = threading.Lock()
    it = iter(order)
    dev = torch.cuda.current_device()

    def worker():
        torch.cuda.set_device(dev)
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            while True:
                with lock:
                    if errors:
                        return
                    c = next(it, None)
                if c is None:
                    return
                try:
                    sd = None
                    if snapshot_root is not None:
                        from pathlib import Path
                        sd = Path(snapshot_root) / f"case_{c.index:04d}"
                    rec = _run_case(base, c, sd, tune)
                except BaseException as e:    # re-raised on the caller's thread
                    with lock:
                        errors.append(e)
                    return
                with lock:
                    records[c.index] = rec
                if on_done is not None:
                    on_done(rec)

    threads = [threading.Thread(target=worker, name=f"fw-batch-{k}")
               for k in range(min(concurrent, len(cases)))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return [records[c.index] for c in cases]


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="run the batch table of a deck")
    ap.add_argument("deck")
    ap.add_argument("--concurrent", type=int, default=None)
    ap.add_argument("--cases", type=int, default=None, help="run only the first N cases")
    ap.add_argument("--steps", type=int, default=None, help="override batch n_steps")
    ap.add_argument("--out", default=None, help="write the table as JSON")
    ap.add_argument("--tune", action="store_true", help="autotune each grid shape first")
    a = ap.parse_args(argv)
    base = load_text_input(a.deck)
    if a.steps is not None:
        base.batch_overrides["n_steps"] = str(a.steps)
    if a.cases is not None:
        base.batch_overrides["n_cases"] = str(a.cases)
    t0 = time.perf_counter()

    def done(r):
        print(f"# case {r['index']:3d}: {r['Nx']}x{r['Ny']}x{r['Nz']} U10 {r['wind_speed_m_s']:.1f}"
              f" {r['steps']} steps {r['wall_s']:.2f} s {r['cell_steps_per_s']:.3e} cell-steps/s"
              f" ROS {r['ros_m_s']:.3f} m/s ({r['stop_reason']})", flush=True)

    recs = run_batch(base, concurrent=a.concurrent, on_done=done, tune=a.tune)
    wall = time.perf_counter() - t0
    total = sum(r["cells"] * r["steps"] for r in recs)
    summary = {"cases": len(recs), "cell_steps": total, "wall_s": wall,
               "cell_steps_per_s": total / wall}
    print(json.dumps(summary), flush=True)
    if a.out:
        rows = [{k: v for k, v in r.items() if k != "result"} for r in recs]
        with open(a.out, "w") as f:
            json.dump({"summary": summary, "cases": rows}, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
