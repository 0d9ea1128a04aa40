"""GPU: the graph-replayed step loops against direct launches.

The sub-step loop of the native step is recorded once per distinct launch
sequence (trip counts, flags, field pointers) and replayed from a hipGraph,
with the per-step scalars read from a device block.  Replays must be bit for
bit the direct launches: every field, the particle state and the step
histories, over runs long enough to cross the ignition window (the drip
source switches off) and the DOM and reinitialisation steps, on the default
stream (the solver's own stream then) and on a caller's stream.
"""
import contextlib

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "rho", "k", "eps", "nu_t",
          "omega", "q_rad", "q_rad_gas", "S_pyro", "Q_comb", "phi")


def _run(name, dtype, n_steps, graphs, stream=None, **kw):
    import torch
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    cfg = Spread3DConfig.resolve(bench.deck(name, dtype), wind_speed_m_s=bench.wind(name), **kw)
    sv = Spread3DSolver(cfg, graphs=graphs)
    try:
        ctx = torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()
        with ctx:
            for _ in range(n_steps):
                sv.step()
        torch.cuda.synchronize()
        out = sv.state_numpy()
        out["diag"] = {k: list(v) for k, v in sv.diag.items()}
        out["front_x"] = list(sv.front_x)
        out["stats"] = sv.graph_stats()
        return out
    finally:
        sv.close()


def _equal(a, b):
    for k in FIELDS + ("p_x", "p_y", "p_z", "p_T_s", "p_m_solid", "p_m_water", "p_alive"):
        assert np.array_equal(a[k], b[k]), k
    assert a["diag"] == b["diag"]
    assert a["front_x"] == b["front_x"]


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_graph_replay_equals_direct_128x32x32(dtype):
    """60 steps at 128 x 32 x 32 (S4): ignition, the end of the drip window,
    DOM every fifth step, six reinitialisations."""
    direct = _run("spread3d_128x32x32_s4", dtype, 60, graphs=False)
    replay = _run("spread3d_128x32x32_s4", dtype, 60, graphs=True)
    assert direct["stats"] is None
    replays, captures = replay["stats"]
    print(f"{dtype}: {replays} replays, {captures} captures over 60 steps", flush=True)
    assert replays + captures == 60 and captures <= 12
    _equal(direct, replay)


def test_graph_replay_equals_direct_256x64x48_on_a_caller_stream():
    """40 steps of the production workload on a caller's stream."""
    import torch
    direct = _run("spread3d_256x64x48_s8", "float32", 40, graphs=False)
    replay = _run("spread3d_256x64x48_s8", "float32", 40, graphs=True,
                  stream=torch.cuda.Stream())
    replays, captures = replay["stats"]
    print(f"{replays} replays, {captures} captures over 40 steps", flush=True)
    assert replays > captures
    _equal(direct, replay)
