"""GPU: the whole step on a grid whose rows are whole 256-cell segments
(Nx = 256), where the species turbulent diffusion's last pass leaves the EDC
flag-scan summaries of the next sub-step, against the oracle: the
reference's integration deck stretched to 25.6 m, Ny = 4, float64, both
projections at rtol 1e-12, 10 steps."""
import os

import numpy as np
import pytest

from tests.golden_io import GOLDEN_DIR

pytestmark = pytest.mark.gpu


def test_step_nx256_fp64_vs_oracle():
    from firewebapp_amd.deck import load_text_input
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver, build_grid
    from oracle.step import OracleSpread3D
    ri = load_text_input(os.path.join(GOLDEN_DIR, "integration_case.txt"))
    ri.outdoor_overrides.update(dict(lx=25.6, ly=0.4, dtype="float64", n_steps=10))
    cfg = Spread3DConfig.resolve(ri, wind_speed_m_s=4.0, projection_cg_rtol=1e-12)
    sv = Spread3DSolver(cfg)
    try:
        assert sv.grid.Nx == 256 and sv.grid.Ny == 4
        sv.run()
        o = OracleSpread3D(cfg, build_grid(cfg)).run(cfg.n_steps)
        st = sv.state_numpy()
        assert sv.step_no == o.step_no == 10
        speed = float(np.sqrt(np.linalg.norm(o.u) ** 2 + np.linalg.norm(o.v) ** 2 +
                              np.linalg.norm(o.w) ** 2))
        errs = {}
        for name in ("u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "rho", "k",
                     "eps", "omega", "S_pyro"):
            a_, b_ = st[name], getattr(o, name)
            den = speed if name in ("u", "v", "w") else max(np.linalg.norm(b_), 1e-300)
            errs[name] = float(np.linalg.norm(a_ - b_) / den)
        print("Nx 256 rel L2 vs oracle:", errs, flush=True)
        for name, e in errs.items():
            assert e < 1e-8, (name, e)
        assert float(np.max(o.omega)) > 0.0          # the EDC burned
        np.testing.assert_allclose(sv.diag["t"], o.diag["t"], rtol=1e-12)
    finally:
        sv.close()


def _summary_numpy(Yf, YO, Yw):
    """Segment summaries by their definition: per 256-cell segment of the
    flat index, -1 without a resetting cell (Y_F > 1e-9 and Y_O2 > 1e-9),
    else the wet-bulb flag of the last one (edc.py:223)."""
    rf, ro, rw = Yf.reshape(-1, 256), YO.reshape(-1, 256), Yw.reshape(-1, 256)
    out = np.full(rf.shape[0], -1, np.int32)
    for s in range(rf.shape[0]):
        hit = np.flatnonzero((rf[s] > 1e-9) & (ro[s] > 1e-9))
        if hit.size:
            y = rw[s, hit[-1]]
            out[s] = 1 if (y > 0.0 and max(1.0 - y / 0.18, 0.0) < 0.5) else 0
    return out


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_fused_diffusion_summary_carries_flag_one(dtype_name):
    """The species diffusion pass that writes the EDC segment summaries, on a
    wet field where they must carry flag 1.  Each x row is one 256-cell
    segment; rows cycle through three kinds: no fuel and nu_t = 0 (so no fuel
    diffuses in: summary -1), fuel cells away from the segment start in wet
    gas (Y_H2O above the 0.09 wet-bulb trigger: 1) and in dry gas (0).  No
    Y_H2O lies within 0.01 of the trigger, so neither diffusion nor the
    float32 quotient can flip a flag."""
    import ctypes as C
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, ptr, stream_ptr
    dtype = getattr(torch, dtype_name)
    nz, ny, nx = 12, 8, 256
    rng = np.random.default_rng(91)
    sh = (nz, ny, nx)
    dz = np.full(nz, 0.1)
    g = DeviceGrid(nx, ny, nz, 0.1, 0.1, dz, dz, dz, dtype=dtype)
    kind = (np.arange(nz * ny) % 3).reshape(nz, ny)
    Yf = np.zeros(sh)
    for k in range(nz):
        for j in range(ny):
            if kind[k, j]:
                Yf[k, j, rng.integers(10, 240, size=3)] = 0.05
    YO = np.full(sh, 0.2)
    wet = 0.10 + 0.08 * rng.random(sh)
    dry = 0.02 + 0.06 * rng.random(sh)
    Yw = np.where((kind == 1)[:, :, None], wet, dry)
    nu = np.where((kind == 0)[:, :, None], 0.0, 1e-3) * np.ones(sh)
    dev = lambda a: torch.as_tensor(a, dtype=dtype, device="cuda").contiguous()
    nu_t = dev(nu)
    fin = [dev(a) for a in (Yf, YO, Yw)]
    fout = [torch.empty_like(t) for t in fin]
    summ = torch.full((nz * ny * nx // 256,), 7, dtype=torch.int32, device="cuda")
    pin = (C.c_void_p * 3)(*[t.data_ptr() for t in fin])
    pout = (C.c_void_p * 3)(*[t.data_ptr() for t in fout])
    lo = (C.c_double * 3)(0.0, 0.0, 0.0)
    hi = (C.c_double * 3)(1.0, 0.232, 1.0)
    _lib.call("fw_turbulent_diffusion_pass3_summary", g.ref, pin, pout, ptr(nu_t),
              0.7, 1e-3, lo, hi, ptr(summ), stream_ptr())
    torch.cuda.synchronize()
    outs = [t.double().cpu().numpy() for t in fout]
    want = _summary_numpy(*outs)
    got = summ.cpu().numpy()
    print("summaries -1/0/1:", [(int((want == v).sum())) for v in (-1, 0, 1)], flush=True)
    assert [(int((want == v).sum())) for v in (-1, 0, 1)] == [32, 32, 32]
    np.testing.assert_array_equal(got, want)
    # the pass itself equals the plain three-field pass
    ref = [torch.empty_like(t) for t in fin]
    _lib.call("fw_turbulent_diffusion_pass3", g.ref, pin,
              (C.c_void_p * 3)(*[t.data_ptr() for t in ref]), ptr(nu_t), 0.7, 1e-3,
              lo, hi, stream_ptr())
    torch.cuda.synchronize()
    for a, b in zip(fout, ref):
        assert torch.equal(a, b)
