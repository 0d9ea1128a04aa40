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
