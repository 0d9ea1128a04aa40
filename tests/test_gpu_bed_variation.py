"""GPU: the seeded bed and inflow variation through the native step.

The reference's integration deck with the particles on, Ny = 4, float64, both
projections at rtol 1e-12, and a lognormal fuel depth field (bed layers to
twice the mean depth), per-particle load and depth jitter and inlet gusts of
+-15 % at a 0.1 s period, against the oracle step fed the same seeded
inputs after 25 steps: fields to relative L2 1e-8, particle state, the step
times and the front history."""
import os

import numpy as np
import pytest

from tests.golden_io import GOLDEN_DIR

pytestmark = pytest.mark.gpu


def test_bed_variation_step_fp64_vs_oracle():
    from firewebapp_amd.deck import load_text_input
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver, bed_inputs, build_grid
    from oracle.step import OracleSpread3D
    ri = load_text_input(os.path.join(GOLDEN_DIR, "integration_case.txt"))
    ri.outdoor_overrides.update(dict(
        ly=0.4, dtype="float64", n_steps=25, fuel_depth_cv=0.3, fuel_depth_corr_length_m=0.3,
        fuel_depth_seed=4, fuel_depth_max_m=0.2, particle_jitter=0.1, particle_jitter_seed=8,
        wind_gust_amplitude=0.15, wind_gust_period_s=0.1, wind_gust_seed=2))
    cfg = Spread3DConfig.resolve(ri, wind_speed_m_s=4.0, projection_cg_rtol=1e-12)
    sv = Spread3DSolver(cfg)
    try:
        g = sv.grid
        assert g.Ny == 4 and abs(g.z_face[g.n_z_bed] - 0.2) < 1e-12
        fill, ms, ht = bed_inputs(cfg, g)
        assert fill is not None and 0.0 < fill.max() <= 1.0
        assert sv.n_particles == ms.size and ht.size == ms.size
        sv.run()
        o = OracleSpread3D(cfg, build_grid(cfg), (fill, ms, ht)).run(cfg.n_steps)
        st = sv.state_numpy()
        assert sv.step_no == o.step_no == 25
        assert o.t > 0.1                       # the gusts went past their first period
        speed = float(np.sqrt(np.linalg.norm(o.u) ** 2 + np.linalg.norm(o.v) ** 2 +
                              np.linalg.norm(o.w) ** 2))
        errs = {}
        for name in ("u", "v", "w", "T_g", "T_s", "Y_fuel", "Y_O2", "Y_H2O", "rho", "k",
                     "eps", "nu_t", "omega", "q_rad", "S_pyro", "alpha_s"):
            a_, b_ = st[name], getattr(o, name)
            den = speed if name in ("u", "v", "w") else max(np.linalg.norm(b_), 1e-300)
            errs[name] = float(np.linalg.norm(a_ - b_) / den)
        print("bed variation rel L2 vs oracle:", errs, flush=True)
        for name, e in errs.items():
            assert e < 1e-8, (name, e)
        for k in ("T_s", "m_solid", "m_water", "m_char"):
            a_, b_ = st["p_" + k], o.buf[k][:o.n_alloc]
            assert np.linalg.norm(a_ - b_) <= 1e-8 * max(np.linalg.norm(b_), 1e-300), k
        assert float(np.max(o.T_s)) > cfg.T_amb + 50.0
        np.testing.assert_allclose(sv.diag["t"], o.diag["t"], rtol=1e-12)
        np.testing.assert_allclose(sv.front_x, o.front_x, rtol=1e-12)
        # the inlet of the last step is the log law times g(t) of its start
        assert not np.allclose(sv.u_inlet.cpu().numpy(), sv.u_inlet_base.cpu().numpy())
    finally:
        sv.close()
