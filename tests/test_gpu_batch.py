"""GPU: a batch table run several cases at a time, each on its own stream
from its own host thread, computes exactly what each case computes alone."""
import os

import numpy as np
import pytest

from tests.golden_io import GOLDEN_DIR

pytestmark = pytest.mark.gpu


def test_concurrent_batch_equals_cases_alone():
    from firewebapp_amd.batch import batch_table, run_batch
    from firewebapp_amd.deck import load_text_input
    from firewebapp_amd.spread_3d import Spread3DConfig, Spread3DSolver
    ri = load_text_input(os.path.join(GOLDEN_DIR, "integration_case.txt"))
    ri.batch_overrides.update({"n_cases": "4", "wind_speed_m_s": "2, 4, 6, 3",
                               "ly": "0.1, 0.4", "n_steps": "15",
                               "initial_moisture_frac": "0.02~0.08", "seed": "5"})
    recs = run_batch(ri, concurrent=4)
    assert [r["index"] for r in recs] == [0, 1, 2, 3]
    for case, rec in zip(batch_table(ri), recs):
        cfg = Spread3DConfig.resolve(case.inputs(ri), wind_speed_m_s=case.wind_speed_m_s)
        sv = Spread3DSolver(cfg)
        try:
            sv.run()
            alone = sv.result()
        finally:
            sv.close()
        got = rec["result"]
        print("case", case.index, rec["Nx"], rec["Ny"], rec["Nz"], "steps", rec["steps"],
              "U10", rec["wind_speed_m_s"], flush=True)
        assert rec["steps"] == 15 and alone.n_steps == 15
        assert got.diag_t == alone.diag_t and got.front_x == alone.front_x
        for k in ("T_g", "T_s", "u", "w", "rho", "Y_fuel", "Y_O2"):
            assert np.array_equal(getattr(got.state_final, k), getattr(alone.state_final, k)), k
