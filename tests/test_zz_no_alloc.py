"""fw_edc_static (the EDC closure with library-owned scratch, no allocation)
against fw_edc (the same kernels with a caller workspace), through the C ABI.

Two calls in a row on one GPU, the second on a smaller grid, so the second
reuses the static summary array the first left behind: both must equal
fw_edc bit for bit (same kernels, same summaries).  The inputs make the
carried wet-bulb flag matter: sparse fuel, wet cells with Y_H2O > 0.09 whose
own flag is set, and resetting cells placed away from the 256-cell segment
starts, so segments inherit their flag through the summaries.  A grid above
the static maximum is rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARGS = (0.34, 1100.0, 0.002, 1, 0, 1.3, 17.0e6)   # chi_rad, cp_g, dt, nsub, ext, s, hoc


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    rho = 0.4 + 0.8 * rng.random(shape)
    T = 350.0 + 1500.0 * rng.random(shape)
    Yf = np.zeros(shape)
    flat = Yf.reshape(-1)
    # a resetting cell every ~700 cells, never on a segment's first cell
    pos = np.arange(37, flat.size, 701)
    pos = pos[pos % 256 != 0]
    flat[pos] = 0.02 + 0.06 * rng.random(pos.size)
    YO = 0.232 * rng.random(shape)
    Yw = 0.2 * rng.random(shape)      # about half the cells above 0.09
    k = 0.8 * rng.random(shape)
    e = 0.8 * rng.random(shape)
    return rho, T, Yf, YO, k, e, Yw


def _run(entry, shape, seed, dtype):
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, ptr, stream_ptr
    nz, ny, nx = shape
    g = DeviceGrid(nx, ny, nz, 1.0, 1.0, [1.0] * nz, [1.0] * nz, [1.0] * nz, dtype=dtype)
    rho, T, Yf, YO, k, e, Yw = (torch.as_tensor(a, dtype=dtype, device="cuda")
                                for a in _inputs(shape, seed))
    om = torch.zeros_like(T)
    head = (g.ref, ptr(rho), ptr(T), ptr(Yf), ptr(YO), ptr(k), ptr(e), ptr(Yw), ptr(om),
            *ARGS)
    if entry == "fw_edc":
        work = torch.empty(max(_lib.lib().fw_edc_workspace_ints(g.ref), 1),
                           dtype=torch.int32, device="cuda")
        _lib.call("fw_edc", *head, ptr(work), stream_ptr())
    else:
        _lib.call("fw_edc_static", *head, stream_ptr())
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (T, Yf, YO, om)]


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_edc_static_matches_workspace_entry(dtype_name):
    import torch
    dtype = getattr(torch, dtype_name)
    for shape, seed in (((12, 16, 512), 5), ((5, 8, 192), 6)):
        print(f"edc_static {dtype_name} shape {shape}", flush=True)
        want = _run("fw_edc", shape, seed, dtype)
        got = _run("fw_edc_static", shape, seed, dtype)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
        # the inputs exercise the carried flag: among the cells without fuel
        # (which inherit it) some were relaxed toward the wet-bulb
        # temperature (flag carried in as 1) and some were not (0)
        _, T0, Yf0, _, _, _, _ = _inputs(shape, seed)
        inherit = (Yf0 <= 1e-9) & (T0 > 373.15)
        relaxed = want[0] != T0.astype(want[0].dtype)
        assert np.any(relaxed & inherit) and np.any(~relaxed & inherit)


def test_edc_static_rejects_oversized_grid():
    import torch
    from firewebapp_amd import _lib
    from firewebapp_amd.device import DeviceGrid, stream_ptr
    nz = 1025                                 # 1025 * 256 * 256 > 67,108,864 cells
    g = DeviceGrid(256, 256, nz, 1.0, 1.0, [1.0] * nz, [1.0] * nz, [1.0] * nz,
                   dtype=torch.float32)
    null = C.c_void_p(0)
    rc = _lib.lib().fw_edc_static(g.ref, null, null, null, null, null, null, null, null,
                                  *ARGS, stream_ptr())
    assert rc == _lib.FW_E_CONFIG
    assert "static scratch" in _lib.lib().fw_last_error().decode()
