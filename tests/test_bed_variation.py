"""Seeded bed and inflow variation (no reference counterpart): the lognormal
fuel depth field, the bed fill and per-particle inputs it implies, the inlet
gust schedule, and the oracle's use of them (CPU)."""
import math

import numpy as np
import pytest

from firewebapp_amd.bed_variation import GustSchedule, bed_variation, fuel_depth_field


def test_depth_field_statistics():
    h = fuel_depth_field(512, 128, 0.1, 0.1, 0.21, 0.3, 1.0, seed=3)
    assert h.shape == (128, 512) and np.all(h > 0.0)
    # sample moments of the log are exact by construction
    s2 = math.log1p(0.09)
    lg = np.log(h)
    assert abs(lg.mean() - (math.log(0.21) - 0.5 * s2)) < 1e-12
    assert abs(lg.std() - math.sqrt(s2)) < 1e-12
    # the depth's own moments follow within sampling error
    assert abs(h.mean() / 0.21 - 1.0) < 0.03
    assert abs(h.std() / h.mean() - 0.3) < 0.03
    # correlation exp(-r^2 / l^2): e^-1 at one correlation length, ~0 at 3 l
    G = (lg - lg.mean()) / lg.std()
    corr = lambda s: float(np.mean(G * np.roll(G, s, axis=0)))
    assert abs(corr(10) - math.exp(-1.0)) < 0.1
    assert abs(corr(30)) < 0.1
    assert corr(1) > corr(5) > corr(10)
    # seeded: same seed same field, another seed another field
    assert np.array_equal(h, fuel_depth_field(512, 128, 0.1, 0.1, 0.21, 0.3, 1.0, seed=3))
    assert not np.array_equal(h, fuel_depth_field(512, 128, 0.1, 0.1, 0.21, 0.3, 1.0, seed=4))
    assert np.all(fuel_depth_field(8, 4, 0.1, 0.1, 0.21, 0.0, 1.0, 0) == 0.21)


def test_bed_fill_follows_column_depth():
    Nz, Ny, Nx, nzb, npc = 12, 6, 20, 5, 4
    z_face = np.concatenate([[0.0], np.cumsum([0.05] * nzb + [0.1] * (Nz - nzb))])
    mask = np.zeros((Nz, Ny, Nx), bool)
    mask[:nzb, :, 3:] = True
    depth = fuel_depth_field(Nx, Ny, 0.1, 0.1, 0.12, 0.3, 0.3, seed=1)
    depth = np.minimum(depth, z_face[nzb])
    fill, ms, ht = bed_variation(mask, z_face, nzb, npc, 0.12, depth, 0.0, 0)
    assert set(np.unique(fill)) <= {0.0, 1.0} and np.all(fill[nzb:] == 0.0)
    assert ms is None
    # whole cells from the ground up; the filled height is the depth rounded
    # to the nearest face (at least the first layer)
    f = fill[:nzb, :, 3:]
    assert np.all(np.diff(f, axis=0) <= 0.0)
    col = (f * 0.05).sum(axis=0)
    assert np.all(np.abs(col - np.maximum(depth[:, 3:], 0.025)) <= 0.025 + 1e-15)
    assert np.all(fill[:, :, :3] == 0.0)
    # one fuel top per particle, in the initialiser's slot order (k, then
    # (j, i) in C order over the filled cells): its column's filled height
    sel = fill[:nzb] > 0.0
    assert ht.size == int(sel.sum()) * npc
    top = np.zeros((Ny, Nx))
    top[:, 3:] = col
    np.testing.assert_allclose(ht, np.repeat(np.broadcast_to(top, (nzb, Ny, Nx))[sel], npc),
                               rtol=0, atol=1e-15)
    # every particle of the initialiser lies below its fuel top
    from oracle import bed as Bd
    buf = Bd.allocate_bed_particle_buffers(ht.size)
    n = Bd.initialize_bed_particles_from_alpha_s(buf, 0.004 * fill, 1.5, 0.05, 300.0, 0.1, 0.1,
                                                 np.diff(z_face), nzb, npc)
    assert n == ht.size and np.all(buf["z"][:n] < ht)


def test_bed_jitter():
    mask = np.zeros((6, 4, 10), bool)
    mask[:3, :, 2:] = True
    z_face = np.arange(7) * 0.1
    fill, ms, ht = bed_variation(mask, z_face, 3, 5, 0.3, None, 0.1, 7)
    assert fill is None and ms.size == ht.size == 3 * 4 * 8 * 5
    assert ms.min() >= 0.9 and ms.max() <= 1.1 and abs(ms.mean() - 1.0) < 0.02
    assert ht.min() >= 0.27 and ht.max() <= 0.33
    rng = np.random.default_rng(7)
    np.testing.assert_array_equal(ms, 1.0 + 0.1 * rng.uniform(-1.0, 1.0, ms.size))
    np.testing.assert_array_equal(ht, 0.3 * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, ms.size)))
    assert bed_variation(mask, z_face, 3, 5, 0.3, None, 0.0, 7) == (None, None, None)


def test_gust_schedule():
    g = GustSchedule(0.15, 5.0, seed=9)
    assert g.factor(0.0) == 1.0
    ts = np.linspace(0.0, 60.0, 6001)
    f = np.array([g.factor(t) for t in ts])
    assert np.all(np.abs(f - 1.0) <= 0.15 + 1e-15)
    assert np.max(np.abs(np.diff(f))) < 0.15 * math.pi / 2 * 0.01 / 5.0 * 2 + 1e-12   # continuous
    rng = np.random.default_rng(9)
    nodes = [0.0] + [0.15 * (2.0 * rng.random() - 1.0) for _ in range(12)]
    for m in range(12):
        assert g.factor(5.0 * m) == pytest.approx(1.0 + nodes[m], abs=1e-15)
    # nodes are drawn in order whatever order t is queried in
    h = GustSchedule(0.15, 5.0, seed=9)
    assert [h.factor(t) for t in ts[::-1]][::-1] == list(f)
    assert GustSchedule(0.0, 5.0, 9).factor(12.3) == 1.0


def test_oracle_gust_matches_schedule():
    from oracle.step import OracleSpread3D
    o = OracleSpread3D.__new__(OracleSpread3D)
    o.gust_A, o.gust_T = 0.15, 0.7
    o._gust_rng, o._gust_nodes = np.random.default_rng(5), [0.0]
    g = GustSchedule(0.15, 0.7, seed=5)
    for t in np.linspace(0.0, 9.0, 377):
        assert o.gust(t) == g.factor(t)


def test_oracle_bed_h_top():
    """The per-particle fuel top: h_top = h_bed everywhere is the scalar
    path bit for bit; h_top = z (no fuel above) is the view factor without
    canopy attenuation."""
    from oracle import bed as Bd
    Nz, Ny, Nx, nzb, npc = 6, 3, 8, 3, 4
    dz = np.array([0.05, 0.05, 0.05, 0.2, 0.2, 0.2])
    z_face = np.concatenate([[0.0], np.cumsum(dz)])
    alpha = np.zeros((Nz, Ny, Nx))
    alpha[:nzb, :, 2:] = 0.004

    def run(h_top, geometric=True):
        buf = Bd.allocate_bed_particle_buffers(nzb * Ny * 6 * npc)
        n = Bd.initialize_bed_particles_from_alpha_s(buf, alpha, 1.5, 0.05, 300.0, 0.1, 0.1,
                                                     dz, nzb, npc, sav=9770.0)
        buf["T_s"][:n] = 300.0 + 500.0 * np.random.default_rng(2).random(n)
        if h_top is not None:
            buf["h_top"] = h_top(buf)[:n].copy()
        Tg = np.full((Nz, Ny, Nx), 900.0)
        Q = np.zeros((Nz, Ny, Nx))
        src, _, _, _ = Bd.step_bed_particles(
            buf, Tg, np.full((Nz, Ny, Nx), 0.2), Q, npc, 0.1, 0.1, dz, z_face, 250.0, 380.0,
            1500.0, 0.9, 300.0, 1.0, geometric, 0.15, 5.0, 0.01, True, True, True, True, 2,
            1.0e5, 0.0)
        return buf["T_s"][:n].copy(), src["Q_g_conv"]

    a = run(None)
    b = run(lambda buf: np.full(buf["z"].shape, 0.15))
    np.testing.assert_array_equal(a[0], b[0])
    c = run(lambda buf: buf["z"].copy())
    d = run(None, geometric=False)
    np.testing.assert_array_equal(c[0], d[0])
    assert not np.array_equal(a[0], d[0])
