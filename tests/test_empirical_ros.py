"""The empirical rate-of-spread module of the active level set (host
scalars) against the known answers the reference's own tests hold
(tests/outdoor/test_empirical_ros_marsden_smedley.py: the Phase 22 CAL and
VAL points, the Tasmania 2009 P = 0.5 boundary, the zero cases, the age
asymptote, dispatch), plus Cheney 1998 Eq. 6 and the blend weight by hand."""
import math

import pytest

from firewebapp_amd import empirical_ros as E


def test_marsden_smedley_known_points():
    assert E.marsden_smedley_ros_m_per_s(2.0, 0.30, 10.0) == pytest.approx(0.04988, abs=1e-5)
    for U, M, age, want in [(4.0, 0.15, 10.0, 0.17832), (6.0, 0.55, 10.0, 0.11484),
                            (1.5, 0.50, 10.0, 0.02104), (0.5, 0.70, 10.0, 0.00306)]:
        assert E.marsden_smedley_ros_m_per_s(U, M, age) == pytest.approx(want, abs=1e-4)


def test_marsden_smedley_limits():
    assert E.marsden_smedley_ros_m_per_s(0.0, 0.30, 10.0) == 0.0
    assert E.marsden_smedley_ros_m_per_s(-1.0, 0.30, 10.0) == 0.0
    assert E.marsden_smedley_ros_m_per_s(2.0, 0.30, 0.0) == 0.0
    r_inf = E.MS_1995_CONST * (7.2 ** E.MS_1995_U_EXP) * math.exp(-E.MS_1995_B_MF * 30.0)
    assert abs(E.marsden_smedley_ros_m_per_s(2.0, 0.30, 100.0) - r_inf) < 1e-4
    assert (E.marsden_smedley_ros_m_per_s(4.0, 0.3, 10.0)
            > E.marsden_smedley_ros_m_per_s(1.0, 0.3, 10.0))
    assert (E.marsden_smedley_ros_m_per_s(2.0, 0.6, 10.0)
            < E.marsden_smedley_ros_m_per_s(2.0, 0.1, 10.0))


def test_p_sustain():
    assert E.marsden_smedley_p_sustain(0.489, 0.30, 1) == pytest.approx(0.5, abs=0.005)
    assert E.marsden_smedley_p_sustain(3.0, 0.5, 1) > E.marsden_smedley_p_sustain(0.5, 0.5, 1)
    assert E.marsden_smedley_p_sustain(2.0, 0.5, 2) > E.marsden_smedley_p_sustain(2.0, 0.5, 1)


def test_cheney_eq6_by_hand():
    # a_ch (0.723 U)^0.987 exp(-0.0707 M%) in m/s for U = 4, M = 4 %, Nat grass
    want = 0.406 * (0.723 * 4.0) ** 0.987 * math.exp(-0.0707 * 4.0)
    assert E.cheney_eq6_ros_m_per_s(4.0, 0.04, 0.406) == pytest.approx(want, rel=1e-14)
    assert E.cheney_eq6_ros_m_per_s(0.0, 0.04, 0.406) == 0.0


def test_dispatch_and_blend():
    assert (E.evaluate_empirical_ros("marsden_smedley", 2.0, 0.30, a_ch=999.0, age_yr=10.0)
            == E.marsden_smedley_ros_m_per_s(2.0, 0.30, 10.0))
    assert (E.evaluate_empirical_ros("cheney_eq6", 3.0, 0.05, 0.343)
            == E.cheney_eq6_ros_m_per_s(3.0, 0.05, 0.343))
    with pytest.raises(ValueError):
        E.get_model("no_such_model")
    with pytest.raises(NotImplementedError):
        E.evaluate_empirical_ros("rothermel", 2.0)
    assert E.list_models() == ["cheney_eq6", "marsden_smedley", "rothermel"]
    assert E.blend_resolved_empirical(1.4, 1.4, 0.5) == 0.0
    assert E.blend_resolved_empirical(0.8, 1.4, 0.5) == 1.0
    assert E.blend_resolved_empirical(1.15, 1.4, 0.5) == pytest.approx(0.5, rel=1e-12)
    assert E.blend_resolved_empirical(1.0, 1.4, 0.0) == 1.0
