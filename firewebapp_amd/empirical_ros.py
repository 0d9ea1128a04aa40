"""Empirical rate-of-spread models for the active level set's blend.

Host-side scalars, evaluated once per run (spread_3d.py:1246-1272): the
empirical ROS at the inflow wind and the weight it gets in the front speed,
v_n = (1 - w) v_resolved + w ROS (spread_3d.py:3252-3255).  Same names and
argument meaning as model_outdoor/empirical_ros.py:

* ``cheney_eq6_ros_m_per_s`` -- Cheney, Gould & Catchpole 1998 (IJWF 8:1),
  Eq. 6: ROS = a_ch (0.723 U)^0.987 exp(-0.0707 M%) [m/min];
* ``marsden_smedley_ros_m_per_s`` -- Marsden-Smedley & Catchpole 1995
  (IJWF 5:215): ROS = 0.678 U_kmh^1.312 exp(-0.0243 M%) (1 - exp(-0.116 age))
  [m/min], U at 1.7 m;
* ``marsden_smedley_p_sustain`` -- the Marsden-Smedley 2001 sustaining
  logistic;
* ``blend_resolved_empirical`` -- 1 below threshold - width, 0 at or above
  the threshold, linear between;
* ``evaluate_empirical_ros`` / ``get_model`` / ``list_models`` -- dispatch by
  model name.
"""
from __future__ import annotations

import math

CHENEY_EQ6_U_EXP = 0.987
CHENEY_EQ6_B_MF = 0.0707
CHENEY_EQ6_U2_RATIO = 0.723
MS_1995_CONST = 0.678 / 60.0
MS_1995_U_EXP = 1.312
MS_1995_B_MF = 0.0243
MS_1995_AGE_LAMBDA = 0.116


def cheney_eq6_ros_m_per_s(U_m_s: float, moisture_frac: float, a_ch: float) -> float:
    if U_m_s <= 0.0:
        return 0.0
    u2 = max(0.0, CHENEY_EQ6_U2_RATIO * U_m_s)
    per_min = (a_ch * (u2 ** CHENEY_EQ6_U_EXP)
               * math.exp(-CHENEY_EQ6_B_MF * (moisture_frac * 100.0)) * 60.0)
    return max(0.0, per_min / 60.0)


def marsden_smedley_ros_m_per_s(U_1p7_m_s: float, moisture_frac: float,
                                age_yr: float) -> float:
    if U_1p7_m_s <= 0.0 or age_yr <= 0.0:
        return 0.0
    u_kmh = 3.6 * U_1p7_m_s
    ros = (MS_1995_CONST * (u_kmh ** MS_1995_U_EXP)
           * math.exp(-MS_1995_B_MF * (moisture_frac * 100.0))
           * (1.0 - math.exp(-MS_1995_AGE_LAMBDA * age_yr)))
    return max(0.0, ros)


def marsden_smedley_p_sustain(U_1p7_m_s: float, moisture_frac: float,
                              productivity: int = 1) -> float:
    u_kmh = 3.6 * U_1p7_m_s
    m = moisture_frac * 100.0
    a = -1.0 + 0.68 * u_kmh - 0.07 * m - 0.0037 * u_kmh * m + 2.1 * productivity
    return 1.0 / (1.0 + math.exp(-a))


def blend_resolved_empirical(U_m_s: float, u_threshold_m_s: float,
                             blend_width_m_s: float) -> float:
    """Weight of the empirical ROS in the blended front speed."""
    if U_m_s >= u_threshold_m_s:
        return 0.0
    if blend_width_m_s <= 0.0:
        return 1.0
    if U_m_s <= u_threshold_m_s - blend_width_m_s:
        return 1.0
    return (u_threshold_m_s - U_m_s) / blend_width_m_s


def _cheney(U_m_s, moisture_frac, a_ch, **_):
    return cheney_eq6_ros_m_per_s(U_m_s, moisture_frac, a_ch)


def _marsden_smedley(U_m_s, moisture_frac, age_yr=10.0, **_):
    return marsden_smedley_ros_m_per_s(U_m_s, moisture_frac, age_yr)


def _rothermel(U_m_s, **_):
    raise NotImplementedError("Rothermel 1972 is not implemented (neither is it in "
                              "the reference); use 'cheney_eq6' for grass")


_MODELS = {"cheney_eq6": _cheney, "marsden_smedley": _marsden_smedley,
           "rothermel": _rothermel}


def list_models() -> list:
    return sorted(_MODELS)


def get_model(name: str):
    """The ROS function of a model name; ValueError for an unknown one."""
    if name not in _MODELS:
        raise ValueError(f"Unknown fuel model '{name}'.  Registered: {list_models()}")
    return _MODELS[name]


def evaluate_empirical_ros(model: str, U_m_s: float, moisture_frac: float = None,
                           a_ch: float = None, **kwargs) -> float:
    if moisture_frac is not None:
        kwargs["moisture_frac"] = moisture_frac
    if a_ch is not None:
        kwargs["a_ch"] = a_ch
    return get_model(model)(U_m_s, **kwargs)
