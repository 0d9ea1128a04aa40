"""Deck format and outdoor configuration (host side, one-shot).

* ``load_text_input``: the ``key = value`` deck reader, outdoor branch of
  src/model/io/text_input.py:308-325, :1022-1033 (comments with ``#``, keys
  containing "sigma" dropped, ``outdoor.*`` values parsed bool first --
  "1"/"0"/yes/no are booleans -- then float, none/null/nan -> None, else
  string; keys lower-cased).
* ``OutdoorEnvConfig`` / ``outdoor_env_from_dict``: model_outdoor/config.py:
  29-242 (case-insensitive field assignment).

Keys this build adds (all additive; absent keys reproduce the reference):
``n_steps`` (fixed outer-step count), ``dom_quadrature`` (4 or 8),
``moisture_min`` / ``moisture_max`` / ``moisture_seed`` (per-seed moisture
draw), ``particle_jitter`` / ``particle_jitter_seed`` (per-particle +-load
and +-depth jitter), ``fuel_depth_cv`` / ``fuel_depth_corr_length_m`` /
``fuel_depth_seed`` / ``fuel_depth_max_m`` (a seeded lognormal fuel depth
field with mean fuel_depth_m; the bed layers span fuel_depth_max_m),
``wind_gust_amplitude`` / ``wind_gust_period_s`` / ``wind_gust_seed`` (seeded
inlet gusts), ``dtype`` (float32 / float64).

A new ``batch.*`` section (raw value strings, keys lower-cased) describes a
sweep of cases over the deck; ``firewebapp_amd.batch`` expands it into the
batch table.  The reference parser ignores unknown sections, so a batch deck
is still a valid single-case deck for it.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, Optional


def _parse_bool(text: str) -> Optional[bool]:
    val = text.strip().lower()
    if val in {"true", "1", "yes", "y"}:
        return True
    if val in {"false", "0", "no", "n"}:
        return False
    return None


def _parse_float(text: str) -> Optional[float]:
    try:
        return float(text)
    except ValueError:
        return None


def _parse_optional_float(text: str) -> Optional[float]:
    val = text.strip().lower()
    if val in {"none", "null", "nan", ""}:
        return None
    return _parse_float(text)


@dataclass
class RomInputs:
    """The part of the reference's RomInputs the Tier-3 step reads."""
    outdoor_overrides: Dict[str, Any] = field(default_factory=dict)
    batch_overrides: Dict[str, str] = field(default_factory=dict)
    other: Dict[str, str] = field(default_factory=dict)


def parse_deck_text(text: str) -> RomInputs:
    inputs = RomInputs()
    for raw in text.splitlines():
        line = raw.strip()
        if not line or line.startswith("#") or line.startswith(";"):
            continue
        if "=" not in line:
            continue
        key, value = [s.strip() for s in line.split("=", 1)]
        if "#" in value:
            value = value[:value.index("#")].strip()
        key_l = key.lower()
        if "sigma" in key_l:
            continue
        if key_l.startswith("outdoor."):
            fld = key_l.split(".", 1)[1]
            b = _parse_bool(value)
            if b is not None:
                inputs.outdoor_overrides[fld] = b
                continue
            val_opt = _parse_optional_float(value)
            if val_opt is not None or value.strip().lower() in {"none", "null", "nan", ""}:
                inputs.outdoor_overrides[fld] = val_opt
                continue
            inputs.outdoor_overrides[fld] = value.strip()
            continue
        if key_l.startswith("batch."):
            inputs.batch_overrides[key_l.split(".", 1)[1]] = value.strip()
            continue
        inputs.other[key_l] = value
    return inputs


def load_text_input(path) -> RomInputs:
    return parse_deck_text(Path(path).read_text(encoding="utf-8"))


@dataclass
class OutdoorEnvConfig:
    wind_speed_m_s: float = 0.0
    ambient_RH_frac: float = 0.30
    ambient_T_K: float = 300.0
    fuel_depth_m: float = 0.30
    bulk_density_kg_m3: float = 3.0
    sav_ratio_1_m: float = 6562.0
    initial_moisture_frac: float = 0.08
    fuel_lag_time_s: float = 3600.0
    terrain: str = "open"
    free_burning_mode: bool = False
    ignition_q_kW_m2: float = 40.0
    ignition_duration_s: float = 30.0
    rothermel_packing_enable: bool = False
    volume_weighted_projection: bool = False
    sem_enable: bool = False
    sem_seed: int = 42
    sem_I_t: float = 0.20
    sem_N: int = 200
    spin_up_s: float = 0.0
    fuel_pert_enable: bool = False
    fuel_pert_amp: float = 0.05
    fuel_pert_kmax: int = 2
    fuel_pert_seed: int = 13
    canopy_C_d: float = 0.30
    canopy_beta_p: float = 1.0
    canopy_beta_d: float = 4.0


def outdoor_env_from_dict(d: dict) -> OutdoorEnvConfig:
    cfg = OutdoorEnvConfig()
    field_map = {f.lower(): f for f in cfg.__dataclass_fields__}
    for key, val in d.items():
        actual = field_map.get(key.lower())
        if actual is not None:
            setattr(cfg, actual, val)
    return cfg
