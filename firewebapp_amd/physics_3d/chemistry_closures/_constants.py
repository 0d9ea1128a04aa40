"""Closure constants (chemistry_closures/_constants.py:15-84 of the reference)."""
S_STOICH = 1.3
Y_O2_AIR = 0.232
A_COMB = 1.0e9
E_COMB = 84_000.0
C_EBU = 1.0
HOC_J = 17_000_000.0

CHEMISTRY_FAMILIES = {
    "biomass": dict(s_stoich=1.3, hoc_J=17_000_000.0, a_comb=1.0e9,
                    e_comb=84_000.0, c_ebu=1.0),
    "methane": dict(s_stoich=4.0, hoc_J=50_000_000.0, a_comb=2.1e11,
                    e_comb=125_000.0, c_ebu=4.0),
}


def resolve_chemistry_family(family: str) -> dict:
    if family not in CHEMISTRY_FAMILIES:
        raise ValueError(f"Unknown chemistry family '{family}'.  "
                         f"Registered: {sorted(CHEMISTRY_FAMILIES.keys())}")
    return dict(CHEMISTRY_FAMILIES[family])
