"""Closure registry (chemistry_closures/__init__.py:32-62 of the reference).

Only 'edc' is on the hot path; other reference closure names are rejected
with the reference's ValueError.
"""
from . import edc

_REGISTRY = {"edc": edc.run}


def available():
    return tuple(sorted(_REGISTRY))


def run(closure_name: str, **kwargs) -> None:
    fn = _REGISTRY.get(closure_name)
    if fn is None:
        raise ValueError(f"combustion_closure={closure_name!r} not in {available()!r}")
    fn(**kwargs)
