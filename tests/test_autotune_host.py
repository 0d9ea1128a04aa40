"""CPU: the autotuner's choice rule (firewebapp_amd.autotune._pick)."""
from firewebapp_amd import autotune


def test_default_kept_within_margin():
    # 4 % faster is within the 5 % margin: the default stays
    assert autotune._pick({4: 1.00, 8: 0.96, 2: 1.2}, 4) == 4


def test_faster_candidate_beyond_margin():
    assert autotune._pick({4: 1.00, 8: 0.90, 2: 1.2}, 4) == 8


def test_fastest_when_default_absent():
    # a default the grid cannot run (not timed) leaves the fastest
    assert autotune._pick({(0, 512, 16): 2.0, (1, 512, 16): 1.5}, (1, 1024, 16)) == (1, 512, 16)
