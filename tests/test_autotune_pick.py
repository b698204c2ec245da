"""The autotuner's choice among timed candidates (host logic, no GPU): the
fastest with an agreeing E_xc, the first kept unless beaten by the margin."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from calchem_amd.autotune import pick  # noqa: E402

A, B, C = (0, 0, 0, 0), (0, 0, 0, 2), (256, 384, 0, 0)


def test_fastest_agreeing_candidate_wins():
    table = [(A, 1.45, -87.0), (B, 1.39, -87.0), (C, 1.30, -86.9)]  # C disagrees in E_xc
    assert pick(table, 1e-7) == (B, 1.39)


def test_first_kept_within_the_margin():
    table = [(A, 1.45, -87.0), (B, 1.43, -87.0)]
    assert pick(table, 1e-7, keep_first_within=0.02) == (A, 1.45)
    assert pick(table, 1e-7, keep_first_within=0.0) == (B, 1.43)


def test_margin_beaten():
    table = [(A, 1.45, -87.0), (B, 1.39, -87.0)]  # 4 % faster
    assert pick(table, 1e-7, keep_first_within=0.02) == (B, 1.39)
