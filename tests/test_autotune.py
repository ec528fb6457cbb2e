"""Autotuner choice rule (host logic, no GPU)."""
from boltzmann_amd.autotune import PUSH_SHAPES, TIE_SHORT, grid_order, pick_within


def _tile(k):
    return PUSH_SHAPES[k][0] * PUSH_SHAPES[k][1]


def test_smaller_tile_wins_a_tie():
    # (1024, 2) is 0.1 % faster than (1024, 1): a tie, the smaller tile is kept
    t = {0: 0.0323, 1: 0.0261, 2: 0.0315, 3: 0.025377, 4: 0.025345}
    assert pick_within(t, _tile) == 3


def test_clear_winner_kept():
    t = {0: 0.0323, 1: 0.0261, 2: 0.0315, 3: 0.0300, 4: 0.0250}
    assert pick_within(t, _tile) == 4


def test_grid_prefers_fewest_blocks_within_tie():
    g = {0: 0.04447, 1: 0.04090, 2: 0.04123, 4: 0.04082}
    assert pick_within(g, lambda k: k if k > 0 else 99) == 1
    g = {0: 0.030, 1: 0.041, 2: 0.040, 4: 0.039}
    assert pick_within(g, lambda k: k if k > 0 else 99) == 0


def test_grid_keeps_the_default_within_two_percent():
    # push-kernel times of a C4 trial: 4 blocks per CU 1 % faster is noise
    g = {0: 0.936, 1: 0.993, 2: 0.966, 4: 0.927}
    assert pick_within(g, grid_order, TIE_SHORT) == 0
    g = {0: 0.936, 1: 0.993, 2: 0.966, 4: 0.900}
    assert pick_within(g, grid_order, TIE_SHORT) == 4
