"""Autotuner choice among probed engine configurations (host logic)."""
from snogray_amd import batch


def cfg(cap, grid=0):
    return {"path_capacity": cap, "grid_blocks": grid}


def test_pick_prefers_larger_capacity_when_rates_tie():
    # a probe that fits in 32 M paths measures 32 M and 64 M alike
    table = [(cfg(1 << 24), 370.0), (cfg(1 << 25), 455.8), (cfg(1 << 26), 453.7),
             (cfg(1 << 26, 8192), 401.0)]
    assert batch.pick(table) == cfg(1 << 26)


def test_pick_keeps_a_clearly_faster_smaller_capacity():
    table = [(cfg(1 << 25), 500.0), (cfg(1 << 26), 450.0)]
    assert batch.pick(table) == cfg(1 << 25)


def test_pick_takes_best_grid_at_the_chosen_capacity():
    table = [(cfg(1 << 26, 0), 440.0), (cfg(1 << 26, 16384), 452.0), (cfg(1 << 25), 455.0)]
    assert batch.pick(table) == cfg(1 << 26, 16384)
