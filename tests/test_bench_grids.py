"""The case grids of the benchmark (host logic, no GPU): the ragged batch of
the headline campaign and the training / held-out split of the learned
warm-start configuration."""
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

import bench  # noqa: E402

MACHS = {round(0.70 + 0.01 * i, 2) for i in range(231)}
AOAS = {-5.0 + 0.5 * j for j in range(21)}


def test_learned_split_is_disjoint_on_the_grid_and_seeded():
    grids = bench.learned_grid(2000, 64)
    assert set(grids) == set(bench.LEARNED_GEOMETRIES)
    for train, test in grids.values():
        assert len(set(train)) == 2000 and len(set(test)) == 64
        assert not set(train) & set(test)
        assert all(m in MACHS and a in AOAS for m, a in train + test)
    assert bench.learned_grid(2000, 64) == grids          # default_rng(1): reproducible


def test_ragged_grid_stays_off_the_training_grid():
    grid = bench.ragged_grid(1024)
    assert len(grid) == 1024
    assert {g for g, *_ in grid} == set(bench.GEOMETRIES)
    assert {f for *_, f in grid} == {"HLLC", "ROE"}
    for g, m, a, f in grid:
        assert 0.7 <= m <= 3.01 and a in AOAS
        assert not (m in bench.TRAIN_MACHS and a in bench.TRAIN_AOAS)
