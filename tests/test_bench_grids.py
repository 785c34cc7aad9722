"""Host logic of the benchmark (no GPU): the case grids of the headline
campaign and of the learned warm-start configuration, the campaign's slot
occupancy from its trace, and the lookup of profiled HBM traffic."""
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


def test_training_grid_holds_no_test_case():
    """The coarse training grid lies on the test grid (test cases on it are
    moved off); the fine one's angles lie between the test grid's steps."""
    grid = bench.train_grid()
    assert len(grid) == len(set(grid)) == 216
    tests = {(m, a) for _, m, a, _ in bench.ragged_grid(1024)}
    assert not tests & set(grid)
    fine = [(m, a) for m in bench.TRAIN_MACHS_FINE for a in bench.TRAIN_AOAS_FINE]
    assert all(a not in AOAS for _, a in fine)
    assert all(0.7 <= m <= 3.0 and -5.0 <= a <= 5.0 for m, a in grid)


def test_ragged_grid_stays_off_the_training_grid():
    grid = bench.ragged_grid(1024)
    assert len(grid) == 1024
    assert {g for g, *_ in grid} == set(bench.GEOMETRIES)
    assert {f for *_, f in grid} == {"HLLC", "ROE"}
    for g, m, a, f in grid:
        assert 0.7 <= m <= 3.01 and a in AOAS
        assert not (m in bench.TRAIN_MACHS and a in bench.TRAIN_AOAS)


def test_occupancy_from_a_campaign_trace():
    # (seconds, step, cases in flight, cases waiting) after every chunk
    trace = [(0.0, 100, 4, 6), (1.0, 200, 4, 2), (2.0, 300, 4, 0), (3.0, 400, 2, 0), (4.0, 500, 1, 0)]
    occ = bench.occupancy(trace, slots=4)
    # in flight over each interval, weighted by its length: (4 + 4 + 2 + 1) / 4
    assert occ["mean_cases_in_flight"] == 11 / 4
    assert occ["queue_empty_at_s"] == 2.0 and occ["tail_s"] == 2.0 and occ["hr_s"] == 4.0
    assert bench.occupancy(trace[:1], slots=4) is None


def test_measured_traffic_needs_the_profiled_stage_path(tmp_path, monkeypatch):
    path = tmp_path / "traffic_bench_x.txt"
    path.write_text("# launch cell,64,1\n# kernel fetch write\nk_cell<1, 100, 0> 100 20\nk_cell<1, 4, 0> 10 2\n"
                    "k_recon 50 80\nk_update<2> 30 10\n")
    monkeypatch.setattr(bench, "traffic_file", lambda tag: str(path))
    pinned = {"path": "cell", "block": 64, "xcd_order": True}
    # flux-templated kernels: every instantiation of the launch set, summed
    assert bench.measured_traffic("k_cell<stage1>", pinned, "x") == 132.0
    assert bench.measured_traffic("k_recon", pinned, "x") == 130.0
    assert bench.measured_traffic("k_update<stage1>", pinned, "x") == 40.0
    # another block size or order on the same path: the profiled bytes, with
    # the launch they were profiled at named
    assert bench.measured_traffic("k_cell<stage1>", dict(pinned, block=256, xcd_order=False), "x") == 132.0
    assert bench.traffic_launch("x") == "launch cell,64,1"
    # another stage path: no traffic figure
    assert bench.measured_traffic("k_cell<stage1>", dict(pinned, path="face"), "x") is None
