"""bench.py's command-line contract: one JSON line with the required keys, one GPU only."""

import json

import pytest

import bench

REQUIRED = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
            "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"}


def test_more_than_one_gpu_is_refused_with_a_message(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_counts_round_up_to_whole_aa_pairs():
    assert [bench.round_up(k, 2) for k in (0, 1, 2, 3, 5, 20)] == [0, 2, 2, 4, 6, 20]
    assert [bench.round_up(k, 1) for k in (0, 3, 5)] == [0, 3, 5]


def test_non_positive_step_counts_are_refused(capsys):
    assert bench.main(["--steps", "0"]) == 2
    assert "--steps" in capsys.readouterr().err


def test_a_failing_leg_is_reported_in_its_field(capsys):
    def leg():
        raise RuntimeError("fixture missing")

    assert bench.guarded("leg", leg) == {"error": "RuntimeError: fixture missing"}
    assert bench.guarded("leg", lambda: {"mlups": 1.0}) == {"mlups": 1.0}


def test_odd_aa_step_counts_are_refused_by_the_solver():
    """_advance of the AA variant only runs whole pairs: an odd count is an error, never a
    silently dropped step (the timed loop and time_step would disagree)."""
    from lbm.backends.hip_solver import HipSolver

    solver = HipSolver.__new__(HipSolver)
    solver.variant, solver.step_multiple, solver.graphs = "aa", 2, False
    with pytest.raises(ValueError, match="multiples of 2"):
        solver._advance(3)
    with pytest.raises(ValueError, match="multiples of 2"):
        solver._advance(-2)


@pytest.mark.gpu
def test_driver_counts_with_an_odd_warmup_give_one_json_line(capsys):
    """The driver's own command line (--steps 20 --warmup 5) at a small size: rc 0, one JSON
    line, the warm-up rounded up to a whole AA pair and the counts run reported."""
    rc = bench.main(["--size", "64", "--steps", "20", "--warmup", "5", "--no-cpu-baseline", "--no-adjoint",
                     "--no-batched", "--no-lds", "--no-two-lattice"])
    assert rc == 0
    lines = [line for line in capsys.readouterr().out.splitlines() if line.startswith("{")]
    assert len(lines) == 1
    record = json.loads(lines[0])
    assert record["steps"] == 20 and record["warmup"] == 6
    assert record["steps_requested"] == 20 and record["warmup_requested"] == 5
    assert REQUIRED <= set(record) and record["value"] > 0


@pytest.mark.gpu
def test_odd_timed_count_is_rounded_up_and_reported(capsys):
    rc = bench.main(["--size", "64", "--steps", "3", "--warmup", "1", "--no-cpu-baseline", "--no-adjoint",
                     "--no-batched", "--no-lds", "--no-two-lattice"])
    assert rc == 0
    record = json.loads([line for line in capsys.readouterr().out.splitlines() if line.startswith("{")][0])
    assert record["steps"] == 4 and record["warmup"] == 2
    assert record["ms_per_step"] == pytest.approx(1e3 * 64**3 * 4 / (record["value"] * 1e6) / 4, rel=1e-2)


@pytest.mark.gpu
def test_json_line_carries_the_contract_keys(capsys):
    rc = bench.main(["--size", "64", "--steps", "4", "--warmup", "2", "--no-cpu-baseline", "--no-adjoint",
                     "--no-batched", "--no-lds", "--no-two-lattice"])
    assert rc == 0
    lines = [line for line in capsys.readouterr().out.splitlines() if line.startswith("{")]
    assert len(lines) == 1
    record = json.loads(lines[0])
    assert REQUIRED <= set(record)
    assert record["n_gpus"] == 1 and record["steps"] == 4 and record["warmup"] == 2 and record["unit"] == "MLUPS"
    assert record["higher_is_better"] is True and record["vs_baseline"] is None and record["dtype"] == "f32"
    assert "workload" in record["config"] and record["value"] > 0
    roof = record["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"], rel=1e-3)
