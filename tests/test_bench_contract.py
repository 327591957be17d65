"""bench.py's command-line contract: one JSON line with the required keys, one GPU only."""

import json

import pytest

import bench

REQUIRED = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
            "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"}


def test_more_than_one_gpu_is_refused_with_a_message(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_aa_steps_must_come_in_pairs(capsys):
    assert bench.main(["--steps", "3"]) == 2
    assert "even" in capsys.readouterr().err


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
