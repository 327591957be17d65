"""The sweep benchmark (the reference's benchmarks/bench_mlups.py protocol) on small grids."""

import pytest

from lbm import bench_mlups


def test_grid_params_are_fully_periodic_trt():
    p = bench_mlups.grid_params(16, 3, "float32")
    assert p.shape == (16, 16, 16) and p.collision == "trt" and p.tau == 0.8
    assert not p.solid.any()
    assert bench_mlups.grid_params(16, 2, "float64").shape == (16, 16)


def test_sweep_refuses_to_run_without_a_device(monkeypatch):
    monkeypatch.setattr(bench_mlups, "available_backends", lambda: ())
    with pytest.raises(SystemExit, match="no HIP backend"):
        bench_mlups.main(["--sizes", "16"])


@pytest.mark.gpu
def test_measure_reports_consistent_rates():
    import torch

    m = bench_mlups.measure(torch, "hip-aa", 64, 3, "float32", target_seconds=0.05)
    assert m.steps % 2 == 0 and m.steps >= 2
    assert m.mlups == pytest.approx(64**3 * m.steps / m.seconds / 1e6)
    assert m.bandwidth_gbs == pytest.approx(m.mlups * 152 / 1e3)
    assert m.fits_in_cache and 0 < m.peak_gbs < 8000.0


@pytest.mark.gpu
def test_sweep_prints_one_row_per_backend_and_size(capsys, tmp_path):
    out = tmp_path / "sweep.csv"
    bench_mlups.main(["--sizes", "64", "128", "--seconds", "0.05", "--csv", str(out)])
    rows = [line for line in capsys.readouterr().out.splitlines() if line.strip().startswith("hip-")]
    assert len(rows) == 6
    assert len(out.read_text().splitlines()) == 7
