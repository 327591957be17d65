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


@pytest.mark.parametrize(("size", "dim"), [(256, 2), (256, 3)])
def test_bodies_cases_are_a_bouzidi_body_with_zou_he_ends(size, dim):
    import math

    import numpy as np

    name, params, ux = bench_mlups.bodies_case(size, dim)
    shape = (size, size) if dim == 2 else (size // 4, size // 4, size)
    assert params.shape == shape and ux.shape == shape
    assert params.solid[..., 0].all() and params.solid[..., -1].all()  # closed first and last x layers
    interior = params.solid[..., 1:-1].copy()
    interior[..., 0, :] = interior[..., -1, :] = False  # side walls
    if dim == 3:
        interior[0] = interior[-1] = False
    radius = size / 16 if dim == 2 else 12.0 * size / 512
    body = math.pi * radius**2 if dim == 2 else 4 / 3 * math.pi * radius**3
    assert interior.sum() == pytest.approx(body, rel=0.3)
    assert len(params.link_rules) > 0 and np.all(ux[params.solid] == 0.0) and ux.max() > 0
    assert params.tau == pytest.approx(1.46 if dim == 2 else 0.62, abs=0.005)  # the same at every size


@pytest.mark.gpu
def test_bodies_comparison_times_both_lattices(capsys):
    import torch

    m = bench_mlups.measure_bodies(torch, "hip-pull", 256, 2, target_seconds=0.05)
    assert m.nodes == 256**2 and m.rules > 0
    assert m.empty_mlups > 0 and m.bodies_mlups > 0 and m.ratio == pytest.approx(m.bodies_mlups / m.empty_mlups)
    bench_mlups.main(["--bodies", "--dim", "3", "--sizes", "64", "--backends", "hip-aa", "--seconds", "0.05"])
    rows = [line for line in capsys.readouterr().out.splitlines() if line.strip().startswith("hip-aa")]
    assert len(rows) == 1 and "duct 64x16x16" in rows[0]
