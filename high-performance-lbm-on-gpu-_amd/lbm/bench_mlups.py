"""Throughput (MLUPS) and achieved memory bandwidth of the HIP backends over grid sizes.

The measurement protocol of the reference's ``benchmarks/bench_mlups.py:66-154``: a fully
periodic TRT grid (tau = 0.8) started from a seed-0 N(0, 0.01) velocity field, warmed up,
the timed run sized to about one second, the device-copy peak re-measured right before
each timed run, and

    MLUPS              = nodes * steps / seconds / 1e6   (host clock around sync -> run -> sync)
    bytes per update   = 2 * Q * sizeof(real)            (72 B D2Q9 fp32, 152 B D3Q19 fp32)
    bandwidth          = MLUPS * 1e6 * bytes per update / 1e9
    peak               = the faster of two 2 GiB device copies: the library's copy kernel and
                         the runtime's device-to-device copy (read + write bytes)

Grids whose population state fits in the 256 MB Infinity Cache can exceed HBM bandwidth;
the table flags them.

``--bodies`` instead times each lattice with a body against the same lattice empty (the
reference's "bodies cost little" comparison, README.md:70-72: a Bouzidi cylinder with a
Zou-He inlet and outlet kept 96 % of the empty-domain throughput at D2Q9 1024^2, 77-83 %
in 3D): a Bouzidi cylinder of radius n/16 in an n x n channel, or a Bouzidi sphere in an
n x n/4 x n/4 duct, with a parabolic Zou-He inlet and a pressure outlet, rule pre-pass
included, against the fully periodic lattice of the same size and tau.

Usage::

    python -m lbm.bench_mlups                                   # D2Q9, all backends, 256..4096
    python -m lbm.bench_mlups --dim 3 --backends hip-pull hip-aa --sizes 64 128 256
    python -m lbm.bench_mlups --dtype float64 --csv results.csv
    python -m lbm.bench_mlups --bodies --backends hip-pull hip-aa            # 1024^2, 4096^2
    python -m lbm.bench_mlups --bodies --dim 3 --backends hip-pull hip-aa    # 512 x 128 x 128
"""

from __future__ import annotations

import argparse
import csv
import time
from dataclasses import asdict, dataclass

import numpy as np

from lbm.backends import BACKENDS, available_backends, make_solver
from lbm.backends.hip_solver import TWO_D_ONLY
from lbm.params import SimulationParams

TARGET_SECONDS = 1.0
"""Each timed run is sized to take about this long."""
INFINITY_CACHE_BYTES = 256 * 2**20
COPY_BYTES = 2 * 2**30


@dataclass(frozen=True)
class Measurement:
    backend: str
    dtype: str
    dim: int
    size: int
    steps: int
    seconds: float
    mlups: float
    bandwidth_gbs: float
    peak_gbs: float
    fraction_of_peak: float
    fits_in_cache: bool
    block: tuple[int, int]
    groups_per_cu: int


def copy_peak(torch, nbytes: int = COPY_BYTES, repeats: int = 10) -> dict[str, float]:
    """Read + write GB/s of two device copies of ``nbytes``: the library copy kernel (one
    16-byte nontemporal access per lane) and the runtime's device-to-device copy (torch
    ``copy_``, the analogue of the reference's ``cp.copyto`` probe); ``peak`` is the larger."""
    from lbm.backends._native import library

    lib = library()
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()

    def rate(copy) -> float:
        copy()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(repeats):
            copy()
        stop.record(stream)
        stop.synchronize()
        return 2 * nbytes * repeats / (start.elapsed_time(stop) / 1e3) / 1e9

    kernel = rate(lambda: lib.call("lbm_device_copy", dst.data_ptr(), src.data_ptr(), nbytes, stream.cuda_stream))
    runtime = rate(lambda: dst.copy_(src))
    del src, dst
    return {"kernel": kernel, "runtime": runtime, "peak": max(kernel, runtime)}


def grid_params(size: int, dim: int, dtype: str) -> SimulationParams:
    """Fully periodic size^dim grid (``bench_mlups.py:112-116``)."""
    sizes = {"nx": size, "ny": size} | ({"nz": size} if dim == 3 else {})
    return SimulationParams(**sizes, tau=0.8, collision="trt", dtype=dtype)


def timed_run(solver, steps: int) -> float:
    """Seconds of ``solver.run(steps)`` between two device synchronisations."""
    solver.synchronize()
    start = time.perf_counter()
    solver.run(steps, check_every=steps)
    solver.synchronize()
    return time.perf_counter() - start


def measure(torch, backend: str, size: int, dim: int, dtype: str, target_seconds: float = TARGET_SECONDS,
            **options) -> Measurement:
    """One backend on a fully periodic size^dim grid (``bench_mlups.py:119-154``)."""
    params = grid_params(size, dim, dtype)
    solver = make_solver(params, backend=backend, **options)
    rng = np.random.default_rng(0)
    solver.initialize(1.0, *(0.01 * rng.standard_normal(params.shape) for _ in range(dim)))
    probe = 2 * solver.step_multiple
    timed_run(solver, probe)
    per_step = timed_run(solver, probe) / probe
    steps = solver.step_multiple * max(1, round(target_seconds / per_step / solver.step_multiple))
    peak = copy_peak(torch)["peak"]
    seconds = timed_run(solver, steps)
    block, cap, state = solver.block, solver.groups_per_cu, solver.lattice_bytes
    del solver
    torch.cuda.empty_cache()
    nodes = size**dim
    mlups = nodes * steps / seconds / 1e6
    bandwidth = mlups * 1e6 * 2 * params.lattice.q * np.dtype(dtype).itemsize / 1e9
    return Measurement(backend, dtype, dim, size, steps, seconds, mlups, bandwidth, peak, bandwidth / peak,
                       state <= INFINITY_CACHE_BYTES, tuple(block), cap)


@dataclass(frozen=True)
class BodiesMeasurement:
    backend: str
    case: str
    nodes: int
    rules: int
    empty_mlups: float
    bodies_mlups: float
    ratio: float


def bodies_case(size: int, dim: int, dtype: str = "float32"):
    """(name, params, initial x velocity) of the with-body lattice of size ``size``: a
    Bouzidi cylinder (radius n/16 at (n/4, n/2)) in an n x n channel, or a Bouzidi sphere
    (radius 12 n/512 at (n/4, n/8, n/8)) in an n x n/4 x n/4 duct, each with a parabolic
    Zou-He inlet, a pressure outlet and no-slip side walls, started from the inflow. The
    Reynolds number grows with the size (20 at 1024^2 and at 512 x 128 x 128) so that tau is
    the same at every size: 1.46 in 2D, 0.62 in 3D."""
    if dim == 2:
        from lbm.cases.cylinder_benchmark import drag_channel

        params, _, profile = drag_channel(reynolds=20.0 * size / 1024, nx=size, ny=size, radius=size / 16,
                                          centre=(size / 4, size / 2), dtype=dtype)
        return f"channel {size}x{size}", params, np.where(params.solid, 0.0, profile[:, None])
    from lbm.cases.sphere_drag import duct_sphere

    side = size // 4
    params, _, inflow = duct_sphere(reynolds=20.0 * size / 512, nx=size, ny=side, nz=side,
                                    radius=12.0 * size / 512, centre=(size / 4, side / 2, side / 2), dtype=dtype)
    return f"duct {size}x{side}x{side}", params, np.where(params.solid, 0.0, inflow[0][:, :, None])


def _rate(torch, params, backend: str, ux, target_seconds: float) -> float:
    solver = make_solver(params, backend=backend)
    solver.initialize(1.0, ux, *([0.0] * (params.d - 1)))
    probe = 2 * solver.step_multiple
    timed_run(solver, probe)
    per_step = timed_run(solver, probe) / probe
    steps = solver.step_multiple * max(1, round(target_seconds / per_step / solver.step_multiple))
    best = max(params.n_nodes * steps / timed_run(solver, steps) / 1e6 for _ in range(2))
    del solver
    torch.cuda.empty_cache()
    return best


def measure_bodies(torch, backend: str, size: int, dim: int, dtype: str = "float32",
                   target_seconds: float = TARGET_SECONDS) -> BodiesMeasurement:
    """MLUPS of the with-body lattice of ``bodies_case`` (best of two timed runs, every node
    counted, rule pre-pass included) and of the fully periodic lattice of the same size."""
    name, params, ux = bodies_case(size, dim, dtype)
    sizes = {"nx": params.nx, "ny": params.ny} | ({"nz": params.nz} if dim == 3 else {})
    empty = SimulationParams(**sizes, tau=params.tau, collision=params.collision, dtype=dtype)
    e = _rate(torch, empty, backend, 0.0, target_seconds)
    b = _rate(torch, params, backend, ux, target_seconds)
    return BodiesMeasurement(backend, name, params.n_nodes, len(params.link_rules), e, b, b / e)


def main(argv=None) -> None:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--backends", nargs="+", default=list(BACKENDS), choices=BACKENDS)
    parser.add_argument("--dim", type=int, default=2, choices=(2, 3))
    parser.add_argument("--sizes", type=int, nargs="+", default=None)
    parser.add_argument("--dtype", default="float32", choices=("float32", "float64"))
    parser.add_argument("--seconds", type=float, default=TARGET_SECONDS, help="length of each timed run")
    parser.add_argument("--csv", default=None, help="write the table to this CSV file")
    parser.add_argument("--bodies", action="store_true",
                        help="time lattices with a body and open ends against the same lattices empty")
    args = parser.parse_args(argv)
    if args.sizes is None and args.bodies:
        args.sizes = [1024, 4096] if args.dim == 2 else [512]
    if args.sizes is None:
        args.sizes = [256, 512, 1024, 2048, 4096] if args.dim == 2 else [64, 96, 128, 192, 256]
    backends = args.backends
    if args.dim == 3:
        skipped = [b for b in backends if b.removeprefix("hip-") in TWO_D_ONLY]
        if skipped:
            print(f"skipping 2D-only backends in 3D: {skipped}")
        backends = [b for b in backends if b not in skipped]
    if not available_backends():
        raise SystemExit("no HIP backend is available on this machine (library or ROCm device missing)")

    import torch

    if args.bodies:
        return _main_bodies(torch, args, backends)
    print(f"{torch.cuda.get_device_name(0)}: {'D2Q9' if args.dim == 2 else 'D3Q19'} {args.dtype}, "
          f"Infinity Cache {INFINITY_CACHE_BYTES // 2**20} MiB")
    tiny = make_solver(grid_params(8, args.dim, args.dtype), backend=backends[0])
    for name, res in tiny.kernel_resources().items():
        print(f"  {name:>22}: {res['registers']:3d} VGPRs/lane, {res['local_bytes']} B scratch (spill), "
              f"{res['shared_bytes']} B LDS/group")
    del tiny
    results = []
    print(f"{'backend':>9} {'N':>6} {'MLUPS':>9} {'GB/s':>8} {'peak':>7} {'% peak':>7} {'block':>9} cap  note")
    for backend in backends:
        for size in args.sizes:
            try:
                m = measure(torch, backend, size, args.dim, args.dtype, args.seconds)
            except torch.cuda.OutOfMemoryError:
                print(f"{backend:>9} {size:6d}  out of device memory")
                torch.cuda.empty_cache()
                continue
            results.append(m)
            note = "state fits in the Infinity Cache" if m.fits_in_cache else ""
            print(f"{m.backend:>9} {m.size:6d} {m.mlups:9.0f} {m.bandwidth_gbs:8.0f} {m.peak_gbs:7.0f} "
                  f"{100 * m.fraction_of_peak:6.1f}% {m.block[0]:>4}x{m.block[1]:<4} {m.groups_per_cu:>3}  {note}",
                  flush=True)
    if args.csv and results:
        with open(args.csv, "w", newline="") as handle:
            writer = csv.DictWriter(handle, fieldnames=list(asdict(results[0])))
            writer.writeheader()
            writer.writerows(asdict(r) for r in results)
        print(f"table written to {args.csv}")


def _main_bodies(torch, args, backends) -> None:
    print(f"{torch.cuda.get_device_name(0)}: {'D2Q9' if args.dim == 2 else 'D3Q19'} {args.dtype}, "
          "with a body and open ends against empty")
    print(f"{'backend':>9} {'case':>22} {'rules':>8} {'empty':>9} {'bodies':>9} {'ratio':>6}")
    results = []
    for backend in backends:
        for size in args.sizes:
            m = measure_bodies(torch, backend, size, args.dim, args.dtype, args.seconds)
            results.append(m)
            print(f"{m.backend:>9} {m.case:>22} {m.rules:8d} {m.empty_mlups:9.0f} {m.bodies_mlups:9.0f} "
                  f"{m.ratio:6.3f}", flush=True)
    if args.csv and results:
        with open(args.csv, "w", newline="") as handle:
            writer = csv.DictWriter(handle, fieldnames=list(asdict(results[0])))
            writer.writeheader()
            writer.writerows(asdict(r) for r in results)
        print(f"table written to {args.csv}")


if __name__ == "__main__":
    main()
