"""Throughput of the HIP lattice Boltzmann step on one MI355X, against the HBM roofline.

Workload (the headline configuration): D3Q19, float32, in-place AA pattern, 256^3
fully periodic box, TRT (Lambda = 3/16) with tau = 0.6, uniform Guo body force
(1e-6, 0, 0), initialised with the decaying ABC flow A = B = C = 0.02,
k = 2 pi / N. A "step" is one time step: every node streamed, bounced and
collided once (an AA pair of half-steps is two steps).

Protocol (as the reference's benchmarks/bench_mlups.py:104-154): warm-up
steps, re-measure the device-copy peak on 2 GiB buffers (above the 256 MB
Infinity Cache; the larger of the library's 16-byte copy kernel and the runtime's
device-to-device copy), then synchronize -> run K steps -> synchronize.
MLUPS = N * K / seconds / 1e6. Algorithmic bytes per node update are
2 * Q * sizeof(float) = 152 B (19 loads + 19 stores), so achieved GB/s =
MLUPS * 152 / 1e3. Each AA kernel launch updates all N nodes once.

The JSON line carries ``roofline`` (achieved GB/s of the dominant kernel from
HIP events around its launches on the launch stream, against the 8 TB/s HBM
spec, with ``traffic`` = measured HBM bytes per launch from rocprofv3 counters
when a summary is present under profiles/) and ``cpu_baseline`` (the NumPy
oracle timed on this host on a bounded sample of the same workload, 1 core).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "high-performance-lbm-on-gpu-_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS & %HBM peak, D3Q19 fp32 AA 256³; fwd+adjoint/fwd MLUPS ratio"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "traffic_d3q19_aa_256.json"


def abc_params(n: int, dtype: str = "float32", tau: float = 0.6):
    from lbm.cases.beltrami import abc_box_params

    return abc_box_params(n, dtype, tau)


def abc_fields(n: int, amplitude: float = 0.02, tau: float = 0.6):
    from lbm.cases.beltrami import exact_fields
    from lbm.lattice import CS2

    return exact_fields(n, amplitude, CS2 * (tau - 0.5), 0.0)


def copy_peak_gbs(torch) -> dict[str, float]:
    """The device-copy peak on 2 GiB buffers (``lbm.bench_mlups.copy_peak``)."""
    from lbm.bench_mlups import copy_peak

    return copy_peak(torch)


def kernel_times_ms(torch, solver, pairs: int = 20) -> dict[str, float]:
    """Average duration of each AA half-step kernel, from events around each launch."""
    stream = torch.cuda.current_stream()
    events = [torch.cuda.Event(enable_timing=True) for _ in range(2 * pairs + 1)]
    events[0].record(stream)
    for p in range(pairs):
        solver._pre_pass("aa_odd")
        solver._call("lbm_aa_odd", solver._lattice_code, solver._dtype_code, solver._f.data_ptr(),
                     solver._kind.data_ptr(), solver._wall.data_ptr(), None, None, *solver._step_args())
        events[2 * p + 1].record(stream)
        solver._pre_pass("aa_even")
        solver._call("lbm_aa_even", solver._lattice_code, solver._dtype_code, solver._f.data_ptr(),
                     solver._kind.data_ptr(), None, None, *solver._step_args())
        events[2 * p + 2].record(stream)
    events[-1].synchronize()
    odd = [events[2 * p].elapsed_time(events[2 * p + 1]) for p in range(pairs)]
    even = [events[2 * p + 1].elapsed_time(events[2 * p + 2]) for p in range(pairs)]
    return {"aa_odd": float(np.mean(odd)), "aa_even": float(np.mean(even))}


def cpu_baseline(n: int = 256, steps: int = 4, n2: int = 2048, steps2: int = 8) -> dict:
    """The NumPy oracle on a bounded sample of the headline workload: D3Q19 fp32 TRT + Guo ABC
    box at n^3, ``steps`` steps after one warm-up step (the first step also pays the page
    faults of NumPy's temporaries), plus the D2Q9 fp32 TRT periodic n2^2 lattice the
    reference quotes its NumPy figure on (README.md:108), ``steps2`` steps after one warm-up."""
    from lbm import SimulationParams
    from oracle import ReferenceSolver

    def timed(params, fields, count) -> tuple[float, float]:
        solver = ReferenceSolver(params)
        solver.initialize(*fields)
        solver.step(1)
        start = time.perf_counter()
        solver.step(count)
        seconds = time.perf_counter() - start
        return params.n_nodes * count / seconds / 1e6, seconds

    mlups, seconds = timed(abc_params(n), abc_fields(n), steps)
    flat = SimulationParams(nx=n2, ny=n2, tau=0.6, collision="trt", dtype="float32")
    mlups2, seconds2 = timed(flat, (), steps2)
    cores = len(os.sched_getaffinity(0))
    return {
        "value": round(mlups, 3),
        "unit": "MLUPS",
        "cores": 1,
        "kind": "port",
        "sample": f"NumPy oracle, D3Q19 fp32 TRT+Guo ABC periodic {n}^3, {steps} steps after 1 warm-up "
        f"({seconds:.1f} s; NumPy elementwise is single-threaded; {cores} cores visible)",
        "d2q9": {"value": round(mlups2, 3), "unit": "MLUPS",
                 "sample": f"D2Q9 fp32 TRT periodic {n2}^2, {steps2} steps after 1 warm-up ({seconds2:.1f} s)"},
    }


def adjoint_ratio(torch, steps: int = 4000, window: int = 2000, k: int = 100) -> dict:
    """Forward-only vs checkpointed forward+adjoint rate on the differentiable configuration:
    D2Q9 fp32 channel 512x128, Zou-He inlet/outlet, IBM cylinder R = 12 (markers 1 lattice
    unit apart, 3 multi-direct passes), Re = 100, body-force field N(0, 1e-7), loss = mean
    drag over the last ``window`` of ``steps`` steps, checkpoints every ``k`` steps."""
    from lbm.adjoint import DragAdjoint
    from lbm.cases.ibm_channel import ibm_channel_params, initial_fields

    params = ibm_channel_params()
    problem = DragAdjoint(params, steps=steps, window=window, checkpoint_every=k, init=initial_fields(params))
    g = np.random.default_rng(0).standard_normal((2, *params.shape)) * 1e-7 * params.fluid
    problem.set_body_force(g)
    problem.forward()
    problem.backward()  # first calls capture the graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    problem.run_forward(store=False)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = problem.forward()
    grad_g, grad_x = problem.backward()
    torch.cuda.synchronize()
    t_both = time.perf_counter() - t0
    nodes = params.nx * params.ny
    return {
        "workload": "D2Q9 fp32 IBM-cylinder channel 512x128, Re 100, 4000 steps, loss = mean drag of the "
        "last 2000, checkpoints every 100 steps",
        "forward_mlups": round(nodes * steps / t_fwd / 1e6, 1),
        "forward_adjoint_mlups": round(nodes * steps / t_both / 1e6, 1),
        "ratio": round(t_fwd / t_both, 4),
        "loss": loss,
        "grad_body_norm": float(np.linalg.norm(grad_g)),
        "grad_markers_norm": float(np.linalg.norm(grad_x)),
    }


def batched_measure(torch, count: int = 64, steps: int = 100) -> dict:
    """Aggregate MLUPS of 64 seeded D2Q9 fp32 Bouzidi-cylinder channels (H in [64, 512],
    L = 4H, Re in [20, 200], Zou-He inlet/outlet) advanced by one persistent launch per
    step (graph-captured with the rule pre-pass, autotuned tiles), against the same
    lattices each run alone with the pull solver (default tiles), each lattice timed with its
    launches replayed from a captured graph and launched eagerly from a prepared list, the
    faster of the two counted."""
    from lbm import make_solver
    from lbm.backends.hip_solver import GRAPH_PAIRS
    from lbm.batched import BatchedSolver
    from lbm.cases.batched_channels import batched_channels, initial_fields

    # untimed warm-up that also captures both graph lengths the timed run replays
    warm = 2 * (GRAPH_PAIRS + 1)

    t_build = time.perf_counter()
    params = batched_channels(count=count, seed=0)
    t_build = time.perf_counter() - t_build
    batch = BatchedSolver(params, autotune=True)
    for i, p in enumerate(params):
        batch.initialize(i, *initial_fields(p))
    batch.run(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    total = batch.total_nodes
    t_alone = 0.0
    for p in params:
        best = float("inf")
        for graphs in (True, False):
            solver = make_solver(p, backend="hip-pull", graphs=graphs)
            solver.initialize(*initial_fields(p))
            solver.run(warm, check_every=warm)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            solver._advance(steps)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
            del solver
        t_alone += best
    return {
        "workload": f"{count} D2Q9 fp32 Bouzidi-cylinder channels (seed 0), {total} nodes, {steps} steps",
        "nodes": total,
        "batched_mlups": round(total * steps / t_batch / 1e6, 1),
        "alone_mlups": round(total * steps / t_alone / 1e6, 1),
        "speedup": round(t_alone / t_batch, 3),
        "block": list(batch.block),
        "groups": batch.groups,
        "groups_per_cu": batch.groups_per_cu,
        "host_build_seconds": round(t_build, 1),
    }


def lds_compare(torch, peak_gbs: float, n: int = 4096, steps: int = 40) -> dict:
    """The LDS-tiled variant against the fused pull kernel, as the reference's benchmark
    compares them (D2Q9 fp32 periodic n^2, TRT tau = 0.6, 72 B per update)."""
    from lbm import SimulationParams, make_solver

    params = SimulationParams(nx=n, ny=n, tau=0.6, collision="trt", dtype="float32")
    out = {"workload": f"D2Q9 fp32 TRT periodic {n}^2, {steps} steps"}
    for variant in ("pull", "lds"):
        solver = make_solver(params, backend=f"hip-{variant}")
        solver._advance(4)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver._advance(steps)
        torch.cuda.synchronize()
        mlups = n * n * steps / (time.perf_counter() - t0) / 1e6
        out[variant] = {"mlups": round(mlups, 1), "fraction_of_copy_peak": round(mlups * 72 / 1e3 / peak_gbs, 4),
                        "block": list(solver.block)}
        del solver
    return out


def two_lattice_measure(torch, n: int, peak_gbs: float, steps: int = 200) -> dict:
    """The two-lattice (pull) kernel on the same workload, as the reference's benchmark times
    both layouts (152 B per update, two population buffers)."""
    from lbm import make_solver

    params = abc_params(n)
    solver = make_solver(params, backend="hip-pull")
    solver.initialize(*abc_fields(n))
    solver._advance(4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver._advance(steps)
    torch.cuda.synchronize()
    mlups = n**3 * steps / (time.perf_counter() - t0) / 1e6
    out = {"workload": f"D3Q19 fp32 TRT+Guo periodic ABC box {n}^3, two-lattice pull, {steps} steps",
           "mlups": round(mlups, 1), "fraction_of_copy_peak": round(mlups * 152 / 1e3 / peak_gbs, 4),
           "block": list(solver.block), "groups_per_cu": solver.groups_per_cu}
    del solver
    return out


def traffic_from_profile():
    if TRAFFIC_FILE.exists():
        data = json.loads(TRAFFIC_FILE.read_text())
        return data.get("hbm_bytes_per_launch")
    return None


def round_up(count: int, multiple: int) -> int:
    """``count`` rounded up to a multiple of ``multiple`` (the AA pattern advances in pairs)."""
    return -(-count // multiple) * multiple


def guarded(name: str, leg) -> dict:
    """Run one auxiliary measurement; a failure is reported in its field and never costs the
    headline record."""
    try:
        out = leg()
        print(f"# {name}: {json.dumps(out)}", flush=True)
        return out
    except Exception as exc:  # noqa: BLE001 - the error text goes into the record
        print(f"# {name} failed: {type(exc).__name__}: {exc}", file=sys.stderr, flush=True)
        try:
            import torch

            torch.cuda.synchronize()
            torch.cuda.empty_cache()
        except Exception:  # noqa: BLE001
            pass
        return {"error": f"{type(exc).__name__}: {exc}"}


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=1000)
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--size", type=int, default=256)
    parser.add_argument("--variant", default="aa", choices=("aa", "pull"))
    parser.add_argument("--block", type=int, nargs=2, default=None)
    parser.add_argument("--groups-per-cu", type=int, default=None, help="resident workgroups per CU (0: no cap)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-adjoint", action="store_true")
    parser.add_argument("--no-batched", action="store_true")
    parser.add_argument("--no-lds", action="store_true", help="skip the LDS-tiled vs fused pull comparison")
    parser.add_argument("--no-two-lattice", action="store_true", help="skip the two-lattice pull timing")
    args = parser.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (got --gpus {args.gpus})", file=sys.stderr)
        return 2
    if args.steps <= 0 or args.warmup < 0:
        print(f"need --steps > 0 and --warmup >= 0; got {args.steps}, {args.warmup}", file=sys.stderr)
        return 2

    import torch

    from lbm import make_solver

    n = args.size
    params = abc_params(n)
    options = {"block": tuple(args.block)} if args.block else {}
    if args.groups_per_cu is not None:
        options["groups_per_cu"] = args.groups_per_cu
    solver = make_solver(params, backend=f"hip-{args.variant}", **options)
    solver.initialize(*abc_fields(n))
    # The AA pattern advances in pairs of steps: odd counts are rounded up to the next pair,
    # as the reference's benchmark sizes its runs in multiples of step_multiple
    # (benchmarks/bench_mlups.py:127-130); the record reports the counts actually run.
    m = solver.step_multiple
    steps, warmup = round_up(args.steps, m), round_up(args.warmup, m)
    print(f"# {torch.cuda.get_device_name(0)}: D3Q19 fp32 {args.variant} {n}^3, block {solver.block}, "
          f"{solver.groups_per_cu or 'no'} resident groups/CU cap; {warmup} warm-up + {steps} timed steps "
          f"(asked for {args.warmup} + {args.steps})", flush=True)

    if warmup:
        solver.run(warmup, check_every=warmup)
    torch.cuda.synchronize()
    copies = copy_peak_gbs(torch)
    peak = copies["peak"]
    print(f"# device-copy peak {peak:.0f} GB/s (2 GiB buffers; copy kernel {copies['kernel']:.0f}, "
          f"runtime copy {copies['runtime']:.0f})", flush=True)

    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    start.record(stream)
    solver._advance(steps)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    gpu_seconds = start.elapsed_time(stop) / 1e3
    seconds = wall
    solver.time_step += steps
    solver._check_health()

    nodes = n**3
    mlups = nodes * steps / seconds / 1e6
    bytes_per_update = 2 * params.lattice.q * 4
    gbs = mlups * bytes_per_update / 1e3
    print(f"# timed {steps} steps: host wall {wall * 1e3:.2f} ms (events {gpu_seconds * 1e3:.2f} ms), "
          f"{mlups:.0f} MLUPS, {gbs:.0f} GB/s = {100 * gbs / peak:.1f}% of copy peak", flush=True)

    block, groups_per_cu = solver.block, solver.groups_per_cu
    times = {}
    if args.variant == "aa":
        times = kernel_times_ms(torch, solver)
        dominant = max(times, key=times.get)
        launch_ms = times[dominant]
    else:
        dominant, launch_ms = "pull_step", seconds * 1e3 / steps
    achieved = nodes * bytes_per_update / (launch_ms / 1e3) / 1e9
    print(f"# per-kernel ms: {json.dumps({k: round(v, 4) for k, v in times.items()})}", flush=True)
    del solver

    record = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "ms_per_step": round(seconds * 1e3 / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (ABC flow initial field)",
        "config": {
            "workload": f"D3Q19 fp32 TRT+Guo periodic ABC box {n}^3, {args.variant.upper()} pattern",
            "nodes": nodes,
            "bytes_per_update": bytes_per_update,
            "block": list(block),
            "groups_per_cu": groups_per_cu,
            "copy_peak_gbs": round(peak, 1),
            "copy_kernel_gbs": round(copies["kernel"], 1),
            "copy_runtime_gbs": round(copies["runtime"], 1),
            "fraction_of_copy_peak": round(gbs / peak, 4),
            "event_ms_per_step": round(gpu_seconds * 1e3 / steps, 5),
            "fits_in_infinity_cache": params.lattice.q * nodes * 4 <= 256 * 2**20,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"lbm_{dominant}",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile(),
            "launch_ms": round(launch_ms, 5),
            "bytes_per_launch": nodes * bytes_per_update,
        },
        "cpu_baseline": None,
        "adjoint": None,
        "batched": None,
        "lds_vs_pull": None,
        "two_lattice": None,
    }
    if not args.no_two_lattice and args.variant == "aa":
        record["two_lattice"] = guarded("two-lattice", lambda: two_lattice_measure(torch, n, peak))
    if not args.no_adjoint:
        record["adjoint"] = guarded("adjoint", lambda: adjoint_ratio(torch))
    if not args.no_lds:
        record["lds_vs_pull"] = guarded("lds vs pull", lambda: lds_compare(torch, peak))
    if not args.no_batched:
        record["batched"] = guarded("batched", lambda: batched_measure(torch))
    if not args.no_cpu_baseline:
        record["cpu_baseline"] = guarded("cpu baseline", cpu_baseline)
    print(json.dumps(record), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
