"""The C-ABI library: it loads without a GPU, exports every entry point the header declares,
its host-side lattice tables match the Python ones, and argument errors surface as ValueError
before any GPU work (tests/test_cuda_backend.py:95-154 of the reference check the same for
their NVRTC build)."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from lbm import BACKENDS, SimulationParams, make_solver
from lbm.backends import BackendUnavailable, gpu_available
from lbm.backends._native import SIGNATURES, library
from lbm.backends.hip_solver import VARIANTS, HipSolver, check_block, check_variant
from lbm.lattice import LATTICES

HEADER = Path(__file__).resolve().parent.parent / "include" / "lbm_hip.h"


def declared_entry_points():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"LBM_API\s+[\w\s\*]+?\b(lbm_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_entry_points()
    assert "lbm_pull_step" in names and "lbm_aa_odd" in names and "lbm_rules_apply" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = library()
    missing = [name for name in declared_entry_points() if not lib.has(name)]
    assert not missing, missing


def test_every_declared_entry_point_has_a_binding():
    bound = set(SIGNATURES) | {"lbm_last_error", "lbm_version"}
    assert set(declared_entry_points()) <= bound


@pytest.mark.parametrize("vs", list(LATTICES.values()), ids=list(LATTICES))
def test_compiled_lattice_tables_match_python(vs):
    q = ctypes.c_int()
    c = (ctypes.c_int * 57)()
    opp = (ctypes.c_int * 19)()
    w = (ctypes.c_double * 19)()
    library().call("lbm_lattice_tables", vs.code, ctypes.byref(q), c, opp, w)
    assert q.value == vs.q
    table = np.array(c[: 3 * vs.q]).reshape(3, vs.q)
    for axis in range(3):
        np.testing.assert_array_equal(table[axis], vs.component(axis))
    assert tuple(opp[: vs.q]) == vs.opposite
    assert tuple(w[: vs.q]) == vs.w


def test_bad_codes_raise_value_error_with_a_message():
    q = ctypes.c_int()
    c = (ctypes.c_int * 57)()
    opp = (ctypes.c_int * 19)()
    w = (ctypes.c_double * 19)()
    with pytest.raises(ValueError, match="lattice"):
        library().call("lbm_lattice_tables", 7, ctypes.byref(q), c, opp, w)


def test_invalid_block_and_grid_are_rejected_by_the_library_before_launch():
    lib = library()
    args = (0, 0, 1, 2, None, None, None, None, 8, 8, 1, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 0)
    with pytest.raises(ValueError, match="block"):
        lib.call("lbm_pull_step", *args, 64, 32, 0, None)
    with pytest.raises(ValueError, match="nz"):
        lib.call("lbm_pull_step", 0, 0, 1, 2, None, None, None, None, 8, 8, 4, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 0, 32, 8, 0, None)
    with pytest.raises(ValueError, match="2D only"):
        lib.call("lbm_lds_step", 1, 0, 1, 2, None, None, None, None, 8, 8, 8, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 0, 32, 8, 0, None)
    with pytest.raises(ValueError, match="terms"):
        lib.call("lbm_rules_apply", 0, 1, 1, 1, 1, 1, 5, 200, None, None)
    with pytest.raises(ValueError, match="terms"):
        lib.call("lbm_rules_apply_direct", 0, 1, 1, 1, 1, None, 1, 5, 0, None)
    with pytest.raises(ValueError, match="null pointer"):
        lib.call("lbm_rules_apply_direct", 0, 1, 1, 1, 1, None, None, 5, 4, None)
    # float64 D3Q19 kernels are compiled for at most 512 lanes per workgroup
    d3 = (1, 1, 1, 2, None, None, None, None, 8, 8, 8, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 0)
    with pytest.raises(ValueError, match="512 threads for float64 D3Q19"):
        lib.call("lbm_pull_step", *d3, 256, 4, 0, None)
    with pytest.raises(ValueError, match="512 threads for float64 D3Q19"):
        lib.call("lbm_aa_odd", 1, 1, 1, None, None, None, None, 8, 8, 8, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 0, 1024, 1, 0,
                 None)


def test_unknown_variant_is_rejected_before_touching_the_gpu():
    params = SimulationParams(nx=4, ny=4, tau=0.8)
    with pytest.raises(ValueError, match="variant"):
        HipSolver(params, variant="warp-specialised")
    with pytest.raises(ValueError, match="variant"):
        check_variant("bogus", 2)


def test_lds_variant_refuses_3d():
    with pytest.raises(ValueError, match="2D only"):
        HipSolver(SimulationParams(nx=8, ny=8, nz=8, tau=0.8), variant="lds")


def test_block_validation():
    assert check_block((32, 8)) == (32, 8)
    with pytest.raises(ValueError, match="block"):
        check_block((64, 32))
    with pytest.raises(ValueError, match="block"):
        check_block((0, 4))


def test_unknown_backend_is_rejected():
    with pytest.raises(ValueError, match="backend"):
        make_solver(SimulationParams(nx=8, ny=8, tau=0.8), backend="opencl")


def test_reference_backend_names_are_accepted_aliases():
    from lbm.backends import ALIASES

    assert set(ALIASES.values()) <= set(BACKENDS)
    assert {v.removeprefix("hip-") for v in BACKENDS} == set(VARIANTS)


@pytest.mark.skipif(gpu_available(), reason="checks the error raised on machines without a GPU")
def test_hip_backend_without_gpu_raises_clear_error():
    with pytest.raises(BackendUnavailable):
        make_solver(SimulationParams(nx=8, ny=8, tau=0.8), backend="hip-pull")


def test_cpu_backends_are_not_part_of_the_package():
    with pytest.raises(BackendUnavailable, match="HIP only"):
        make_solver(SimulationParams(nx=8, ny=8, tau=0.8), backend="numpy")


def test_force_scratch_size_covers_the_partial_sums():
    import ctypes

    from lbm.backends._native import library

    lib = library()
    size = ctypes.c_int64()
    lib.call("lbm_affine_force_scratch_size", 2048, ctypes.byref(size))
    assert size.value == 0  # one workgroup writes the result directly
    lib.call("lbm_affine_force_scratch_size", 2049, ctypes.byref(size))
    assert size.value == 3 * 8 * 3  # one group per 1024 rows
    lib.call("lbm_affine_force_scratch_size", 10**7, ctypes.byref(size))
    assert size.value == 3 * 8 * 1024
    with pytest.raises(ValueError, match="bad arguments"):
        lib.call("lbm_affine_force_scratch_size", -1, ctypes.byref(size))


def test_fused_force_and_pre_pass_refuse_sizes_beyond_one_workgroup():
    lib = library()
    common = (0, 1, 1, 1, 1, 1)
    with pytest.raises(ValueError, match="2048 rows"):
        lib.call("lbm_affine_force_rules", *common, 2049, 4, 2, 1, 1, 1, 1, 1, 10, 4, None)
    with pytest.raises(ValueError, match="4096 rules"):
        lib.call("lbm_affine_force_rules", *common, 100, 4, 2, 1, 1, 1, 1, 1, 4097, 4, None)


def test_batched_descriptors_lay_out_the_padded_pools_the_host_expects():
    """lbm_batched_describe (host only) pads every lattice's population planes to the
    stride lbm.batched.plane_stride computes, and places the lattices back to back: the
    library's descriptors and the Python pool views agree on every offset."""
    import struct

    from lbm.batched import PLANE_ALIGN, plane_stride

    lib = library()
    size = ctypes.c_int64()
    lib.call("lbm_batched_desc_size", 0, ctypes.byref(size))
    sizes = [(40, 30), (77, 300), (64, 64), (1000, 3), (33, 1)]
    nx = np.array([s[0] for s in sizes], dtype=np.int32)
    ny = np.array([s[1] for s in sizes], dtype=np.int32)
    nz = np.ones(len(sizes), dtype=np.int32)
    wp, wm = np.full(len(sizes), 1.25), np.full(len(sizes), 1.1)
    force = np.zeros(3 * len(sizes))
    raw = np.zeros(len(sizes) * size.value, dtype=np.uint8)
    lib.call("lbm_batched_describe", 0, 0, len(sizes), nx.ctypes.data, ny.ctypes.data, nz.ctypes.data,
             wp.ctypes.data, wm.ctypes.data, force.ctypes.data, raw.ctypes.data)
    data = raw.tobytes()
    f = kind = 0
    for lat, (x, y) in enumerate(sizes):
        d = data[lat * size.value : (lat + 1) * size.value]
        n, ns = struct.unpack_from("<II", d, 12)  # Grid {nx, ny, nz, n, ns}
        f_offset, kind_offset = struct.unpack_from("<QQ", d, 24)
        assert n == x * y and ns == plane_stride(n) and ns % PLANE_ALIGN == 0 and ns - n < PLANE_ALIGN
        assert (f_offset, kind_offset) == (f, kind)
        f += 9 * ns
        kind += n


def test_grids_past_two_to_the_31_nodes_and_empty_rule_sets_are_handled_before_any_gpu_work():
    lib = library()
    huge = (1, 0, None, None, None, None, None, None, 2048, 2048, 2048, 1.0, 1.0, 0.0, 0.0, 0.0, 0, 0, 256, 1, 0, None)
    with pytest.raises(ValueError, match=r"2\^31 nodes"):
        lib.call("lbm_pull_step", *huge)
    # no rules: nothing to read, nothing launched, success
    lib.call("lbm_rules_apply", 0, None, None, None, None, None, 0, 4, None, None)
    lib.call("lbm_rules_apply_direct", 0, None, None, None, None, None, None, 0, 4, None)
