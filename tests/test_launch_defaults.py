"""Default launch shapes (host logic, no GPU): the measured D3Q19 shapes per variant and
precision, D2Q9 row-wide workgroups, the occupancy-cap rule, and the lane limits the
kernels are compiled for."""

import numpy as np
import pytest

from lbm import SimulationParams
from lbm.backends.hip_solver import default_block, default_groups_per_cu


def box(n, d, dtype="float32", solid=False):
    sizes = {"nx": n, "ny": n} | ({"nz": n} if d == 3 else {})
    shape = (n,) * d
    mask = np.zeros(shape, dtype=bool)
    if solid:
        mask[(0,) * d] = True
    return SimulationParams(**sizes, tau=0.6, dtype=dtype, solid=mask)


@pytest.mark.parametrize(("variant", "dtype", "block", "cap"), [
    ("aa", "float32", (256, 2), 2), ("pull", "float32", (256, 1), 4),
    ("aa", "float64", (256, 1), 2), ("pull", "float64", (256, 1), 2),
])
def test_d3q19_shapes_at_256(variant, dtype, block, cap):
    p = box(256, 3, dtype)
    assert default_block(p, variant) == block
    assert default_groups_per_cu(p, variant) == cap


@pytest.mark.parametrize(("variant", "dtype", "block", "cap"), [
    ("aa", "float32", (256, 1), 4), ("pull", "float32", (256, 1), 4),
    ("aa", "float64", (256, 1), 2), ("pull", "float64", (256, 1), 2),
])
def test_d3q19_shapes_with_solid_nodes(variant, dtype, block, cap):
    p = box(256, 3, dtype, solid=True)
    assert default_block(p, variant) == block
    assert default_groups_per_cu(p, variant) == cap


def test_d2q9_rows_and_the_cap_rule():
    big = box(4096, 2)  # 604 MB (AA) / 1.2 GB (pull): beyond twice the Infinity Cache
    assert default_block(big, "pull") == (1024, 1) and default_groups_per_cu(big, "pull") == 1
    assert default_groups_per_cu(big, "aa") == 1
    mid = box(2048, 2)  # 151 MB / 302 MB: uncapped
    assert default_groups_per_cu(mid, "aa") == 0 and default_groups_per_cu(mid, "pull") == 0
    walls = box(4096, 2, solid=True)  # masked lattices run uncapped, in rows of 512 lanes
    assert default_groups_per_cu(walls, "pull") == 0
    assert default_block(walls, "pull") == default_block(walls, "aa") == (512, 1)
    assert default_block(walls, "lds") == (256, 4)
    assert default_block(big, "lds") == (256, 4) and default_groups_per_cu(big, "lds") == 0
    small = box(512, 2, solid=True)  # masked up to 2^18 nodes: rows of 512 lanes, two per group
    assert default_block(small, "pull") == default_block(small, "aa") == (512, 2)
    channel = SimulationParams(nx=880, ny=164, tau=0.6, solid=np.zeros((164, 880), dtype=bool) | (np.arange(164) == 0)[:, None])
    assert default_block(channel, "pull") == (448, 2)  # two even 448-lane tiles per row
    assert default_block(box(1024, 2, solid=True), "aa") == (512, 1)
    narrow = SimulationParams(nx=45, ny=29, tau=0.6)
    assert default_block(narrow, "pull") == (64, 16)


@pytest.mark.parametrize("n", [8, 20, 64, 100, 256, 300])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_default_blocks_respect_the_kernel_lane_limits(n, dtype):
    for d in (2, 3):
        p = box(n, d, dtype)
        for variant in ("pull", "aa") + (("lds",) if d == 2 else ()):
            bx, by = default_block(p, variant)
            limit = 512 if (d == 3 and dtype == "float64") else 1024
            assert bx % 64 == 0 and 1 <= bx * by <= limit


@pytest.mark.parametrize(("nx", "widest", "width"), [
    (384, 256, 192), (320, 256, 192), (256, 256, 256), (512, 256, 256), (640, 256, 256),
    (100, 1024, 128), (1536, 1024, 768), (5120, 1024, 1024), (1100, 1024, 576),
])
def test_rows_split_into_even_tiles(nx, widest, width):
    from lbm.backends.hip_solver import _even_tiles

    bx = _even_tiles(nx, widest)
    assert bx == width and bx % 64 == 0 and bx <= widest
    tiles = -(-nx // bx)
    assert tiles == -(-nx // widest)  # no more tiles than the widest width needs
    assert tiles * bx - nx < 64 * tiles  # less than a wavefront idle per tile


def test_long_rows_take_even_tiles():
    p = box(384, 3)
    assert default_block(p, "aa") == (192, 2) and default_block(p, "pull") == (192, 1)
    walls = box(384, 3, solid=True)
    assert default_block(walls, "aa") == default_block(walls, "pull") == (192, 1)
    assert default_block(box(1536, 2), "pull") == (768, 1)
    assert default_block(box(1536, 2, solid=True), "pull") == (512, 1)  # three 512-lane rows
    assert default_block(box(1200, 2, solid=True), "pull") == (448, 1)
