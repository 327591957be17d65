"""Write the golden fixtures of tests/test_golden.py from the reference's own NumPy solver.

Run in the development container only (the reference tree is not shipped):

    python tests/golden/make_golden.py /path/to/reference/src

It imports the reference package ``lbm`` from the given ``src`` directory (and
nothing from this repository), runs small cases on its NumPy backend, and
stores inputs and outputs as compressed .npz data next to this script:

* ``states.npz``   populations after 40 steps of the mixed wall/lid/obstacle/force
                   geometries (tests/test_cuda_backend.py:21-84 of the reference),
                   2D 45x29 and 3D 12x10x8, BGK and TRT, float64 and float32;
* ``zou_he.npz``   the Zou-He/Hecht-Harting rule rows of a 2D and a 3D open channel
                   (open_boundaries.py:96-186), and the populations after 40 steps;
* ``bouzidi.npz``  Bouzidi rule rows, 40-step populations and the momentum-exchange
                   force (ForceProbe) of a 2D disc and a 3D sphere case;
* ``ibm.npz``      immersed-boundary populations and body force after 40 steps and the
                   force field of the first step, 2D and 3D.
"""

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
STEPS = 40


def main(src: str) -> None:
    sys.path.insert(0, src)
    from lbm import SimulationParams, make_solver
    from lbm.boundaries import channel_walls
    from lbm.forces import ForceProbe
    from lbm.geometry import circle, sphere, voxelize, wall_links
    from lbm.ibm import ImmersedBoundary
    from lbm.lattice import default_lattice
    from lbm.link_rules import bouzidi_rules
    from lbm.open_boundaries import open_channel

    def random_state(params, seed=0):
        rng = np.random.default_rng(seed)
        rho = 1 + 0.01 * rng.standard_normal(params.shape)
        u = 0.02 * rng.standard_normal((params.d, *params.shape))
        return rho, *u

    def run(params, seed=0, steps=STEPS):
        s = make_solver(params, backend="numpy")
        s.initialize(*random_state(params, seed))
        s.run(steps)
        return s

    # ---- mixed geometries
    states = {}
    for d in (2, 3):
        if d == 2:
            nx, ny, nz = 45, 29, None
            solid = channel_walls(nx, ny)
            solid[10:16, 12:19] = True
            wall = np.zeros((2, ny, nx))
            wall[0, -1, :] = 0.05
            force = (1e-5, -2e-6)
        else:
            nx, ny, nz = 12, 10, 8
            solid = channel_walls(nx, ny, nz)
            solid[0] = solid[-1] = True
            solid[3:5, 3:6, 4:8] = True
            wall = np.zeros((3, nz, ny, nx))
            wall[0, :, -1, :] = 0.05
            wall[2, :, -1, :] = -0.02
            force = (1e-5, -2e-6, 3e-6)
        states[f"solid_{d}d"] = solid
        states[f"wall_{d}d"] = wall
        for collision in ("bgk", "trt"):
            for dtype in ("float64", "float32"):
                params = SimulationParams(nx=nx, ny=ny, nz=nz, tau=0.7, collision=collision, force=force,
                                          solid=solid, wall_velocity=wall, dtype=dtype)
                states[f"f_{d}d_{collision}_{dtype}"] = run(params).populations()
    np.savez_compressed(HERE / "states.npz", **states)

    # ---- Zou-He open channels (tests/test_open_boundaries.py:133-146)
    zou = {}
    for d in (2, 3):
        lattice = default_lattice(d)
        nx, ny, nz = 12, 9, 5
        walls = channel_walls(nx, ny) if d == 2 else channel_walls(nx, ny, nz)
        inflow = np.zeros((d, *walls.shape[:-1]))
        inflow[0], inflow[1] = 0.04, 0.01
        if d == 3:
            inflow[2] = -0.02
        solid, rules = open_channel(lattice, walls, inflow, rho_out=1.003)
        v = rules.values
        zou.update({f"node_{d}d": rules.node, f"direction_{d}d": rules.direction, f"vdir_{d}d": v.direction,
                    f"vnode_{d}d": v.node, f"coeff_{d}d": v.coeff, f"constant_{d}d": v.constant})
        params = SimulationParams(nx=nx, ny=ny, nz=nz if d == 3 else None, tau=0.7, collision="trt", solid=solid,
                                  dtype="float64", link_rules=rules)
        zou[f"f_{d}d"] = run(params, seed=1).populations()
    np.savez_compressed(HERE / "zou_he.npz", **zou)

    # ---- Bouzidi + force probe (tests/test_geometry_link_rules.py:280-320)
    bz = {}
    for d in (2, 3):
        lattice = default_lattice(d)
        if d == 2:
            nx, ny, nz = 45, 29, None
            solid, body = channel_walls(nx, ny), circle(20.3, 13.6, 5.2)
            wall = np.zeros((2, ny, nx))
            wall[0, -1, :] = 0.05
            force = (1e-5, -2e-6)
        else:
            nx, ny, nz = 14, 12, 10
            solid, body = channel_walls(nx, ny, nz), sphere(6.3, 5.6, 4.8, 2.6)
            solid[0] = solid[-1] = True
            wall = np.zeros((3, nz, ny, nx))
            wall[0, :, -1, :] = 0.05
            wall[2, :, -1, :] = -0.02
            force = (1e-5, -2e-6, 3e-6)
        solid = solid | voxelize(body, solid.shape)
        links = wall_links(lattice, body, solid)
        rules = bouzidi_rules(lattice, links, solid)
        for dtype in ("float64", "float32"):
            params = SimulationParams(nx=nx, ny=ny, nz=nz, tau=0.7, collision="trt", force=force, solid=solid,
                                      wall_velocity=wall, dtype=dtype, link_rules=rules)
            s = run(params)
            bz[f"f_{d}d_{dtype}"] = s.populations()
            bz[f"force_{d}d_{dtype}"] = ForceProbe.build(params, links).evaluate(s)
        bz[f"link_q_{d}d"] = links.q
        bz[f"coeff_{d}d"] = rules.values.coeff
    np.savez_compressed(HERE / "bouzidi.npz", **bz)

    # ---- IBM (tests/test_ibm.py:59-89)
    ib = {}
    for d in (2, 3):
        if d == 2:
            nx, ny, nz = 45, 29, None
            solid, body = channel_walls(45, 29), circle(20.3, 13.6, 5.2)
            wall, force = np.zeros((2, 29, 45)), (1e-5, -2e-6)
            wall[0, -1] = 0.05
        else:
            nx, ny, nz = 16, 14, 12
            solid, body = channel_walls(16, 14, 12), sphere(7.3, 6.6, 5.8, 2.4)
            wall, force = np.zeros((3, 12, 14, 16)), (1e-5, -2e-6, 3e-6)
            wall[0, :, -1] = 0.05
        for dtype in ("float64", "float32"):
            params = SimulationParams(nx=nx, ny=ny, nz=nz, tau=0.7, collision="trt", force=force, solid=solid,
                                      wall_velocity=wall, dtype=dtype, immersed=ImmersedBoundary.from_ball(body))
            first = make_solver(params, backend="numpy")
            first.initialize(*random_state(params))
            first.run(1)
            ib[f"field1_{d}d_{dtype}"] = np.array(first._field)
            s = run(params)
            ib[f"f_{d}d_{dtype}"] = s.populations()
            ib[f"body_force_{d}d_{dtype}"] = s.immersed_force
    np.savez_compressed(HERE / "ibm.npz", **ib)
    for name in ("states", "zou_he", "bouzidi", "ibm"):
        print(name, (HERE / f"{name}.npz").stat().st_size, "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
