"""Launches of the grid-barrier kernels (the coloring fixpoint and the solver
passes). By default they are plain launches sized to what every CU's
occupancy holds at once; cooperative launches are the alternative. Both give
bit-identical worlds and batches. A barrier wait beyond its bound fails the
step, names the cause and turns the world to cooperative launches."""
import numpy as np
import pytest

from melon3d_amd import scenes
from parity import bodies_array

DT = 1.0 / 60.0


def _raft(dev):
    """The 15 x 15 raft of touching boxes on the ground with five boxes
    dropping onto it (test/test_main.cpp:960-985): one island of about 650
    contacts, solved on the grid path with greedy colors and overflow
    levels."""
    w = dev.create_world(dev.m3d_world_def_default())
    scenes.add_ground(dev, w)
    for i in range(15):
        for k in range(15):
            scenes.add_box(dev, w, (-7.0 + i, 0.5, -7.0 + k), (0.501, 0.5, 0.501))
    for i in range(5):
        scenes.add_box(dev, w, (2.0 * i - 4.0, 3.0 + 0.7 * i, 0.5 * i - 1.0), (0.45, 0.45, 0.45))
    return w


def _same_bodies(dev, a, b, what):
    sa, sb = bodies_array(dev.export_state(a)), bodies_array(dev.export_state(b))
    names = ("position", "rotation", "linear velocity", "angular velocity", "in use", "awake")
    for k, name in enumerate(names):
        assert np.array_equal(sa[k], sb[k]), f"{what}: {name} differs"


@pytest.mark.gpu
def test_plain_and_cooperative_launches_agree(dev):
    """The raft stepped 60 times with plain and with cooperative launches of
    the grid kernels: every body state bit-identical, every step succeeds,
    and each world keeps the mode it was given."""
    worlds = []
    try:
        for coop in (0, 1):
            w = _raft(dev)
            assert dev.m3dx_set_grid_launch(w, coop, 0) == 0
            assert dev.m3dx_grid_launch_mode(w) == coop
            worlds.append(w)
        colored = 0
        for _ in range(60):
            for w in worlds:
                dev.m3d_world_step(w, DT, 4)
            colored += dev.stats(worlds[0]).coloredIslands
        assert [dev.m3dx_step_failures(w) for w in worlds] == [0, 0], dev.last_error()
        assert [dev.m3dx_grid_launch_mode(w) for w in worlds] == [0, 1]
        assert colored > 0
        _same_bodies(dev, worlds[0], worlds[1], "raft after 60 steps")
    finally:
        for w in worlds:
            dev.m3d_world_destroy(w)


@pytest.mark.gpu
def test_plain_and_cooperative_launches_agree_on_batches(dev):
    """A batch of 48 worlds (piles and pyramids, the substep loop captured
    as a graph once its workgroup count is settled) stepped 30 times with
    plain and with cooperative launches: bit-identical bodies."""
    S, steps = 48, 30
    worlds = []
    try:
        for coop in (0, 1):
            bw, _ = scenes.batch(dev, S, first=300)
            assert dev.m3dx_set_grid_launch(bw, coop, 0) == 0
            worlds.append(bw)
        for _ in range(steps):
            for bw in worlds:
                dev.m3d_world_step(bw, DT, 4)
        assert [dev.m3dx_step_failures(bw) for bw in worlds] == [0, 0], dev.last_error()
        assert dev.m3dx_step_graph_launches(worlds[0]) > 0
        _same_bodies(dev, worlds[0], worlds[1], f"batch after {steps} steps")
    finally:
        for bw in worlds:
            dev.m3d_world_destroy(bw)


@pytest.mark.gpu
def test_barrier_wait_bound_fails_the_step(dev):
    """A bound of one wall-clock tick on every grid-barrier wait, set from the
    host: the first workgroup to wait gives up, sets the step's abort word,
    and every later wait of the step returns at once, so every launch drains.
    The step is counted as failed, the error names the barrier, and the world
    launches its grid kernels cooperatively from then on."""
    w = _raft(dev)
    try:
        assert dev.m3dx_set_grid_launch(w, 0, 1) == 0
        assert dev.m3dx_grid_launch_mode(w) == 0
        dev.m3d_world_step(w, DT, 4)
        assert dev.m3dx_step_failures(w) == 1
        assert "grid barrier" in dev.last_error()
        assert dev.m3dx_grid_launch_mode(w) == 1
    finally:
        dev.m3d_world_destroy(w)

