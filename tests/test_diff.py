"""Adjoint of the XPBD solve (differentiable worlds).

The adjoint is pinned three ways: the float replay it differentiates equals
the engine's own solve (so the gradient is the engine's, step for step); its
directional derivatives equal FP64 central differences of the double replay
(fixed contact sets, h = 1e-4 relative); and the torch.autograd.Function
returns the same gradients as the C ABI.
"""
import numpy as np
import pytest

from melon3d_amd import diff

DT = 1.0 / 60.0


def _record(dev, scene, steps):
    w = scene.build(dev)
    assert dev.m3dx_diff_record(w, 1) == 0, dev.last_error()
    for _ in range(steps):
        dev.m3d_world_step(w, DT, 4)
    dev.m3dx_diff_record(w, 0)
    assert dev.m3dx_diff_steps(w) == steps
    return w


@pytest.mark.gpu
def test_record_needs_sleep_off(dev):
    w = dev.create_world()
    try:
        assert dev.m3dx_diff_record(w, 1) == -1
        assert "enableSleep" in dev.last_error()
    finally:
        dev.m3d_world_destroy(w)


@pytest.mark.gpu
def test_float_replay_is_the_engine(dev):
    """Replaying the recorded steps in float reproduces the engine's states:
    each step from its recorded start equals the next recorded start, and the
    chained replay equals the engine's final state."""
    scene = diff.DiffScene(worlds=16)
    steps = 40
    w = _record(dev, scene, steps)
    try:
        import ctypes as C
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        n = scene.body_count()
        idx = scene.dynamic_index().ravel()
        for t in range(steps - 1):
            out = np.zeros((n, 13), dtype=np.float32)
            nxt = np.zeros((n, 13), dtype=np.float32)
            assert dev.m3dx_diff_replay(w, t, 1, 0, None, None, None, ptr(out)) == 0
            assert dev.m3dx_diff_state(w, t + 1, ptr(nxt)) >= 0
            assert np.array_equal(out[idx], nxt[idx]), t
        final = diff.replay(dev, w, scene, steps, precision=0, recorded=True)
        engine = diff.final_positions(dev, w, scene)
        assert np.array_equal(final[scene.dynamic_index(), 0:3].astype(np.float64), engine)
    finally:
        dev.m3d_world_destroy(w)


def _per_world_fd(dev, w, scene, steps, dv, df, dm, h):
    """Central differences of each world's loss along (dv, df, dm): the
    worlds are independent, so one double replay per offset serves all."""
    def losses(eps):
        st = diff.replay(dev, w, scene, steps, precision=1, v0=scene.v0 + eps * dv,
                         friction=scene.friction + eps * df, mass=scene.mass + eps * dm)
        pos = st[scene.dynamic_index(), 0:3]
        mm = (scene.mass + eps * dm)[..., None]
        com = (mm * pos).sum(1) / mm.sum(1)
        return ((com - scene.target) ** 2).sum(1)
    return (losses(h) - losses(-h)) / (2.0 * h)


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [3, 10])
def test_gradient_matches_fp64_central_differences(dev, steps):
    """Per-world directional derivatives of the loss along random directions
    in (v0, friction, mass): adjoint vs FP64 central differences of the
    double replay (fixed contact sets), rel err <= 1e-3. The loss is
    piecewise smooth (contact activation, the friction cone, the static /
    dynamic friction switch and re-anchoring are branches); a world counts
    only where central differences at h = 1e-4 and 5e-5 agree to 1e-4, i.e.
    no branch flips inside the stencil. At 3 steps every world qualifies."""
    scene = diff.DiffScene(worlds=16)
    w = _record(dev, scene, steps)
    try:
        engine = diff.final_positions(dev, w, scene)
        m = scene.mass[..., None]
        msum = scene.mass.sum(1)[:, None, None]
        com = (m * engine).sum(1) / scene.mass.sum(1)[:, None]
        g_com = 2.0 * (com - scene.target)
        g_pos = g_com[:, None, :] * m / msum
        gv, gf, gm = diff.backward(dev, w, scene, g_pos)
        # the loss depends on mass directly too (the COM weights)
        dcom_dm = (engine - com[:, None, :]) / msum
        gm_total = gm + (g_com[:, None, :] * dcom_dm).sum(2)
        rng = np.random.default_rng(11)
        dv = rng.normal(size=scene.v0.shape)
        df = rng.normal(size=scene.friction.shape) * 0.05
        dm = rng.normal(size=scene.mass.shape) * 10.0
        adj = (gv * dv).sum((1, 2)) + (gf * df).sum(1) + (gm_total * dm).sum(1)
        fd1 = _per_world_fd(dev, w, scene, steps, dv, df, dm, 1e-4)
        fd2 = _per_world_fd(dev, w, scene, steps, dv, df, dm, 5e-5)
        smooth = np.abs(fd1 - fd2) <= 1e-4 * np.maximum(np.abs(fd1), 1e-9)
        rel = np.abs(adj - fd1) / np.maximum(np.abs(fd1), 1e-12)
        if steps <= 3:
            assert smooth.all(), smooth
        assert smooth.sum() >= 1
        assert (rel[smooth] <= 1e-3).all(), (rel, smooth)
    finally:
        dev.m3d_world_destroy(w)


@pytest.mark.gpu
def test_torch_function_matches_c_abi(dev):
    torch = pytest.importorskip("torch")
    scene = diff.DiffScene(worlds=4)
    steps = 10
    v0 = torch.tensor(scene.v0, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(scene.friction, dtype=torch.float64, requires_grad=True)
    mass = torch.tensor(scene.mass, dtype=torch.float64, requires_grad=True)
    pos = diff.rollout(dev, scene, v0, mu, mass, steps=steps)
    loss = scene.loss(pos, mass)
    loss.backward()
    w = _record(dev, scene, steps)
    try:
        engine = diff.final_positions(dev, w, scene)
        assert np.allclose(pos.detach().numpy(), engine)
        m = scene.mass[..., None]
        com = (m * engine).sum(1) / scene.mass.sum(1)[:, None]
        g_pos = (2.0 * (com - scene.target))[:, None, :] * m / scene.mass.sum(1)[:, None, None]
        gv, gf, gm = diff.backward(dev, w, scene, g_pos)
        assert np.allclose(v0.grad.numpy(), gv, rtol=1e-5, atol=1e-7)
        assert np.allclose(mu.grad.numpy(), gf, rtol=1e-5, atol=1e-7)
    finally:
        dev.m3d_world_destroy(w)


def test_scene_draws():
    scene = diff.DiffScene(worlds=8)
    assert scene.v0.shape == (8, 64, 3) and abs(scene.v0.std() - 2.0) < 0.3
    assert scene.friction.min() >= 0.2 and scene.friction.max() <= 0.9
    assert scene.sphere.sum() == 32
    idx = scene.dynamic_index()
    assert idx[1, 0] == 66 and idx[0, 63] == 64
