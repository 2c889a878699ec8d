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
    assert dev.m3dx_step_failures(w) == 0, dev.last_error()
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


@pytest.mark.gpu
def test_parallel_forward_is_the_engine(dev):
    """The parallel adjoint's forward replay (level-parallel, one workgroup per
    world) reproduces the engine bit for bit: each step from its recorded start
    equals the next recorded start, and 40 chained steps equal the engine's
    final state."""
    import ctypes as C
    scene = diff.DiffScene(worlds=16)
    steps = 40
    w = _record(dev, scene, steps)
    try:
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        n = scene.body_count()
        idx = scene.dynamic_index().ravel()
        for t in range(steps - 1):
            out = np.zeros((n, 13), dtype=np.float32)
            nxt = np.zeros((n, 13), dtype=np.float32)
            assert dev.m3dx_diff_forward(w, t, 1, ptr(out)) == 0, dev.last_error()
            assert dev.m3dx_diff_state(w, t + 1, ptr(nxt)) >= 0
            assert np.array_equal(out[idx], nxt[idx]), t
        out = np.zeros((n, 13), dtype=np.float32)
        assert dev.m3dx_diff_forward(w, 0, steps, ptr(out)) == 0
        engine = diff.final_positions(dev, w, scene)
        assert np.array_equal(out[scene.dynamic_index(), 0:3].astype(np.float64), engine)
    finally:
        dev.m3d_world_destroy(w)


def _grads(dev, w, scene, tape=False):
    import ctypes as C
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    n = scene.body_count()
    rng = np.random.default_rng(5)
    seed = rng.normal(size=(n, 13)).astype(np.float32)
    adj = np.zeros((n, 13), dtype=np.float32)
    mg = np.zeros(n, dtype=np.float32)
    fg = np.zeros(n, dtype=np.float32)
    fn = dev.m3dx_diff_backward_tape if tape else dev.m3dx_diff_backward
    assert fn(w, ptr(seed), ptr(adj), ptr(mg), ptr(fg)) == 0, dev.last_error()
    return adj, mg, fg


@pytest.mark.gpu
def test_parallel_adjoint_matches_tape(dev):
    """The parallel adjoint (forward tangents per op component, reverse level
    order) and the reverse-mode tape (one lane per world) are independent
    evaluations of the same vector-Jacobian product: with a random seed on
    every state component over 6 steps they agree to float rounding (1e-4 of
    each world's largest component). The checkpoint interval changes nothing
    (recomputed states are bit-exact): intervals 1, 4 and 10 give identical
    gradients."""
    scene = diff.DiffScene(worlds=16)
    steps = 6
    w = _record(dev, scene, steps)
    try:
        at, mt, ft = _grads(dev, w, scene, tape=True)
        res = []
        for k in (1, 4, 10):
            assert dev.m3dx_diff_set_checkpoint(w, k) == 0
            res.append(_grads(dev, w, scene))
        for r in res[1:]:
            for a, b in zip(r, res[0]):
                assert np.array_equal(a, b)
        ap, mp, fp = res[0]
        idx = scene.dynamic_index()
        # static bodies: their poses are constants of the parallel adjoint (its
        # state adjoint for them is the seed), the tape also differentiates
        # them; dynamic bodies' state and mass, and every body's friction
        for name, a, b in (("state", ap[idx], at[idx]), ("mass", mp[idx], mt[idx]),
                           ("friction", fp.reshape(scene.worlds, -1), ft.reshape(scene.worlds, -1))):
            a = a.reshape(scene.worlds, -1).astype(np.float64)
            b = b.reshape(scene.worlds, -1).astype(np.float64)
            scale = np.abs(b).max(axis=1, keepdims=True) + 1e-12
            err = (np.abs(a - b) / scale).max()
            assert err <= 1e-4, (name, err)
        ground = np.arange(scene.worlds) * (scene.body_count() // scene.worlds)
        seed = np.random.default_rng(5).normal(size=(scene.body_count(), 13)).astype(np.float32)
        assert np.array_equal(ap[ground], seed[ground])
    finally:
        dev.m3d_world_destroy(w)


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [3, 10])
def test_adjoint_dot_fp64_cpu_tangents(dev, steps):
    """The dot-product test against the FP64 CPU replay gradient: for a random
    seed y on the final state and a random direction u on the dynamic bodies'
    initial velocities (linear and angular), mass and friction, <y, J u> from
    FP64 forward tangents on the host CPU along the engine's own float path
    (from the recorded inputs) equals <J^T y, u> from the device adjoint,
    per world, to 1e-3 relative."""
    import ctypes as C
    scene = diff.DiffScene(worlds=16)
    w = _record(dev, scene, steps)
    try:
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        n = scene.body_count()
        idx = scene.dynamic_index()
        rng = np.random.default_rng(21)
        seed = np.zeros((n, 13), np.float32)
        seed[idx] = rng.normal(size=(scene.worlds, diff.BODIES, 13))
        u = np.zeros((n, 13))
        u[idx, 7:13] = rng.normal(size=(scene.worlds, diff.BODIES, 6))
        dmass = np.zeros(n)
        dmass[idx] = rng.normal(size=idx.shape) * 10.0
        dfric = rng.normal(size=n) * 0.05
        adj = np.zeros((n, 13), np.float32)
        mg = np.zeros(n, np.float32)
        fg = np.zeros(n, np.float32)
        assert dev.m3dx_diff_backward(w, ptr(seed), ptr(adj), ptr(mg), ptr(fg)) == 0, dev.last_error()
        st0 = np.zeros((n, 13), np.float32)
        assert dev.m3dx_diff_state(w, 0, ptr(st0)) >= 0
        st0 = st0.astype(np.float64)
        out = np.zeros((n, 13))
        dout = np.zeros((n, 13))
        assert dev.m3dx_diff_jvp_cpu(w, 0, steps, 0, ptr(st0), ptr(u), None, ptr(dmass), None, ptr(dfric),
                                     ptr(out), ptr(dout)) == 0, dev.last_error()
        per = n // scene.worlds
        jvp = (seed.astype(np.float64) * dout).reshape(scene.worlds, per * 13).sum(1)
        vjp = ((adj.astype(np.float64) * u).reshape(scene.worlds, per * 13).sum(1) +
               (mg.astype(np.float64) * dmass).reshape(scene.worlds, per).sum(1) +
               (fg.astype(np.float64) * dfric).reshape(scene.worlds, per).sum(1))
        rel = np.abs(vjp - jvp) / np.maximum(np.abs(jvp), 1e-9)
        assert (rel <= 1e-3).all(), (rel, jvp, vjp)
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
        # mass enters twice: through the trajectory (the adjoint's gm) and
        # through the centre-of-mass weights, d com / d m_k = (p_k - com) / M
        M = scene.mass.sum(1)[:, None]
        g_com = 2.0 * (com - scene.target)
        g_weights = np.einsum("wc,wkc->wk", g_com, engine - com[:, None, :]) / M
        assert np.allclose(mass.grad.numpy(), gm + g_weights, rtol=1e-5, atol=1e-9)
    finally:
        dev.m3d_world_destroy(w)


@pytest.mark.gpu
def test_torch_function_second_backward_and_release(dev):
    """The recorded world lives as long as the graph: a second backward
    (retain_graph) gives the same gradients, and dropping the graph frees
    the world."""
    torch = pytest.importorskip("torch")
    import gc

    scene = diff.DiffScene(worlds=2)
    v0 = torch.tensor(scene.v0, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(scene.friction, dtype=torch.float64, requires_grad=True)
    mass = torch.tensor(scene.mass, dtype=torch.float64, requires_grad=True)
    pos = diff.rollout(dev, scene, v0, mu, mass, steps=3)
    release = pos.grad_fn.owner._finalizer  # holds no reference to the owner
    loss = scene.loss(pos, mass)
    loss.backward(retain_graph=True)
    g1 = v0.grad.clone()
    v0.grad = None
    loss.backward()
    assert torch.equal(g1, v0.grad)
    assert release.alive
    del pos, loss
    gc.collect()
    assert not release.alive


def _raft(dev, settle):
    """The 15 x 15 raft of touching boxes on the ground with five boxes
    dropping onto it (test/test_main.cpp:960-985, sleep off): one island of
    about 650 contacts, which the engine solves with the colored schedule
    (greedy colors in 8-lane packets, then the overflow pass). Returns the
    world, its dynamic bodies' indices in the recorded state arrays and their
    ids."""
    from melon3d_amd import scenes

    wd = dev.m3d_world_def_default()
    wd.enableSleep = False
    w = dev.create_world(wd)
    scenes.add_ground(dev, w)
    ids = [scenes.add_box(dev, w, (-7.0 + i, 0.5, -7.0 + k), (0.501, 0.5, 0.501)) for i in range(15) for k in range(15)]
    ids += [scenes.add_box(dev, w, (2.0 * i - 4.0, 3.0 + 0.7 * i, 0.5 * i - 1.0), (0.45, 0.45, 0.45)) for i in range(5)]
    for _ in range(settle):
        dev.m3d_world_step(w, DT, 4)
    return w, np.arange(1, 231), ids


def _record_raft(dev, settle, steps):
    w, dyn, ids = _raft(dev, settle)
    assert dev.m3dx_diff_record(w, 1) == 0, dev.last_error()
    colored = []
    for _ in range(steps):
        dev.m3d_world_step(w, DT, 4)
        colored.append(dev.stats(w).coloredIslands)
    assert dev.m3dx_diff_record(w, 0) == 0, dev.last_error()
    assert dev.m3dx_diff_steps(w) == steps
    assert dev.m3dx_step_failures(w) == 0, dev.last_error()
    assert min(colored) >= 1, colored
    return w, dyn, ids


@pytest.mark.gpu
def test_colored_raft_replay_is_the_engine(dev):
    """Colored steps replayed in float reproduce the engine bit for bit: the
    single-lane replay and the parallel adjoint's forward kernels, each step
    from its recorded start against the next recorded start, and 8 chained
    steps against the engine's final poses. The steps record the boxes
    landing on the raft (steps 41-48), so packets of mixed point counts and
    the overflow pass both occur."""
    import ctypes as C
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    steps = 8
    w, dyn, ids = _record_raft(dev, 40, steps)
    try:
        n = 231
        for t in range(steps - 1):
            nxt = np.zeros((n, 13), np.float32)
            assert dev.m3dx_diff_state(w, t + 1, ptr(nxt)) >= 300
            out = np.zeros((n, 13), np.float32)
            assert dev.m3dx_diff_replay(w, t, 1, 0, None, None, None, ptr(out)) == 0, dev.last_error()
            assert np.array_equal(out[dyn], nxt[dyn]), t
            par = np.zeros((n, 13), np.float32)
            assert dev.m3dx_diff_forward(w, t, 1, ptr(par)) == 0, dev.last_error()
            assert np.array_equal(par[dyn], nxt[dyn]), t
        out = np.zeros((n, 13), np.float32)
        assert dev.m3dx_diff_forward(w, 0, steps, ptr(out)) == 0
        engine = np.array([[getattr(dev.m3d_body_position(w, b), a) for a in "xyz"] for b in ids], np.float32)
        assert np.array_equal(out[dyn, 0:3], engine)
    finally:
        dev.m3d_world_destroy(w)


@pytest.mark.gpu
def test_colored_raft_adjoint(dev):
    """The adjoint of colored steps (4 steps of boxes landing on the raft).

    - The parallel adjoint equals the reverse-mode tape to float rounding and
      is the same for checkpoint intervals 1 and 10.
    - Dot-product test: <J^T y, u> from the adjoint equals <y, J u> from FP64
      forward tangents on the host CPU along the same float path, for random
      directions in the initial velocities, mass and friction.
    - The derivative code against FP64 central differences: FP64 tangents
      along the double replay equal central differences of the double
      replay of a linear loss on the final poses, to 1e-2 of the
      directions' RMS derivative. The stencil is h = 1e-10 (velocities in
      m/s): the velocity pass applies friction only where the tangential
      speed exceeds 1e-9 (src/xpbd.cpp:326), and on resting boxes the
      previous friction impulse leaves it just below that, so any larger
      stencil crosses into the friction branch and measures that branch's
      derivative instead (central differences at 1e-6 .. 1e-9 agree with
      each other and not with either path's tangents). At this stencil the
      double replay's rounding (about 1e-15 m on coordinates of up to 7 m)
      leaves central differences good to a few 1e-3.
    The float and double replays themselves drift apart by about 1e-4 m over
    these steps, which moves their derivatives by about 1 %: the adjoint is
    the derivative of the engine's own float path. The thrown lattice below
    compares the adjoint with central differences directly."""
    import ctypes as C
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    steps = 4
    w, dyn, _ = _record_raft(dev, 40, steps)
    try:
        n = 231
        rng = np.random.default_rng(17)
        seed = np.zeros((n, 13))
        seed[dyn] = rng.normal(size=(len(dyn), 13))
        seed32 = seed.astype(np.float32)
        # tape against the parallel adjoint, and checkpoint intervals
        at, mt, ft = (np.zeros((n, 13), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))
        assert dev.m3dx_diff_backward_tape(w, ptr(seed32), ptr(at), ptr(mt), ptr(ft)) == 0, dev.last_error()
        res = []
        for k in (1, 10):
            assert dev.m3dx_diff_set_checkpoint(w, k) == 0
            r = (np.zeros((n, 13), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32))
            assert dev.m3dx_diff_backward(w, ptr(seed32), ptr(r[0]), ptr(r[1]), ptr(r[2])) == 0, dev.last_error()
            res.append(r)
        for a, b in zip(res[0], res[1]):
            assert np.array_equal(a, b)
        ap, mp, fp = res[0]
        for name, a, b in (("state", ap[dyn], at[dyn]), ("mass", mp[dyn], mt[dyn]), ("friction", fp, ft)):
            a = a.astype(np.float64)
            b = b.astype(np.float64)
            err = np.abs(a - b).max() / (np.abs(b).max() + 1e-12)
            assert err <= 1e-4, (name, err)
        # FP64 adjoint, then the dot-product test against FP64 CPU tangents
        adj = np.zeros((n, 13))
        mg = np.zeros(n)
        fg = np.zeros(n)
        assert dev.m3dx_diff_backward_f64(w, ptr(seed), ptr(adj), ptr(mg), ptr(fg)) == 0, dev.last_error()
        st0 = np.zeros((n, 13), np.float32)
        assert dev.m3dx_diff_state(w, 0, ptr(st0)) >= 300
        st0 = st0.astype(np.float64)
        for _ in range(3):
            u = np.zeros((n, 13))
            u[dyn, 7:13] = rng.normal(size=(len(dyn), 6))
            dmass = np.zeros(n)
            dmass[dyn] = rng.normal(size=len(dyn)) * 10.0
            dfric = rng.normal(size=n) * 0.05
            out = np.zeros((n, 13))
            dout = np.zeros((n, 13))
            assert dev.m3dx_diff_jvp_cpu(w, 0, steps, 0, ptr(st0), ptr(u), None, ptr(dmass), None, ptr(dfric),
                                         ptr(out), ptr(dout)) == 0, dev.last_error()
            jvp = (seed * dout).sum()
            vjp = (adj * u).sum() + (mg * dmass).sum() + (fg * dfric).sum()
            assert abs(vjp - jvp) <= 1e-9 * abs(jvp), (vjp, jvp)
        # FP64 tangents of the double replay against its central differences
        y = np.zeros((n, 13))
        y[dyn, 0:7] = rng.normal(size=(len(dyn), 7))
        zero = np.zeros(n)

        def loss(x0):
            o = np.zeros((n, 13))
            assert dev.m3dx_diff_replay(w, 0, steps, 1, ptr(np.ascontiguousarray(x0)), None, None, ptr(o)) == 0
            return (y * o).sum()

        tangents, fds = [], []
        for _ in range(6):
            u = np.zeros((n, 13))
            u[dyn, 7:13] = rng.normal(size=(len(dyn), 6))
            out = np.zeros((n, 13))
            dout = np.zeros((n, 13))
            assert dev.m3dx_diff_jvp_cpu(w, 0, steps, 1, ptr(st0), ptr(u), None, ptr(zero), None, ptr(zero),
                                         ptr(out), ptr(dout)) == 0, dev.last_error()
            tangents.append((y * dout).sum())
            fds.append((loss(st0 + 1e-10 * u) - loss(st0 - 1e-10 * u)) / 2e-10)
        tangents, fds = np.array(tangents), np.array(fds)
        scale = np.sqrt((fds ** 2).mean())
        assert np.abs(tangents - fds).max() <= 1e-2 * scale, (tangents, fds)
    finally:
        dev.m3d_world_destroy(w)


def _lattice(dev, rng):
    """A 6 x 6 x 6 lattice of unit spheres and cubes (alternating, 2 cm
    gaps) thrown with random linear and angular velocities, sleep off: one
    island of more than 500 contacts, solved with the colored schedule, whose
    contacts are in flight (sliding) rather than resting."""
    import ctypes as C

    from melon3d_amd.api import M3D_BODY_DYNAMIC, M3D_SHAPE_BOX, M3D_SHAPE_SPHERE, m3d_vec3

    wd = dev.m3d_world_def_default()
    wd.enableSleep = False
    w = dev.create_world(wd)
    for i in range(6):
        for j in range(6):
            for k in range(6):
                bd = dev.m3d_body_def_default()
                bd.type = M3D_BODY_DYNAMIC
                bd.position = m3d_vec3(1.02 * i, 1.02 * j, 1.02 * k)
                bd.linearVelocity = m3d_vec3(*rng.normal(size=3))
                bd.angularVelocity = m3d_vec3(*rng.normal(size=3))
                sd = dev.m3d_shape_def_default()
                if (i + j + k) % 2 == 0:
                    sd.type = M3D_SHAPE_SPHERE
                    sd.radius = 0.5
                else:
                    sd.type = M3D_SHAPE_BOX
                    sd.halfExtents = m3d_vec3(0.5, 0.5, 0.5)
                sd.friction = 0.6
                dev.m3d_body_create(w, C.byref(bd), C.byref(sd))
    return w


@pytest.mark.gpu
def test_colored_lattice_adjoint_matches_fp64_central_differences(dev):
    """Directional derivatives of a linear loss on the final poses after 2
    colored steps of the thrown lattice, along random directions in the
    initial linear and angular velocities, against FP64 central differences
    of the double replay (h = 1e-5), wherever central differences at h and
    h / 2 agree to 1e-4 (no branch flips inside the stencil; at least half
    the directions qualify):
    - the adjoint (FP64 along the engine's float path): rel err <= 1e-3;
    - FP64 tangents along the double replay itself: rel err <= 1e-6.
    Over a third step a branch decision of the float and double replays
    differs (their final poses part by 0.16 m) and the two paths'
    derivatives by about 10 %, so the horizon stays at 2 steps."""
    import ctypes as C
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rng = np.random.default_rng(29)
    w = _lattice(dev, rng)
    try:
        steps = 2
        assert dev.m3dx_diff_record(w, 1) == 0, dev.last_error()
        colored = []
        for _ in range(steps):
            dev.m3d_world_step(w, DT, 4)
            colored.append(dev.stats(w).coloredIslands)
        assert dev.m3dx_diff_record(w, 0) == 0, dev.last_error()
        assert min(colored) >= 1, colored
        n = 216
        st0 = np.zeros((n, 13), np.float32)
        assert dev.m3dx_diff_state(w, 0, ptr(st0)) >= 300
        st0 = st0.astype(np.float64)
        y = rng.normal(size=(n, 13))
        y[:, 7:13] = 0.0
        g = np.zeros((n, 13))
        assert dev.m3dx_diff_backward_f64(w, ptr(y), ptr(g), None, None) == 0, dev.last_error()
        zero = np.zeros(n)

        def loss(x0):
            o = np.zeros((n, 13))
            assert dev.m3dx_diff_replay(w, 0, steps, 1, ptr(np.ascontiguousarray(x0)), None, None, ptr(o)) == 0
            return (y * o).sum()

        smooth = 0
        for _ in range(8):
            u = np.zeros((n, 13))
            u[:, 7:13] = rng.normal(size=(n, 6))
            fd1 = (loss(st0 + 1e-5 * u) - loss(st0 - 1e-5 * u)) / 2e-5
            fd2 = (loss(st0 + 5e-6 * u) - loss(st0 - 5e-6 * u)) / 1e-5
            if abs(fd1 - fd2) > 1e-4 * abs(fd1):
                continue
            smooth += 1
            adj = (g * u).sum()
            assert abs(adj - fd1) <= 1e-3 * abs(fd1), (adj, fd1, fd2)
            out = np.zeros((n, 13))
            dout = np.zeros((n, 13))
            assert dev.m3dx_diff_jvp_cpu(w, 0, steps, 1, ptr(st0), ptr(u), None, ptr(zero), None, ptr(zero),
                                         ptr(out), ptr(dout)) == 0, dev.last_error()
            tan = (y * dout).sum()
            assert abs(tan - fd1) <= 1e-6 * abs(fd1), (tan, fd1, fd2)
        assert smooth >= 4, smooth
    finally:
        dev.m3d_world_destroy(w)


@pytest.mark.gpu
def test_recording_refuses_joints(dev):
    """Joints are not replayed by the adjoint: a world holding one is
    refused instead of returning the gradient of a solve without them. The
    steps still run as usual."""
    import ctypes as C

    from melon3d_amd import scenes
    from melon3d_amd.api import m3d_ball_joint_def, m3d_vec3

    wd = dev.m3d_world_def_default()
    wd.enableSleep = False
    w = dev.create_world(wd)
    try:
        scenes.add_ground(dev, w)
        a = scenes.add_box(dev, w, (0.0, 0.5, 0.0), (0.5, 0.5, 0.5))
        b = scenes.add_box(dev, w, (0.0, 3.0, 0.0), (0.3, 0.3, 0.3))
        bj = m3d_ball_joint_def()
        bj.bodyA, bj.bodyB = a, b
        bj.localAnchorA = m3d_vec3(0.0, 1.5, 0.0)
        bj.localAnchorB = m3d_vec3(0.0, -1.0, 0.0)
        dev.m3d_joint_create_ball(w, C.byref(bj))
        assert dev.m3dx_diff_record(w, 1) == 0
        for _ in range(5):
            dev.m3d_world_step(w, 1.0 / 60.0, 4)
        assert dev.m3dx_step_failures(w) == 0
        assert dev.m3dx_diff_steps(w) == 0
        assert dev.m3dx_diff_record(w, 0) == -1
        assert "joints" in dev.last_error()
        seed = np.zeros(3 * 13)
        out = np.zeros(3 * 13)
        ptr = lambda v: v.ctypes.data_as(C.c_void_p)
        assert dev.m3dx_diff_backward_f64(w, ptr(seed), ptr(out), None, None) == -1
        assert "joints" in dev.last_error()
        # the engine kept stepping: the box rests on the ground
        assert abs(dev.m3d_body_position(w, a).y - 0.5) < 0.01
    finally:
        dev.m3d_world_destroy(w)


def test_scene_draws():
    scene = diff.DiffScene(worlds=8)
    assert scene.v0.shape == (8, 64, 3) and abs(scene.v0.std() - 2.0) < 0.3
    assert scene.friction.min() >= 0.2 and scene.friction.max() <= 0.9
    assert scene.sphere.sum() == 32
    idx = scene.dynamic_index()
    assert idx[1, 0] == 66 and idx[0, 63] == 64
