"""The native scene driver (used by bench.py) builds the same scenes as the
Python builders, and its churn continues identically on an imported world."""
import numpy as np

from melon3d_amd import scenes
from parity import bodies_array


def _same(lib, wa, wb):
    a, b = bodies_array(lib.export_state(wa)), bodies_array(lib.export_state(wb))
    return len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))


def test_native_rain_matches_python(ref):
    a, b = ref.create_world(), ref.create_world()
    scenes.rain_pile(ref, a, count=300, seed=12345, column=(10.0, 10.0, 20.0))
    scenes.native_rain_pile(ref, b, count=300, seed=12345, column=(10.0, 10.0, 20.0))
    assert _same(ref, a, b)
    ref.m3d_world_destroy(a)
    ref.m3d_world_destroy(b)


def test_native_stream_matches_python_and_clones(ref):
    a, b = ref.create_world(), ref.create_world()
    sa = scenes.Stream(ref, a, count=400, churn=7, seed=777, arena=30.0)
    sb = scenes.NativeStream(ref, b, count=400, churn=7, seed=777, arena=30.0)
    for _ in range(3):
        sa.churn_step()
        sb.churn_step()
        ref.m3d_world_step(a, 1.0 / 60.0, 4)
        ref.m3d_world_step(b, 1.0 / 60.0, 4)
    assert _same(ref, a, b)
    # continue b's stream on an imported copy: identical to continuing b
    c = ref.import_state(ref.export_state(b))
    sc = sb.clone_onto(ref, c)
    for _ in range(2):
        sb.churn_step()
        sc.churn_step()
        ref.m3d_world_step(b, 1.0 / 60.0, 4)
        ref.m3d_world_step(c, 1.0 / 60.0, 4)
    assert _same(ref, b, c)
    for w in (a, b, c):
        ref.m3d_world_destroy(w)
