"""scene.bin frame dumps (examples/dump_scene.cpp:212-237).

The golden file is the one the reference's own example writes
(oracle/_ref/dump_scene: the example compiled with the reference sources):
its header, body table and 500 frames pin the reader, the writer and the
restated showcase scene (the same scene built through the C API on the
reference build must reproduce the file byte for byte), and the device's
dump of the same scene is diffed against it frame by frame.
"""
import os
import subprocess

import numpy as np
import pytest

from melon3d_amd import scene_dump as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "oracle", "_ref", "dump_scene")


@pytest.fixture(scope="module")
def golden(tmp_path_factory):
    if not os.path.exists(EXAMPLE):
        pytest.skip("oracle/_ref not built (reference tree absent when build() ran)")
    d = tmp_path_factory.mktemp("golden")
    r = subprocess.run([EXAMPLE], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "500 frames, 78 bodies" in r.stdout
    path = str(d / "scene.bin")
    return path, S.read_scene(path)


def test_reader_parses_the_examples_file(golden):
    path, (infos, poses) = golden
    assert os.path.getsize(path) == 8 + 78 * 32 + 500 * 78 * 7 * 4
    assert infos.shape == (78,) and poses.shape == (500, 78, 7)
    # ground first (static box 60 x 1 x 60), then the 21-box pyramid
    assert infos[0]["shapeType"] == 2 and infos[0]["isStatic"] == 1
    assert tuple(infos[0]["halfExtents"]) == (60.0, 1.0, 60.0)
    assert (infos[1:22]["colorId"] == 1).all()
    q = poses[..., 3:]
    # unit rotations up to the engine's first-order renormalization
    assert np.allclose(np.linalg.norm(q, axis=-1), 1.0, atol=1e-3)


def test_showcase_on_the_reference_reproduces_the_examples_file(golden, ref, tmp_path):
    path, (infos, poses) = golden
    mine = str(tmp_path / "mine.bin")
    S.dump(ref, mine, frames=500, substeps=8)
    assert open(mine, "rb").read() == open(path, "rb").read()


def test_writer_round_trip_and_checks(tmp_path):
    infos = np.zeros(3, dtype=S.BODY_DTYPE)
    infos["shapeType"] = [0, 1, 2]
    infos["halfExtents"][2] = (1.0, 2.0, 3.0)
    rng = np.random.default_rng(0)
    frames = rng.normal(size=(4, 3, 7)).astype(np.float32)
    p = str(tmp_path / "x.bin")
    w = S.SceneWriter(p, infos, 4)
    for f in frames:
        w.write_frame(f)
    w.close()
    i2, f2 = S.read_scene(p)
    assert np.array_equal(i2, infos) and np.array_equal(f2, frames)
    r = S.compare(frames, f2)
    assert r["identical"] and r["first_beyond"] == -1
    g = frames.copy()
    g[2, 1, 0] += 1e-3
    r = S.compare(frames, g)
    assert r["first_beyond"] == 2 and abs(r["max_pos"] - 1e-3) < 1e-6
    w = S.SceneWriter(str(tmp_path / "short.bin"), infos, 2)
    w.write_frame(frames[0])
    with pytest.raises(ValueError):
        w.close()
    with open(str(tmp_path / "short.bin"), "ab") as f:
        f.write(b"\0" * 4)
    with pytest.raises(ValueError):
        S.read_scene(str(tmp_path / "short.bin"))


@pytest.mark.gpu
def test_device_dump_against_the_examples_file(golden, dev, tmp_path):
    """The device runs the showcase (joints, a pyramid demolition, mixed
    rain; 8 substeps) for 500 frames: every frame within 1e-5 m / 1e-5 rad of
    the reference example's own dump."""
    path, (infos, poses) = golden
    mine = str(tmp_path / "device.bin")
    dp = S.dump(dev, mine, frames=500, substeps=8)
    i2, p2 = S.read_scene(mine)
    assert np.array_equal(i2, infos)
    r = S.compare(poses, p2)
    print(f"device vs reference example: max |dx| {r['max_pos']:.3g} m, max rotation {r['max_rot']:.3g} rad, "
          f"bit-identical {r['identical']}")
    assert r["first_beyond"] == -1, (r["first_beyond"], r["max_pos"], r["max_rot"])
    assert np.array_equal(dp, p2)
