"""The viewer caller's host logic: the orbit camera's closed form
(lib.rs:160-169) and the depth range rule of State::frame (lib.rs:777)."""
import math

import numpy as np
import pytest

from nanotracer_amd.viewer import OrbitCamera, TrainingChannel, TrainSnapshot


def test_orbit_position_closed_form():
    o = OrbitCamera(target=(1.0, 2.0, 3.0), distance=4.0)
    assert o.position() == (5.0, 2.0, 3.0)
    up = OrbitCamera(distance=2.0, elevation=math.pi / 2).position()
    assert up[1] == pytest.approx(2.0, abs=1e-6) and abs(up[0]) < 1e-6 and abs(up[2]) < 1e-6
    side = OrbitCamera(distance=3.0, azimuth=math.pi / 2).position()
    assert side[2] == pytest.approx(3.0, abs=1e-6) and abs(side[0]) < 1e-6
    rng = np.random.default_rng(3)
    for _ in range(20):
        az, el, d = rng.uniform(-3, 3), rng.uniform(-1.5, 1.5), rng.uniform(0.5, 20)
        p = np.array(OrbitCamera(distance=d, azimuth=az, elevation=el).position())
        assert np.linalg.norm(p) == pytest.approx(d, rel=1e-5)


def test_depth_max_rule():
    assert OrbitCamera(distance=5.0).depth_max() == 50.0
    assert OrbitCamera(distance=20.0).depth_max() == 80.0
    assert OrbitCamera(distance=12.5).depth_max() == 50.0


def test_channel_reports_only_newer_snapshots():
    ch = TrainingChannel()
    assert ch.poll() is None
    ch.publish(TrainSnapshot(None, 0, 1.0, 10, 1))
    assert ch.poll().version == 1 and ch.poll() is None
    ch.publish(TrainSnapshot(None, 3, 0.5, 12, 2))
    assert ch.poll().iteration == 3 and ch.history == [(0, 1.0, 10), (3, 0.5, 12)]
