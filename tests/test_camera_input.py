"""Camera::processInput (lib/src/DataStructure/Camera.cpp:13-90) through cupbr_camera_process_input:
host code, no GPU. The expected values are the closed forms of the reference's update rules."""
import math

import numpy as np
import pytest

import cupbr_amd as cb

f32 = np.float32


def _front(yaw, pitch):
    return np.array([math.cos(yaw) * math.cos(pitch), math.sin(pitch), math.sin(yaw) * math.cos(pitch)])


def _fresh(w=200, h=100):
    return cb.Camera(w, h), cb.CameraController(w, h)


def test_initial_frame_rebuilds_the_axes_from_yaw_and_pitch():
    """Camera(w, h) holds yaw 3.14159/2, pitch 0 (Camera.h:80-81); the first call records the cursor
    (no turn, not moved) and rebuilds z = front(yaw, pitch), x = aspect * normalize((z.z, 0, -z.x)),
    y = z x x / aspect."""
    cam, ctl = _fresh()
    cam.process_input(ctl, 0, (310.0, 95.0), 0.016)
    assert not cam.moved()
    yaw = float(f32(3.14159) / f32(2.0))
    z = _front(yaw, 0.0)
    np.testing.assert_allclose(cam.z_axis, z, atol=1e-6)
    x = 2.0 * np.array([z[2], 0, -z[0]]) / math.hypot(z[2], z[0])
    np.testing.assert_allclose(cam.x_axis, x, atol=1e-6)
    np.testing.assert_allclose(cam.y_axis, np.cross(z, x) / 2.0, atol=1e-6)
    np.testing.assert_allclose(cam.position, 0.0)


@pytest.mark.parametrize("keys,speed", [(cb.KEY_W, 4.0), (cb.KEY_W | cb.KEY_LEFT_SHIFT, 16.0),
                                        (cb.KEY_W | cb.KEY_LEFT_CONTROL, 1.0),
                                        (cb.KEY_W | cb.KEY_LEFT_SHIFT | cb.KEY_LEFT_CONTROL, 4.0)])
def test_keys_move_along_the_axes(keys, speed):
    """W/S move along z, D/A along x, at 4 units/s, x4 with shift, /4 with control; moved is set."""
    cam, ctl = _fresh()
    cam.process_input(ctl, 0, (0.0, 0.0), 0.0)  # settles the axes
    z0, x0 = cam.z_axis.astype(np.float64), cam.x_axis.astype(np.float64)
    dt = 0.25
    cam.process_input(ctl, keys, (0.0, 0.0), dt)
    assert cam.moved()
    np.testing.assert_allclose(cam.position, speed * dt * z0, rtol=1e-6, atol=1e-7)
    cam.process_input(ctl, cb.KEY_S, (0.0, 0.0), dt)
    np.testing.assert_allclose(cam.position, (speed - 4.0) * dt * z0, rtol=1e-6, atol=1e-6)
    p = cam.position.astype(np.float64)
    cam.process_input(ctl, cb.KEY_D | cb.KEY_A, (0.0, 0.0), dt)  # opposite keys cancel
    np.testing.assert_allclose(cam.position, p, atol=1e-6)
    cam.process_input(ctl, cb.KEY_D, (0.0, 0.0), dt)
    np.testing.assert_allclose(cam.position, p + 4.0 * dt * x0, rtol=1e-6, atol=1e-6)
    cam.process_input(ctl, 0, (0.0, 0.0), dt)
    assert not cam.moved()


def test_cursor_turns_yaw_and_pitch_with_the_clamp():
    """Offsets (last - current) x 0.002 add to yaw and pitch; pitch stays within +-3.14159/2."""
    cam, ctl = _fresh()
    cam.process_input(ctl, 0, (100.0, 100.0), 0.0)
    y0 = ctl.yaw
    cam.process_input(ctl, 0, (90.0, 120.0), 0.0)  # 10 px left, 20 px down
    assert cam.moved()
    assert ctl.yaw == pytest.approx(y0 + 10 * 0.002, abs=1e-6)
    assert ctl.pitch == pytest.approx(-20 * 0.002, abs=1e-6)
    np.testing.assert_allclose(cam.z_axis, _front(ctl.yaw, ctl.pitch), atol=1e-6)
    cam.process_input(ctl, 0, (90.0, -5000.0), 0.0)  # far up: clamped
    assert ctl.pitch == pytest.approx(3.14159 / 2, abs=1e-6)
    cam.process_input(ctl, 0, (90.0, 5000.0), 0.0)
    assert ctl.pitch == pytest.approx(-3.14159 / 2, abs=1e-6)
    # a sub-threshold cursor offset (|offset| < 1e-5 px before scaling) does not count as moved
    cam.process_input(ctl, 0, (90.0, 5000.0), 0.0)
    assert not cam.moved()


def test_moved_camera_restarts_the_renderers_running_mean_flag():
    """The renderer restarts accumulation on Camera::moved() (PBRenderer.cu:174-177); the flag is
    what process_input sets, so a caller passes the camera straight to render()."""
    cam, ctl = _fresh(64, 64)
    cam.process_input(ctl, 0, (5.0, 5.0), 0.01)
    cam.process_input(ctl, cb.KEY_A, (5.0, 5.0), 0.01)
    assert cam.moved()
    cam.process_input(ctl, 0, (5.0, 5.0), 0.01)
    assert not cam.moved()
