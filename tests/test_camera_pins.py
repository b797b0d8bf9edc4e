"""Camera pins (CPU): the orthographic projection against its closed form, the
oracle's camera-parameter tangents against central differences, and the
library's camera chain rule (psdr_camera_param_vjp, host) against the oracle.

Camera gradients follow the reference's autodiff: a parameter reaches the image
only through ray origins (src/scene.cc:134-136, src/path_integrator.cc:240-244),
so what must be right is d(origin)/d(parameters) for camera rays
(src/camera.h:235-258, 314-335) and for rays through a point
(src/camera.h:260-276, 337-361), and the matrix chain rule behind it.
"""
import ctypes

import numpy as np
import pytest

import torch_psdr_amd._native as N
from torch_psdr_amd import ops
from tests import helpers as H

# a camera whose up is not perpendicular to the view direction (non-unit right axis)
TILTED = dict(look_from=(0.6, 1.4, 4.2))
CAMS = {
    "persp": lambda r: H.cam_params(r, **TILTED),
    "ortho": lambda r: H.ortho_params(r, view_width=3.7, **TILTED),
}


def test_orthographic_rays_closed_form(oracle):
    # look along -z from (0, 0, 5): right = x, up' = y, a 4 x 4 window on 8 x 8 pixels
    cp = [0.0, 0.0, 5.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 4.0, 8.0, 8.0, 1.0]
    o, d, pix = oracle.camera_rays(cp, 2)
    np.testing.assert_array_equal(d, np.tile(np.float32([0, 0, -1]), (128, 1)))
    np.testing.assert_allclose(o[:, 2], 5.0, atol=1e-6)
    # film point (x + u0, y + u1) -> world (-2 + (x + u0) / 2, 2 - (y + u1) / 2)
    fx = (o[:, 0] + 2.0) * 2.0
    fy = (2.0 - o[:, 1]) * 2.0
    px, py = pix % 8, pix // 8
    assert ((fx >= px - 1e-5) & (fx < px + 1 + 1e-5)).all()
    assert ((fy >= py - 1e-5) & (fy < py + 1 + 1e-5)).all()
    # raster projection of world points, valid inside the window only
    xy, valid = oracle.project_points(cp, np.array([[0.3, -1.1, 0.0], [1.9, 1.9, -3.0], [2.5, 0.0, 0.0],
                                                    [0.0, 0.0, 6.0]], np.float32))
    assert valid.tolist() == [True, True, False, False]
    assert xy[0].tolist() == [int(np.floor((0.3 + 2) * 2)), int(np.floor((2 + 1.1) * 2))]
    assert xy[1].tolist() == [7, 0]


def test_orthographic_window_follows_view_width(oracle):
    # the window's world width is view_width * |right| (look_at() does not
    # normalise right, src/camera.h:36-37) and its height H / W of that
    cp = H.ortho_params(16, view_width=3.0)
    cp[10] = 8.0  # H = 8, W = 16
    o, d, pix = oracle.camera_rays(cp, 1)
    look = -np.array([0.0, 1.0, 5.0]) / np.linalg.norm([0.0, 1.0, 5.0])
    np.testing.assert_allclose(d, np.tile(look, (128, 1)), atol=1e-6)
    np.testing.assert_allclose((o - np.float32([0, 1, 5])) @ look, 0.0, atol=1e-5)
    up = np.array([0.0, 1.0, 0.0])
    right = np.cross(up, -look)
    nup = np.cross(-look, right)
    s = np.linalg.norm(right)
    xs = (o - np.float32([0, 1, 5])) @ (right / s) / s
    ys = (o - np.float32([0, 1, 5])) @ (nup / s) / s
    assert xs.min() > -1.5 and xs.max() < 1.5 and xs.max() - xs.min() > 1.5 * 15 / 16
    assert ys.min() > -0.75 and ys.max() < 0.75 and ys.max() - ys.min() > 0.75 * 7 / 8


@pytest.mark.parametrize("kind", sorted(CAMS))
def test_camera_ray_origin_tangent_matches_central_difference(oracle, kind):
    cp = CAMS[kind](6)
    rng = np.random.default_rng(3)
    for trial in range(4):
        w = rng.normal(size=10)
        w /= np.linalg.norm(w)
        _, _, _, do = oracle.camera_rays(cp, 2, d_params=w)
        h = 1e-2
        plus = list(cp)
        minus = list(cp)
        for i in range(10):
            plus[i] += h * w[i]
            minus[i] -= h * w[i]
        fd = (oracle.camera_rays(plus, 2)[0].astype(np.float64) - oracle.camera_rays(minus, 2)[0]) / (2 * h)
        scale = max(np.abs(fd).max(), 1e-3)
        assert np.abs(do - fd).max() <= 3e-3 * scale, (kind, trial, np.abs(do - fd).max(), scale)
        if kind == "persp":
            np.testing.assert_allclose(do, np.tile(w[:3], (do.shape[0], 1)), atol=1e-6)


@pytest.mark.parametrize("kind", sorted(CAMS))
def test_point_ray_origin_tangent_matches_central_difference(oracle, kind):
    cp = CAMS[kind](6)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1, 1, size=(32, 3)).astype(np.float32)
    for trial in range(4):
        w = rng.normal(size=10)
        w /= np.linalg.norm(w)
        dp = rng.normal(size=(32, 3)).astype(np.float32)
        _, _, do = oracle.point_rays(cp, pts, w, dp)
        h = 1e-2
        plus = [v + h * w[i] if i < 10 else v for i, v in enumerate(cp)]
        minus = [v - h * w[i] if i < 10 else v for i, v in enumerate(cp)]
        zero = np.zeros_like(dp)
        op = oracle.point_rays(plus, pts + h * dp, w, zero)[0].astype(np.float64)
        om = oracle.point_rays(minus, pts - h * dp, w, zero)[0]
        fd = (op - om) / (2 * h)
        scale = max(np.abs(fd).max(), 1e-3)
        assert np.abs(do - fd).max() <= 3e-3 * scale, (kind, trial, np.abs(do - fd).max(), scale)
    if kind == "ortho":
        # the origin is the point's foot on the camera plane: it moves with the
        # point only within the plane
        look = oracle.point_rays(cp, pts, np.zeros(10), zero)[1][0]
        _, _, do = oracle.point_rays(cp, pts, np.zeros(10), dp)
        np.testing.assert_allclose(do @ look, 0.0, atol=1e-5)
        np.testing.assert_allclose(do, dp - np.outer(dp @ look, look), atol=1e-5)


@pytest.mark.parametrize("kind", sorted(CAMS))
def test_camera_matrix_tangents_match_central_difference(oracle, kind):
    cp = CAMS[kind](8)
    rng = np.random.default_rng(5)
    w = rng.normal(size=10)
    w /= np.linalg.norm(w)
    dC, dW, dR = oracle.camera_tangents(cp, w)
    h = 1e-2
    plus = [v + h * w[i] if i < 10 else v for i, v in enumerate(cp)]
    minus = [v - h * w[i] if i < 10 else v for i, v in enumerate(cp)]
    (cP, wP, _, rP), _ = oracle.camera_matrices(plus)
    (cM, wM, _, rM), _ = oracle.camera_matrices(minus)
    # raster2camera rows 0, 1 only: its depth row holds the far-plane scale,
    # whose fp32 rounding swamps a finite difference
    pairs = [(dC, cP, cM), (dW, wP, wM)] + ([(dR[:2], rP[:2], rM[:2])] if kind == "ortho" else [])
    for got, a, b in pairs:
        fd = (a.astype(np.float64) - b) / (2 * h)
        assert np.abs(got - fd).max() <= 3e-3 * max(np.abs(fd).max(), 1e-3)


@pytest.mark.parametrize("kind", sorted(CAMS))
def test_library_camera_matrices_match_oracle(oracle, kind):
    cp = CAMS[kind](8)
    lib = N.load()
    k = ops.make_camera(cp)
    outs = [np.zeros((4, 4), np.float32) for _ in range(4)]
    pos = np.zeros(3, np.float32)
    N.check(lib.psdr_camera_matrices(ctypes.byref(k), *[o.ctypes.data for o in outs], pos.ctypes.data))
    ref, rpos = oracle.camera_matrices(cp)
    for a, b in zip(outs, ref):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(pos, rpos)


@pytest.mark.parametrize("kind", sorted(CAMS))
def test_library_camera_chain_rule_matches_oracle_tangents(oracle, kind):
    # grad_k = sum(acc * d(matrices)/d(param k)) over the 25 matrix entries an
    # origin depends on: camera2world columns 0, 1, 3, world2camera and
    # raster2camera rows 0, 1
    cp = CAMS[kind](8)
    lib = N.load()
    rng = np.random.default_rng(6)
    acc = rng.normal(size=25).astype(np.float32)
    if kind == "persp":
        acc[:6] = 0.0
        acc[9:] = 0.0  # a perspective origin is column 3 of camera2world alone
    grad = np.zeros(10, np.float32)
    k = ops.make_camera(cp)
    N.check(lib.psdr_camera_param_vjp(ctypes.byref(k), acc.ctypes.data, grad.ctypes.data))
    for i in range(10):
        e = np.zeros(10)
        e[i] = 1.0
        dC, dW, dR = [m.astype(np.float64) for m in oracle.camera_tangents(cp, e)]
        ref = float(acc[0:3] @ dC[:3, 0] + acc[3:6] @ dC[:3, 1] + acc[6:9] @ dC[:3, 3] +
                    acc[9:13] @ dW[0] + acc[13:17] @ dW[1] + acc[17:21] @ dR[0] + acc[21:25] @ dR[1])
        assert abs(grad[i] - ref) <= 1e-4 * max(np.abs(acc).max(), 1.0) * max(1.0, abs(ref)), (i, grad[i], ref)
    if kind == "persp":
        np.testing.assert_allclose(grad[:3], acc[6:9], rtol=1e-6)
        np.testing.assert_allclose(grad[3:], 0.0, atol=1e-6)
