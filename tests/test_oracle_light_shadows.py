"""Oracle pins for the point-light shadow moments: light_shadow.rgen:27-86,
light_shadow.rchit:15-18, light_shadow.rmiss:12-15 (distances of 1024 Fibonacci
rays per light) and light_shadow_sum.comp:16-71 (8x8 octahedral moments).

What pins them:
* distances: a light inside the empty 10 x 4 x 10 m room, where every ray's hit
  distance is the closed form min over axes of (wall - p) / d;
* misses: a range shorter than the nearest wall gives dist = range for every ray,
  so each texel's mean is the range and its second moment the range squared;
* disabled lights write zeros, and the border columns of the moments image are
  never written;
* weights and moments: a float64 re-evaluation from the distances, with the
  direction rounded through RGBA8 and pow(x, 50) in float64;
* orientation: near a wall, the texel facing the wall sees the short distances.
"""
import numpy as np
import pytest

from anerend_amd import scenegen
from oracle import oracle as O

LO = np.array([-5.0, -0.5, -5.0])   # the empty room of room_scene (10 x 4 x 10 m around y = 1.5)
HI = np.array([5.0, 3.5, 5.0])


def _light(pos, rng, on=True):
    return [pos[0], pos[1], pos[2], rng, 1.0, 0.9, 0.8, 1.0 if on else 0.0]


@pytest.fixture(scope="module")
def room():
    return O.OracleScene(scenegen.room_scene(1, num_boxes=0))


def _ray_image(dist, l):
    """Distances of light l in ray order (ray y * 32 + x sits at pixel (32 l + x, y))."""
    return dist[:, 32 * l:32 * l + 32].reshape(-1)


def test_distances_inside_the_empty_room(room):
    pos = [(0.0, 1.5, 0.0), (1.0, 0.5, -2.0), (-4.2, 3.0, 3.3)]
    lights = np.array([_light(p, 40.0) for p in pos], np.float32)
    dist, _, _ = room.light_shadows(lights)
    d = O.fibonacci_dirs(1024).astype(np.float64)
    for l, p in enumerate(pos):
        p = np.array(p)
        with np.errstate(divide="ignore"):
            t = np.where(d > 0, (HI - p) / d, np.where(d < 0, (LO - p) / d, np.inf)).min(1)
        np.testing.assert_allclose(_ray_image(dist, l), t, rtol=2e-6)


def test_all_misses_give_the_range(room):
    # the nearest wall is 2 m away (the ceiling); range 1.5 reaches nothing
    lights = np.array([_light((0.0, 1.5, 0.0), 1.5)], np.float32)
    dist, mom, _ = room.light_shadows(lights)
    assert np.all(dist == np.float32(1.5))
    interior = mom[:, 1:9, :]
    np.testing.assert_allclose(interior[..., 0], 1.5, rtol=1e-6)
    np.testing.assert_allclose(interior[..., 1], 2.25, rtol=1e-6)


def test_disabled_light_and_borders(room):
    lights = np.array([_light((0.0, 1.5, 0.0), 40.0, on=False), _light((0.5, 1.0, 0.5), 40.0)], np.float32)
    dist, mom, _ = room.light_shadows(lights, moments_fill=-7.0)
    assert np.all(_ray_image(dist, 0) == 0.0)
    assert np.all(mom[:, 1:9, :] == 0.0)
    # columns 10 l and 10 l + 9 are borders
    for c in (0, 9, 10, 19):
        assert np.all(mom[:, c, :] == -7.0)
    assert np.all(mom[:, 11:19, :] > 0.0)


def _weights64():
    """float64 restatement of light_shadow_sum.comp:45-61: RGBA8-rounded direction,
    dot with the texel direction, pow 50, kept above 0.001."""
    d = O.fibonacci_dirs(1024).astype(np.float64)
    q = np.rint(np.clip((d + 1.0) * 0.5, 0.0, 1.0) * 255.0) / 255.0 * 2.0 - 1.0
    u = (2.0 * np.arange(8) + 1.0) / 16.0 * 2.0 - 1.0
    ox, oy = (a.ravel() for a in np.meshgrid(u, u))   # t = ty * 8 + tx
    z = 1.0 - np.abs(ox) - np.abs(oy)
    sx = np.where(ox >= 0, 1.0, -1.0)
    sy = np.where(oy >= 0, 1.0, -1.0)
    fx = np.where(z < 0, (1.0 - np.abs(oy)) * sx, ox)
    fy = np.where(z < 0, (1.0 - np.abs(ox)) * sy, oy)
    o = np.stack([fx, z, fy], 1)                          # octDecode swaps y and z
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    w = np.maximum(0.0, o @ q.T) ** 50
    return w


def test_weights_match_float64():
    w32 = O.light_shadow_weights().astype(np.float64)
    w64 = _weights64()
    near = np.abs(w64 - 1e-3) < 1e-6               # threshold cases may round either way
    keep = (w64 > 1e-3) & ~near
    assert np.all(w32[keep] > 0.0) and np.all(w32[(w64 <= 1e-3) & ~near] == 0.0)
    np.testing.assert_allclose(w32[keep], w64[keep], rtol=2e-5)
    # 63-69 rays per texel; the RGBA8 rounding leaves |dir| != 1, so weights can exceed 1
    assert 60 <= (w32 > 0).sum(1).min() and (w32 > 0).sum(1).max() <= 72


def test_moments_match_float64(room):
    scene = scenegen.room_scene(3, num_boxes=24)
    osc = O.OracleScene(scene)
    rng = np.random.default_rng(4)
    pos = rng.uniform([-4.5, 0.0, -4.5], [4.5, 3.0, 4.5], (6, 3))
    lights = np.array([_light(p, r) for p, r in zip(pos, rng.uniform(2.0, 12.0, 6))], np.float32)
    dist, mom, _ = osc.light_shadows(lights)
    w64 = _weights64()
    w64 = np.where(w64 > 1e-3, w64, 0.0)
    for l in range(len(lights)):
        dl = _ray_image(dist, l).astype(np.float64)
        sw = w64.sum(1)
        m1 = (w64 @ dl) / sw
        m2 = (w64 @ (dl * dl)) / sw
        got = mom[:, 10 * l + 1:10 * l + 9, :].reshape(64, 2)  # [ty][tx] -> t = ty * 8 + tx
        np.testing.assert_allclose(got[:, 0], m1, rtol=2e-4)
        np.testing.assert_allclose(got[:, 1], m2, rtol=4e-4)


def test_orientation_near_a_wall(room):
    # 1 m from the +x wall, 9 m from the -x wall: octEncode(+x) = (1, 0) is image
    # column tx = 7, octEncode(-x) = (-1, 0) is tx = 0 (rows ty = 3, 4 straddle v = 0)
    lights = np.array([_light((4.0, 1.5, 0.0), 40.0)], np.float32)
    _, mom, _ = room.light_shadows(lights)
    mean = mom[:, 1:9, 0]
    assert mean[3:5, 7].max() < 1.3
    assert mean[3:5, 0].min() > 8.5
