"""Pins of the oracle's post-process (post_process.hlsl:25-72, linear_to_srgb
common.hlsl:251-259) by values the mathematics fixes, not by its own formula.

- sRGB: the published encodings linear 0.5 -> 188, 0.18 -> 118, 0.214 at
  the 127/128 boundary, 0.0031308 (the segment joint) -> 0.04045 -> 10,
  1 -> 255.
- Reinhard: energy 1 -> 0.5 -> 188 in view mode NONE; energy 3 -> 0.75.
- Per mode: which of accumulation, Reinhard and sRGB apply; each expected
  byte below changes if one rule is flipped (accumulating a debug view,
  tonemapping or sRGB-encoding the wrong mode).
- Accumulation: N frames with sample_count 1..N give the running mean of
  the frames (the closed form of the recursive blend).
"""
import numpy as np
import pytest

import oracle_api as O


def px(*vals):
    return np.array([[*vals, 1.0]], np.float32)


def post(e, mode, accumulate=0, acc=None, n=1):
    acc = np.zeros_like(e) if acc is None else acc
    return O.post_process(e, acc, n, accumulate, mode)


def test_srgb_published_values():
    # view mode 6 (base colour): sRGB without Reinhard
    _, b, f = post(px(0.5, 0.0031308, 1.0), 6)
    assert tuple(b[0, :3]) == (188, 10, 255)
    assert abs(f[0, 0] - 0.735357) < 2e-6          # 1.055 * 0.5^(1/2.4) - 0.055
    assert abs(f[0, 1] - 0.040449936) < 1e-7       # the two segments meet at 0.04045
    # view mode 9 (emissive) is a colour view too; linear 0.214 encodes to
    # 0.49996, just below the 127/128 boundary, and 0.2142 just above it
    _, b, _ = post(px(0.2140, 0.2142, 0.18), 9)
    assert tuple(b[0, :3]) == (127, 128, 118)


def test_reinhard_only_for_the_path_tracer():
    _, b, f = post(px(1.0, 3.0, 0.0), 0)
    assert abs(f[0, 0] - 0.735357) < 2e-6 and tuple(b[0, :3]) == (188, 225, 0)   # srgb(0.5), srgb(0.75)
    _, b, _ = post(px(1.0, 3.0, 0.0), 6)       # no tonemap in a colour view: clamps to 1
    assert tuple(b[0, :3]) == (255, 255, 0)


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 7, 8, 10, 11])
def test_data_views_are_written_linearly(mode):
    _, b, f = post(px(0.5, -0.25, 2.0), mode)
    assert tuple(b[0]) == (128, 0, 255, 255)
    assert np.array_equal(f[0], np.float32([0.5, -0.25, 2.0, 1.0]))


def test_badness_becomes_magenta():
    e = np.array([[np.nan, 0.0, 0.0, 1.0], [0.0, np.inf, 0.0, 1.0], [0.0, 0.0, -np.inf, 1.0]], np.float32)
    acc, b, _ = post(e, 0)
    assert np.array_equal(acc[:, :3], np.tile(np.float32([1, 0, 1]), (3, 1)))
    assert np.all(b[:, :3] == (188, 0, 188))
    acc, b, _ = post(e, 4)   # also in a data view, before any tonemapping
    assert np.all(b[:, :3] == (255, 0, 255))


def test_accumulator_is_the_running_mean():
    rng = np.random.default_rng(3)
    frames = rng.uniform(0.0, 4.0, (12, 64, 4)).astype(np.float32)
    acc = np.zeros((64, 4), np.float32)
    for k, e in enumerate(frames):
        acc, _, _ = post(e, 0, accumulate=1, acc=acc, n=k + 1)
        mean = frames[:k + 1, :, :3].astype(np.float64).mean(axis=0)
        assert np.allclose(acc[:, :3], mean, rtol=2e-6, atol=0), k
        assert np.allclose(acc[:, 3], 1.0, rtol=1e-6)
    # weight 1 on the first frame discards whatever the accumulator held
    acc, _, _ = post(frames[0], 0, accumulate=1, acc=np.full((64, 4), 7.0, np.float32), n=1)
    assert np.array_equal(acc[:, :3], frames[0, :, :3])


@pytest.mark.parametrize("mode", range(1, 12))
def test_debug_views_do_not_accumulate(mode):
    old = np.full((4, 4), 9.0, np.float32)
    e = px(0.25, 0.5, 0.75).repeat(4, axis=0)
    acc, _, _ = post(e, mode, accumulate=1, acc=old, n=5)
    assert np.array_equal(acc, e)


def test_accumulate_off_writes_the_frame():
    old = np.full((1, 4), 9.0, np.float32)
    acc, _, _ = post(px(0.25, 0.5, 0.75), 0, accumulate=0, acc=old, n=5)
    assert np.array_equal(acc, px(0.25, 0.5, 0.75))
