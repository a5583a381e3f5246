"""Image comparison used by the frame parity tests.

mean_rel_err is the mean over pixels of the per-pixel relative error
|got - ref|_1 / max(|ref|_1, floor) over the RGB channels, with the floor
DARK_FLOOR x the reference image's mean |ref|_1: a pixel far darker than the
image's mean is measured against that level instead of its own near-zero
value (a black pixel of the reference that the GPU also leaves black counts
0).  The bar is MEAN_REL_TOL = 1e-3.  bit_identical is reported beside it:
the fraction of pixels whose three channels are equal as FP32 words.
"""
import numpy as np

MEAN_REL_TOL = 1e-3
DARK_FLOOR = 1e-2


def mean_rel_err(got: np.ndarray, ref: np.ndarray, floor: float = DARK_FLOOR) -> float:
    g = np.asarray(got, np.float64)[..., :3].reshape(-1, 3)
    r = np.asarray(ref, np.float64)[..., :3].reshape(-1, 3)
    ref_l1 = np.abs(r).sum(axis=1)
    level = floor * max(float(ref_l1.mean()), 1e-30)
    return float(np.mean(np.abs(g - r).sum(axis=1) / np.maximum(ref_l1, level)))


def bit_identical(got: np.ndarray, ref: np.ndarray) -> float:
    g = np.ascontiguousarray(got, np.float32)[..., :3].reshape(-1, 3)
    r = np.ascontiguousarray(ref, np.float32)[..., :3].reshape(-1, 3)
    return float(np.mean(np.all(g.view(np.uint32) == r.view(np.uint32), axis=1)))


def assert_image_close(got, ref, what: str = "", tol: float = MEAN_REL_TOL):
    err = mean_rel_err(got, ref)
    same = bit_identical(got, ref)
    print(f"{what}: mean relative per-pixel error {err:.3e}, bit-identical pixels {same:.5f}")
    assert err <= tol, f"{what}: mean relative per-pixel error {err:.3e} > {tol:g} (bit-identical pixels {same:.5f})"
    return err, same
