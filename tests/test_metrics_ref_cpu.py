"""Known answers for the restated PSNR harness (tests/metrics.rs:40-114)."""
import math

import numpy as np

from metrics_ref import calculate_image_metrics_rgba8, srgb_to_linear_f64


def test_srgb_to_linear_known_values():
    assert srgb_to_linear_f64(0) == 0.0
    assert srgb_to_linear_f64(255) == 1.0
    assert srgb_to_linear_f64(10) == (10 / 255.0) / 12.92   # below the 0.04045 knee
    assert abs(srgb_to_linear_f64(128) - 0.2158605001138992) < 1e-15


def test_identical_images_report_zero_psnr():
    img = np.random.default_rng(1).integers(0, 256, (8, 8, 4), dtype=np.uint8)
    m = calculate_image_metrics_rgba8(img, img, 4)
    assert m["overall_mse"] == 0.0 and m["overall_psnr"] == 0.0   # metrics.rs:80-81


def test_one_channel_difference_closed_form():
    a = np.zeros((4, 8, 4), np.uint8)
    b = a.copy()
    b[0, 0, 3] = 51                        # alpha: 0.2 on one pixel of 32
    m = calculate_image_metrics_rgba8(a, b, 4)
    want = (0.2 ** 2 / 32) / 4
    assert abs(m["overall_mse"] - want) < 1e-18
    assert abs(m["overall_psnr"] - 20 * math.log10(1 / math.sqrt(want))) < 1e-12
    # With 3 channels alpha is not counted.
    assert calculate_image_metrics_rgba8(a, b, 3)["overall_psnr"] == 0.0
    b[1, 1, 0] = 255                       # red: linear 1.0 on one pixel
    m = calculate_image_metrics_rgba8(a, b, 3)
    assert abs(m["overall_mse"] - (1.0 / 32) / 3) < 1e-18
