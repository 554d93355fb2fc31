"""The reference's image-quality harness, restated for the tests.

calculate_image_metrics_rgba8 follows tests/metrics.rs:40-114 line by line:
per-channel MSE over all pixels, RGB in linear light through the harness's
float64 srgb_to_linear (tests/common/mod.rs:17-24), alpha as a/255; overall
MSE is the mean of the first `channels` channel MSEs; PSNR = 20 log10(1 /
sqrt(mse)), and 0.0 when the MSE is 0 (metrics.rs:80-81).  compare_psnr's
tolerance is 0.0025 dB (metrics.rs:245).
"""
import math

import numpy as np

PSNR_TOLERANCE_DB = 0.0025


def srgb_to_linear_f64(c: int) -> float:
    v = c / 255.0
    return v / 12.92 if v <= 0.04045 else math.pow((v + 0.055) / 1.055, 2.4)


_LUT = np.array([srgb_to_linear_f64(c) for c in range(256)], np.float64)


def _psnr(mse: float) -> float:
    return 0.0 if mse == 0.0 else 20.0 * math.log10(1.0 / math.sqrt(mse))


def calculate_image_metrics_rgba8(original: np.ndarray, compressed: np.ndarray, channels: int) -> dict:
    """(h, w, 4) uint8 images -> {"overall_psnr", "overall_mse", "channel_mse", "channel_psnr"}."""
    if original.shape != compressed.shape or original.shape[-1] != 4:
        raise ValueError("images must be RGBA of the same size")
    o = original.reshape(-1, 4)
    c = compressed.reshape(-1, 4)
    n = float(o.shape[0])
    mse = [0.0] * 4
    for ch in range(channels):
        if ch < 3:
            d = _LUT[o[:, ch]] - _LUT[c[:, ch]]
        else:
            d = o[:, ch].astype(np.float64) / 255.0 - c[:, ch].astype(np.float64) / 255.0
        mse[ch] = float(np.sum(d * d)) / n
    overall = sum(mse) / channels
    return {"overall_psnr": _psnr(overall), "overall_mse": overall, "channel_mse": mse,
            "channel_psnr": [_psnr(m) for m in mse]}
