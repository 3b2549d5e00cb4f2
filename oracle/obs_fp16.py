"""TEST INFRASTRUCTURE: FP16 observation storage restated in numpy.

The solver's optional pixel storage (two IEEE binary16 values per observation
times a per-camera FP32 scale) has no counterpart in the reference's solver:
the reference only studied FP16 pixel storage offline, with a separate tool
(tools/quantize_bal.cu:99-116 writes a quantised BAL file; docs/RUNGS.md:44,50
record the cost change).  This module states the encoding the build uses so
tests can check the library's encoder bit for bit, and decode the pixels the
device sees.  numpy's float64 -> float16 conversion is its own IEEE 754
round-to-nearest-even implementation, independent of the library's.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker; the product never does.
"""
import numpy as np


def camera_scales(obs_camera, obs_pixel, num_cameras):
    """scale[c] = float32(m_c / 32768), m_c = the largest finite |pixel| of
    camera c's observations; 1 when m_c is 0, absent, or the scale is not a
    positive finite float32."""
    a = np.abs(np.asarray(obs_pixel, dtype=np.float64)).reshape(-1, 2)
    a = np.where(np.isfinite(a), a, 0.0).max(axis=1) if len(a) else np.zeros(0)
    m = np.zeros(num_cameras)
    np.maximum.at(m, np.asarray(obs_camera, dtype=np.int64), a)
    with np.errstate(over="ignore"):
        s = (m / 32768.0).astype(np.float32)
    ok = np.isfinite(s) & (s > 0)
    return np.where(ok, s, np.float32(1.0)).astype(np.float32)


def encode(obs_camera, obs_pixel, num_cameras):
    """(scale float32[C], px16 float16[O, 2]): px16 = binary16(pixel / scale),
    the quotient formed in float64 and rounded once to binary16."""
    scale = camera_scales(obs_camera, obs_pixel, num_cameras)
    s = scale.astype(np.float64)[np.asarray(obs_camera, dtype=np.int64)][:, None]
    with np.errstate(over="ignore", invalid="ignore"):
        px16 = (np.asarray(obs_pixel, dtype=np.float64).reshape(-1, 2) / s).astype(np.float16)
    return scale, px16


def decode(obs_camera, scale, px16):
    """Pixels as the device reads them: float64(half) * float64(scale), which is
    exact (an 11-bit times a 24-bit significand)."""
    s = np.asarray(scale, dtype=np.float32).astype(np.float64)[np.asarray(obs_camera, dtype=np.int64)]
    return np.asarray(px16, dtype=np.float16).astype(np.float64) * s[:, None]
