"""Learned cutoff surrogate: an MLP (6 -> 128 -> 128 -> 1, GELU) trained on
the GPU tracer's own cutoff maps, and the bracketed scan that uses it.

Features per (location, arrival direction): sin/cos latitude, sin/cos
longitude, zenith / 90 deg and azimuth / 180 deg - 1.  Target: log10 of the
upper cutoff rigidity (the allowed side of the highest forbidden -> allowed
transition on the scan grid).  At inference only the predicted bin +-3 is
traced, all bins of all brackets in parallel; the full scan is the fallback
wherever the bracket does not contain the transition.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from . import cutoff as CU

NPARAM = 17537
D_IN = 6


def features(lat_deg, lon_deg, zen_deg, az_deg):
    la = np.radians(np.asarray(lat_deg, np.float64))
    lo = np.radians(np.asarray(lon_deg, np.float64))
    z = np.asarray(zen_deg, np.float64) / 90.0
    a = np.asarray(az_deg, np.float64) / 180.0 - 1.0
    return np.stack([np.sin(la), np.cos(la), np.sin(lo), np.cos(lo), z, a], axis=-1).astype(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (128, D_IN)) * math.sqrt(6.0 / D_IN)
    W2 = rng.uniform(-1, 1, (128, 128)) * math.sqrt(6.0 / 128)
    w3 = rng.uniform(-1, 1, 128) * math.sqrt(3.0 / 128)
    return np.concatenate([W1.ravel(), np.zeros(128), W2.ravel(), np.zeros(128), w3, np.zeros(1)]).astype(np.float32)


def bin_target(bins, grid):
    """log10 of the cutoff rigidity for a scan result (-1: above Rmax)."""
    b = np.where(np.asarray(bins) < 0, len(grid) - 1, bins)
    return np.log10(np.asarray(grid)[b]).astype(np.float32)


def predicted_bin(y, grid):
    lg = np.log10(np.asarray(grid))
    return np.clip(np.searchsorted((lg[1:] + lg[:-1]) / 2, y), 0, len(grid) - 1).astype(np.int32)


class Surrogate:
    """The network's parameters and training data live on the device."""

    def __init__(self, ctx, theta=None, seed=0):
        self.ctx = ctx
        th = init_params(seed) if theta is None else np.ascontiguousarray(theta, np.float32)
        L.check(ctx.lib.amps_gpu_mlp_init(ctx.h, L._ptr(th, C.c_float)), "mlp_init")

    def params(self):
        th = np.zeros(NPARAM, np.float32)
        st = C.c_int64()
        L.check(self.ctx.lib.amps_gpu_mlp_get_params(self.ctx.h, L._ptr(th, C.c_float), C.byref(st)), "mlp_get_params")
        return th, st.value

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        L.check(self.ctx.lib.amps_gpu_mlp_set_data(self.ctx.h, len(T), L._ptr(X, C.c_float),
                                                   L._ptr(T, C.c_float)), "mlp_set_data")

    def grad(self, n):
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_float()
        L.check(self.ctx.lib.amps_gpu_mlp_grad(self.ctx.h, n, L._ptr(g, C.c_float), C.byref(loss)), "mlp_grad")
        return g, loss.value

    def train(self, nsteps, batch=4096, lr=1e-3, record_loss=False):
        lo = np.zeros(nsteps, np.float32) if record_loss else None
        L.check(self.ctx.lib.amps_gpu_mlp_train(self.ctx.h, nsteps, batch, lr, L._ptr(lo, C.c_float)), "mlp_train")
        return lo

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(len(X), np.float32)
        L.check(self.ctx.lib.amps_gpu_mlp_predict(self.ctx.h, len(X), L._ptr(X, C.c_float),
                                                  L._ptr(y, C.c_float)), "mlp_predict")
        return y


def random_tasks(n, seed=6, r_Re=1.05, max_zenith_deg=60.0):
    """n random (location, arrival direction) pairs of the surrogate's
    training set: locations uniform on the sphere of radius r_Re dipole
    radii, zenith uniform in cos over [0, max_zenith], azimuth uniform.
    Returns (x0 [n, 3] m, dirs [n, 3], features [n, 6])."""
    rng = np.random.default_rng(seed)
    lat = np.degrees(np.arcsin(rng.uniform(-1, 1, n)))
    lon = rng.uniform(0, 360, n)
    zen = np.degrees(np.arccos(rng.uniform(math.cos(math.radians(max_zenith_deg)), 1.0, n)))
    az = rng.uniform(0, 360, n)
    la, lo = np.radians(lat), np.radians(lon)
    r0 = r_Re * CU.RE_DIPOLE
    x0 = np.stack([r0 * np.cos(la) * np.cos(lo), r0 * np.cos(la) * np.sin(lo), r0 * np.sin(la)], 1)
    dirs = -CU.local_dirs_paired(x0, az, 90.0 - zen)
    return x0, dirs, features(lat, lon, zen, az)


def training_set(ctx, params, n, grid, seed=6, precision=L.FP32):
    """The surrogate's training set from the tracer itself: the upper cutoff
    bin of n random pairs by the downward scan.  Returns (features, targets,
    bins, steps)."""
    x0, dirs, X = random_tasks(n, seed)
    _, bins, steps = CU.cutoff_scan(ctx, params, x0, dirs, grid, mode=CU.SCAN_UPPER, precision=precision)
    return X, bin_target(bins, grid), bins, steps


def bracket_bins(pb, nR, half=3):
    """Fixed-width brackets [lo, lo + 2 half] around the predicted bins,
    shifted inward at the ends of the grid."""
    w = 2 * half + 1
    lo = np.clip(np.asarray(pb) - half, 0, nR - w).astype(np.int32)
    return lo, (lo + w - 1).astype(np.int32)


def bracketed_map(ctx, sur, params, x0, dirs, feats, grid, half=3, precision=L.FP32, one_launch=True):
    """Cutoff bins with the surrogate bracket and the full-scan fallback.

    Every bin of every bracket is traced as an independent trajectory; the
    bins come from the bracket's samples as CutoffForDirectionUpperScan_GV
    would read them.  Points whose bracket shows no transition (the top
    sample forbidden below Rmax, or every sample allowed above Rmin) are
    rescanned over the whole grid.  one_launch: the rescan runs inside the
    same persistent launch, each pair's other bins starting as soon as its
    bracket is done (AMPS_SCAN_BRACKETED); otherwise a second launch follows
    the bracket scan.  Returns (bins, fallback mask, predicted bins, total
    steps)."""
    y = sur.predict(feats)
    pb = predicted_bin(y, grid)
    lo, hi = bracket_bins(pb, len(grid), half)
    if one_launch:
        states, bins, steps = CU.cutoff_scan(ctx, params, x0, dirs, grid, mode=CU.SCAN_BRACKETED, lo=lo, hi=hi,
                                             precision=precision)
        outside = np.ones(states.shape, bool)
        cols = np.arange(len(grid))[None, :]
        outside &= (cols < lo[:, None]) | (cols > hi[:, None])
        fb = ((states >= 0) & outside).any(axis=1)
        return bins, fb, pb, steps
    _, bins, steps = CU.cutoff_scan(ctx, params, x0, dirs, grid, mode=CU.SCAN_FULL, lo=lo, hi=hi,
                                    precision=precision)
    fb = bins < -1
    if fb.any():
        idx = np.flatnonzero(fb)
        _, b2, s2 = CU.cutoff_scan(ctx, params, x0[idx], dirs[idx], grid, mode=CU.SCAN_FULL,
                                   precision=precision)
        bins = bins.copy()
        bins[idx] = b2
        steps += s2
    return bins, fb, pb, steps
