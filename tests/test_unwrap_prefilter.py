"""The list builder's unwrapped FP32 prefilter (md_kernels.hip brick_build,
`unwrap`): on a grid whose axes are all tiled into bricks (brick_dims,
md_device.h: an axis is tiled when it has more cells than a brick's edge
plus its two halo cells), each staged particle's FP32 copy is its periodic
image next to its halo cell, and the prefilter takes plain differences
instead of the minimum image (potential.rs:186-200).  The claim it rests on:
for a centre particle and a particle in its stencil (halo coordinates at
most one apart) within r_list (<= one cell edge), the unwrapped difference
equals the minimum-image difference.

Checked here by brute force on random periodic configurations along one
axis, with the halo unwrapping restated from the kernel (a halo coordinate
org + h below 0, or at or past the grid size, shifts the particle by -L or
+L); and, for contrast, a whole axis (the halo is the entire axis, its
stencil wraps across the seam), where plain differences are wrong and the
kernel keeps the minimum image.
"""
import numpy as np
import pytest

L = 7.3


def min_image(d):
    # potential.rs:186-200: one correction against L / 2
    return np.where(d < -L / 2, d + L, np.where(d > L / 2, d - L, d))


def config(gd, rng, n=3000):
    edge = L / gd
    x = rng.uniform(0.0, L, size=n)
    cell = np.minimum((x / edge).astype(int), gd - 1)
    rl = edge * rng.uniform(0.6, 1.0)  # r_list <= one cell edge
    return x, cell, rl


def tiled_violations(gd, seed):
    rng = np.random.default_rng(seed)
    bad = checked = 0
    for _ in range(30):
        x, cell, rl = config(gd, rng)
        be = int(rng.integers(1, gd - 2))          # tiled: gd > be + 2
        ncen = int(rng.integers(1, be + 1))        # a brick, sub-brick or cell
        hd = ncen + 2
        org = int(rng.integers(-1, gd - 1))        # the region's first halo cell
        h = (cell - org) % gd                      # halo coordinate
        staged = h < hd
        g = org + h
        xu = x + np.where(g < 0, -L, np.where(g >= gd, L, 0.0))
        ci = np.flatnonzero(staged & (h >= 1) & (h <= ncen))  # centre particles
        cj = np.flatnonzero(staged)
        dh = h[cj][None, :] - h[ci][:, None]
        dm = min_image(x[cj][None, :] - x[ci][:, None])
        close = (np.abs(dh) <= 1) & (np.abs(dm) <= rl) & (ci[:, None] != cj[None, :])
        du = xu[cj][None, :] - xu[ci][:, None]
        checked += int(close.sum())
        bad += int(np.count_nonzero(np.abs(du - dm)[close] > 1e-9))
    return bad, checked


@pytest.mark.parametrize("gd", [4, 5, 7, 16, 31])
def test_unwrapped_difference_is_the_minimum_image(gd):
    bad, checked = tiled_violations(gd, seed=gd)
    assert checked > 500
    assert bad == 0


def test_whole_axis_needs_the_minimum_image():
    # a whole axis: the halo is every cell (org 0, no shift) and the
    # stencil of the last cell wraps to the first, so plain differences of
    # such pairs are a box length off
    rng = np.random.default_rng(1)
    gd = 5
    x, cell, rl = config(gd, rng)
    dc = (cell[None, :] - cell[:, None]) % gd
    adjacent = (dc == 0) | (dc == 1) | (dc == gd - 1)
    dm = min_image(x[None, :] - x[:, None])
    close = adjacent & (np.abs(dm) <= rl)
    np.fill_diagonal(close, False)
    du = x[None, :] - x[:, None]
    assert np.count_nonzero(np.abs(du - dm)[close] > 1e-9) > 0
