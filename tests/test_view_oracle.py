"""Pins the CPU restatement of the viewer kernels (oracle/view_oracle.py) with
closed forms: isolines of linear ramps are straight pixel lines at the
column / row the interpolation puts them on, a saddle square draws its two
diagonals, Bresenham's line stops one pixel short of its end point
(src/cuda/cudaIsosurface.cu:23-74), and the export is the identity."""
import numpy as np
import pytest

import view_oracle as vo

f32 = np.float32


def test_export_is_identity():
    rng = np.random.default_rng(0)
    ux, uy = rng.normal(size=(5, 7)).astype(f32), rng.normal(size=(5, 7)).astype(f32)
    img = vo.export_velocity(ux, uy)
    assert img.shape == (5, 7, 2)
    np.testing.assert_array_equal(img[..., 0], ux)
    np.testing.assert_array_equal(img[..., 1], uy)


def test_magnitude_matches_hypot():
    rng = np.random.default_rng(1)
    for a, b in rng.uniform(-0.3, 0.3, (200, 2)).astype(f32):
        assert vo.magnitude(a, b) == pytest.approx(float(np.hypot(np.float64(a), np.float64(b))),
                                                   rel=6e-8)
        assert vo.magnitude(a, f32(0)) == abs(a)


@pytest.mark.parametrize("bounds", [(0, 24, 0, 16), (3, 20, 2, 11)])
def test_vertical_line_of_x_ramp(bounds):
    """values = x, iso = 10.5: every square at x = 10 is case 6 with both
    points at x + 0.5; the line is one texture column spanning the window."""
    nx, ny, W, H = 24, 16, 96, 64
    v = np.tile(np.arange(nx, dtype=f32), (ny, 1))
    img = vo.isosurface(v, W, H, bounds, 10.5)
    x0, x1, y0, y1 = bounds
    ppx, ppy = f32(W) / f32(x1 - x0), f32(H) / f32(y1 - y0)
    xo = f32(10 - x0) + f32(0.5)
    col = int(ppx * xo - f32(1.0) + f32(0.5))
    rows_sq = min(y1, ny - 1) - y0
    top = int(f32(H) - f32(0.5))                       # yo = 0
    bottom = int(f32(H) - ppy * f32(rows_sq) - f32(0.5))  # end point, not drawn
    lit = np.argwhere(img == 255)
    assert set(lit[:, 1]) == {col}
    assert sorted(lit[:, 0]) == list(range(bottom + 1, top + 1))


def test_horizontal_line_of_y_ramp():
    """values = lattice y, iso = 7.5: display row ny-9 holds the squares
    between lattice rows 8 and 7 (case 3, points at yo + 0.5); after the flip
    to GL's lower-left origin the line is one texture row, drawn from the
    left edge of the window up to one pixel short of the right end."""
    nx, ny, W, H = 20, 16, 80, 64
    v = np.tile(np.arange(ny, dtype=f32)[:, None], (1, nx))
    img = vo.isosurface(v, W, H, (0, nx, 0, ny), 7.5)
    ppy = f32(H) / f32(ny)
    ppx = f32(W) / f32(nx)
    yo = f32(ny - 9) + f32(0.5)
    row = int(f32(H) - ppy * yo - f32(0.5))
    last = int(ppx * f32(nx - 1) - f32(1.0) + f32(0.5))
    lit = np.argwhere(img == 255)
    assert set(lit[:, 0]) == {row}
    assert sorted(lit[:, 1]) == list(range(0, last))


def test_saddle_draws_two_diagonals():
    """One square with v0 = v2 = 1, v1 = v3 = 0 (case 5) on a 40 x 40 texture:
    segments (0.5, 0)-(0, 0.5) and (0.5, 1)-(1, 0.5), i.e. the pixel
    diagonals (20, 39) -> (0, 19) and (20, 0) -> (39, 19), each without its
    end pixel."""
    v = np.array([[0, 1], [1, 0]], f32)   # row 0 = lattice y = 0 (bottom)
    img = vo.isosurface(v, 40, 40, (0, 1, 0, 1), 0.5)
    want = {(39 - k, 20 - k) for k in range(20)} | {(k, 20 + k) for k in range(19)}
    assert {tuple(p) for p in np.argwhere(img == 255)} == want
    # The other saddle (case 10): segments (0, 0.5)-(0.5, 1) and (0.5, 0)-(1, 0.5),
    # pixels (0, 19) -> (20, 0) and (20, 39) -> (39, 19): 20 steps each, the
    # start pixels drawn and the end pixels not.
    v10 = np.array([[1, 0], [0, 1]], f32)
    img = vo.isosurface(v10, 40, 40, (0, 1, 0, 1), 0.5)
    assert (img == 255).sum() == 40
    assert img[19, 0] == 255 and img[39, 20] == 255
    assert img[0, 20] == 0 and img[19, 39] == 0


def test_window_outside_domain_draws_nothing_beyond_edge():
    """Squares exist only for x < nx - 1, y < ny - 1: a window reaching past
    the domain draws the squares inside it and nothing else."""
    nx, ny = 12, 10
    v = np.tile(np.arange(nx, dtype=f32), (ny, 1))
    a = vo.isosurface(v, 60, 50, (0, 20, 0, 15), 10.5)   # x = 10 is the last square column
    assert (a == 255).any()
    b = vo.isosurface(v, 60, 50, (0, 20, 0, 15), 11.5)   # would need squares at x = 11
    assert not (b == 255).any()
