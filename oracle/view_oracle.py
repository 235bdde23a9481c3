"""CPU restatement of the GUI's device kernels (the checker of flb_view.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product never
imports it.  Pure-Python loops in float32 arithmetic (numpy scalars), for the
small fields the tests use.

Follows:
  * CudaUtils::save2DDataToOpenGL          src/cuda/cudaUtils.cu:107-118
  * d_CreateIsoSurfaceVector2D             src/cuda/cudaIsosurface.cuh:33-198
  * getInterpolatedPoint, drawLine,
    marchingSquares                        src/cuda/cudaIsosurface.cu:16-86
Two deliberate differences from the reference, both also in the GPU kernel:
the velocity magnitude is sqrt(ux^2 + uy^2) rounded once from double (the
reference's device hypotf may differ by an ulp), and the fill-level mode
reads phi as a scalar (the reference reads a second plane past phi's end,
src/cuda/runCuda.cu:424).
"""
import numpy as np

f32 = np.float32

# Segments per marching-squares case in drawing order, as (edge, edge) pairs.
# Edge 0: vertices 0-1, 1: 1-2, 2: 3-2, 3: 0-3 (cudaIsosurface.cuh:93-196).
SEGMENTS = {
    1: [(0, 3)], 14: [(0, 3)],
    2: [(0, 1)], 13: [(0, 1)],
    3: [(3, 1)], 12: [(3, 1)],
    4: [(1, 2)], 11: [(1, 2)],
    5: [(0, 3), (2, 1)],
    6: [(0, 2)], 9: [(0, 2)],
    7: [(3, 2)], 8: [(3, 2)],
    10: [(3, 2), (0, 1)],
}
CORNERS = [(f32(0), f32(0)), (f32(1), f32(0)), (f32(1), f32(1)), (f32(0), f32(1))]
EDGES = {0: (0, 1), 1: (1, 2), 2: (3, 2), 3: (0, 3)}


def export_velocity(ux, uy):
    """(ny, nx, 2) float32 image: texel (x, y) = (ux, uy) of cell (x, y)."""
    return np.stack([np.asarray(ux, f32), np.asarray(uy, f32)], axis=-1)


def magnitude(ux, uy):
    return f32(np.sqrt(np.float64(ux) * np.float64(ux) + np.float64(uy) * np.float64(uy)))


def interp(p0, p1, v0, v1, iso):
    a = f32(f32(iso - v0) / f32(v1 - v0))
    b = f32(f32(1) - a)
    return (f32(f32(b * p0[0]) + f32(a * p1[0])), f32(f32(b * p0[1]) + f32(a * p1[1])))


def to_int(v):
    """C float -> int conversion: truncation toward zero."""
    return int(np.trunc(np.float64(v)))


def draw_line(img, p0, p1, ppn, xo, yo):
    height, width = img.shape
    xp0, xp1 = f32(p0[0] + xo), f32(p1[0] + xo)
    yp0, yp1 = f32(p0[1] + yo), f32(p1[1] + yo)
    x0 = to_int(f32(f32(f32(ppn[0] * xp0) - f32(float(to_int(xp0) != 0))) + f32(0.5)))
    x1 = to_int(f32(f32(f32(ppn[0] * xp1) - f32(float(to_int(xp1) != 0))) + f32(0.5)))
    h = f32(height)
    y0 = to_int(f32(f32(h - f32(ppn[1] * yp0)) - f32(0.5)))
    y1 = to_int(f32(f32(h - f32(ppn[1] * yp1)) - f32(0.5)))
    dx, dy = abs(x1 - x0), -abs(y1 - y0)
    err = dx + dy
    sx = 1 if x0 < x1 else -1
    sy = 1 if y0 < y1 else -1
    while x0 != x1 or y0 != y1:
        if 0 <= x0 < width and 0 <= y0 < height:
            img[y0, x0] = 255
        e2 = 2 * err
        if e2 >= dy:
            err += dy
            x0 += sx
        if e2 <= dx:
            err += dx
            y0 += sy


def isosurface(values, width, height, bounds, iso, img=None):
    """values: (ny, nx) float32 vertex values (|u| or phi), row 0 = lattice y = 0.
    bounds = (x_min, x_max, y_min, y_max) in display rows from the top.
    Returns the (height, width) uint8 image (rows = texture rows)."""
    v = np.asarray(values, f32)
    ny, nx = v.shape
    x_min, x_max, y_min, y_max = bounds
    if img is None:
        img = np.zeros((height, width), np.uint8)
    iso = f32(iso)
    ppn = (f32(f32(width) / f32(x_max - x_min)), f32(f32(height) / f32(y_max - y_min)))
    for y in range(y_min, min(y_max, ny - 1)):
        top = ny - 1 - y
        for x in range(x_min, min(x_max, nx - 1)):
            vv = [v[top, x], v[top, x + 1], v[top - 1, x + 1], v[top - 1, x]]
            idx = sum(int(vv[k] > iso) << k for k in range(4))
            for e0, e1 in SEGMENTS.get(idx, []):
                pts = []
                for e in (e0, e1):
                    a, b = EDGES[e]
                    pts.append(interp(CORNERS[a], CORNERS[b], vv[a], vv[b], iso))
                draw_line(img, pts[0], pts[1], ppn, f32(x - x_min), f32(y - y_min))
    return img


def magnitude_field(ux, uy):
    """Vectorised magnitude(): float32 of sqrt in double, no fused operations."""
    x = np.asarray(ux, np.float64)
    y = np.asarray(uy, np.float64)
    return np.sqrt(x * x + y * y).astype(f32)
