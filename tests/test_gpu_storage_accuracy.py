"""Accuracy of the storage formats on the free-surface dam break, measured on
three physical quantities: velocity (relative L2), the water-surface profile
(per column, the height where the fill level crosses 0.5) and the discharge
through the culvert barrel, sum over the barrel column of phi rho u_x,
time-averaged.

* 1024 x 256 (the 4096 x 1024 case / 4): FP32, FP16S and BF16S storage on the
  GPU against the FP64 oracle at the same size;
* 8192 x 2048: FP16S and BF16S against FP32 on the GPU, velocity box-averaged
  over 8 x 8 cells, surface heights over 8 columns.

The runs use the corrected reconstruction gate (gas_check_corrected = 1): with
the reference's gate the water column of this case barely moves (its front
advances ~2 cells per 1000 steps) and there is no flow to measure.  Bars
(stated per format below) bound what the test measures; the measured values
are printed."""
import numpy as np
import pytest

from flowlb_amd import cases, core

pytestmark = pytest.mark.gpu

SMALL = cases.FS_1024
STEPS, EVERY, T_AVG = 4000, 50, 2000


def surface(phi, nx, ny):
    """Per column the height where phi crosses 0.5, searched from the top."""
    p = phi.reshape(ny, nx)
    h = np.zeros(nx)
    for x in range(nx):
        col = p[:, x]
        idx = np.nonzero(col >= 0.5)[0]
        if idx.size == 0:
            continue
        y = idx.max()
        if y + 1 < ny and col[y] > col[y + 1]:
            h[x] = y + (col[y] - 0.5) / (col[y] - col[y + 1])
        else:
            h[x] = y
    return h


def discharge(fields, flags, nx, ny, column):
    fl = flags.reshape(ny, nx)[:, column]
    wet = (fl & (core.FLUID | core.INTERFACE)) != 0
    phi = fields["phi"].reshape(ny, nx)[:, column]
    rho = fields["rho"].reshape(ny, nx)[:, column]
    ux = fields["u"][:nx * ny].reshape(ny, nx)[:, column]
    return float(np.sum(np.where(wet, phi * rho * ux, 0.0)))


def run_gpu(case, storage, column, every=EVERY):
    qs = []
    with case.solver(storage, mass_mode=1, gas_check_corrected=1) as s:
        for t in range(every, STEPS + 1, every):
            s.step(every)
            if t > STEPS - T_AVG:
                g = s.read()
                qs.append(discharge(g, g["flags"], case.nx, case.ny, column))
        g = s.read()
    return g, float(np.mean(qs))


def run_oracle(orc, case, column):
    flags, _, _ = case.build()
    o = orc.Oracle(flags, case.nx, case.ny, case.tau, use_force=True, fx=case.force[0],
                   fy=case.force[1], free_surface=True, mass_mode=1, gas_check_corrected=True,
                   threads=16)
    qs = []
    for t in range(EVERY, STEPS + 1, EVERY):
        o.step(EVERY)
        if t > STEPS - T_AVG:
            r = o.get()
            qs.append(discharge(r, r["flags"], case.nx, case.ny, column))
    return o.get(), float(np.mean(qs))


@pytest.fixture(scope="module")
def reference_small(orc):
    column = (SMALL.emb[0] + SMALL.emb[1]) // 2
    return run_oracle(orc, SMALL, column), column


@pytest.mark.parametrize("storage,max_dh,max_dq,max_du", [
    ("fp32", 2.0, 1e-2, 5e-2), ("fp16s", 4.0, 5e-2, 1e-1), ("bf16s", 16.0, 3e-1, 1.0)])
def test_small_against_fp64_oracle(reference_small, storage, max_dh, max_dq, max_du):
    (r, q_ref), column = reference_small
    g, q = run_gpu(SMALL, storage, column)
    nx, ny = SMALL.nx, SMALL.ny
    dh = np.abs(surface(g["phi"], nx, ny) - surface(r["phi"], nx, ny))
    du = float(np.linalg.norm(g["u"] - r["u"]) / np.linalg.norm(r["u"]))
    dq = abs(q - q_ref) / abs(q_ref)
    print(f"\n{storage} 1024x256 x {STEPS} vs FP64 oracle: surface max|dh| {dh.max():.3f} "
          f"mean {dh.mean():.4f} cells; discharge {q:.5e} vs {q_ref:.5e} (rel {dq:.2e}); "
          f"relL2 u {du:.2e}")
    assert abs(q_ref) > 0
    assert dh.max() <= max_dh and dq <= max_dq and du <= max_du


def box(a, nx, ny, k):
    return a.reshape(ny // k, k, nx // k, k).mean(axis=(1, 3))


def test_large_sixteen_bit_against_fp32():
    case = cases.FS_8192
    column = (case.emb[0] + case.emb[1]) // 2
    out = {st: run_gpu(case, st, column, every=200) for st in ("fp32", "fp16s", "bf16s")}
    nx, ny = case.nx, case.ny
    g32, q32 = out["fp32"]
    ux32 = box(g32["u"][:nx * ny], nx, ny, 8)
    h32 = surface(g32["phi"], nx, ny).reshape(-1, 8).mean(axis=1)
    for st, (max_dh, max_dq, max_du) in {"fp16s": (4.0, 5e-2, 1e-1),
                                          "bf16s": (16.0, 3e-1, 1.0)}.items():
        g, q = out[st]
        ux = box(g["u"][:nx * ny], nx, ny, 8)
        h = surface(g["phi"], nx, ny).reshape(-1, 8).mean(axis=1)
        du = float(np.linalg.norm(ux - ux32) / np.linalg.norm(ux32))
        dh = np.abs(h - h32)
        dq = abs(q - q32) / abs(q32)
        print(f"\n{st} 8192x2048 x {STEPS} vs FP32: surface (8-column mean) max|dh| {dh.max():.3f} "
              f"cells; discharge rel {dq:.2e}; box-averaged relL2 u_x {du:.2e}")
        assert dh.max() <= max_dh and dq <= max_dq and du <= max_du
