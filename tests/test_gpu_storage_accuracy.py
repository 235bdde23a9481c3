"""Accuracy of the storage formats on the free-surface dam break, measured on
physical quantities: the discharge through the culvert barrel (sum over the
barrel column of phi rho u_x, time-averaged), the water-surface profile (per
column, the height where the fill level crosses 0.5), the position of the
surge front and the velocity.

* 1024 x 256 (the 4096 x 1024 case / 4): FP32, FP16S and BF16S storage on the
  GPU against the FP64 oracle at the same size, 4000 steps;
* 8192 x 2048: FP16S and BF16S against FP32 on the GPU over the first 3000
  steps -- the collapse of the 1200-cell water column up to the moment the
  surge reaches the embankment -- comparing front position, surface heights
  averaged over 8 columns and velocity averaged over 8 x 8 boxes.  Past that
  moment the surge meets the wall at lattice speeds near 0.3, which SRT at
  tau = 0.51 does not survive (the FP32 run turns to NaN between steps 4000
  and 5000), so the barrel discharge is measured on the small
  case only.  FP16S stays on top of FP32 there (same front, 0.07 cells mean
  surface difference, 0.7 % velocity difference).  BF16S departs visibly (front
  7 cells behind, 3.9 cells mean surface difference, 21 % velocity difference):
  under 1200 cells of water rho - 1 ~ 0.07, where BF16's rounding step of the
  shifted populations (~1e-5) is of the order of the hydrostatic density step
  between neighbouring rows (3 g = 6e-5), so the pressure field is noisy.  The
  bars hold those measured levels with margin; BF16S is not recommended for
  deep free-surface runs.

The runs use the corrected reconstruction gate (gas_check_corrected = 1): with
the reference's gate the water column of this case barely moves (its front
advances ~2 cells per 1000 steps) and there is no flow to measure.  At
tau = 0.51 the breaking wave of the small case is chaotic: any rounding
difference decorrelates the instantaneous fields (FP32 and FP64 differ by
~100 cells in the local surface height and ~30 % in velocity after 4000
steps), so that test bounds the statistics that survive -- the time-averaged
discharge and the mean surface-height difference -- and prints the rest.  A
storage format passes when it stays as close to the reference run as FP32
does."""
import numpy as np
import pytest

from flowlb_amd import cases, core

pytestmark = pytest.mark.gpu

SMALL = cases.FS_1024
STEPS, EVERY, T_AVG = 4000, 50, 2000          # 1024 x 256
BIG_STEPS = 3000                                    # 8192 x 2048


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


def run_gpu(case, storage, column, steps=STEPS, every=EVERY, t_avg=T_AVG):
    qs = []
    with case.solver(storage, mass_mode=1, gas_check_corrected=1) as s:
        for t in range(every, steps + 1, every):
            s.step(every)
            if t > steps - t_avg:
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


@pytest.mark.parametrize("storage,max_mean_dh,max_dq", [
    ("fp32", 1.5, 3e-3), ("fp16s", 1.5, 3e-3), ("bf16s", 2.0, 5e-3)])
def test_small_against_fp64_oracle(reference_small, storage, max_mean_dh, max_dq):
    (r, q_ref), column = reference_small
    g, q = run_gpu(SMALL, storage, column)
    nx, ny = SMALL.nx, SMALL.ny
    dh = np.abs(surface(g["phi"], nx, ny) - surface(r["phi"], nx, ny))
    du = float(np.linalg.norm(g["u"] - r["u"]) / np.linalg.norm(r["u"]))
    dq = abs(q - q_ref) / abs(q_ref)
    print(f"\n{storage} 1024x256 x {STEPS} vs FP64 oracle: surface max|dh| {dh.max():.3f} "
          f"mean {dh.mean():.4f} cells; discharge {q:.5e} vs {q_ref:.5e} (rel {dq:.2e}); "
          f"relL2 u {du:.2e}")
    assert abs(q_ref) > 0.5             # the barrel flows
    assert dh.mean() <= max_mean_dh and dq <= max_dq
    assert du < 1.0                     # decorrelated, not blown up


def box(a, nx, ny, k):
    return a.reshape(ny // k, k, nx // k, k).mean(axis=(1, 3))


def front(flags, nx, ny):
    """Largest x of a wet column (the surge front)."""
    wet = (flags.reshape(ny, nx) & (core.FLUID | core.INTERFACE)) != 0
    return int(np.nonzero(wet.any(axis=0))[0].max())


Q_COLUMN, Q_FROM, Q_EVERY = 3620, 1000, 100     # 8192 x 2048 discharge sampling


def test_large_sixteen_bit_against_fp32():
    """8192 x 2048, 3000 steps of the collapsing column: the storage formats
    against FP32 on the front position, the surface heights, box-averaged
    velocity and the time-averaged discharge through column 3620, which the
    front crosses before step 1000 (mean over steps 1000-3000, every 100).
    The barrel itself (x = 4000-4400) is reached only after ~4500 steps, when
    this SRT run at tau = 0.51 has already turned to NaN.  Measured: FP16S
    discharge 0.51 % below FP32's; the bar, 1 %, guards that level (the 0.5 %
    target is against the FP64 oracle: bench.py --workload accuracy_8192x2048).
    BF16S (discharge 13.6 % off FP32 at this depth) is printed and held only
    to its measured levels."""
    case = cases.FS_8192
    nx, ny = case.nx, case.ny
    out = {}
    for st in ("fp32", "fp16s", "bf16s"):
        qs = []
        with case.solver(st, mass_mode=1, gas_check_corrected=1) as s:
            m0 = s.mass()
            s.step(Q_FROM)
            for t in range(Q_FROM, BIG_STEPS + 1, Q_EVERY):
                if t > Q_FROM:
                    s.step(Q_EVERY)
                g = s.read()
                qs.append(discharge(g, g["flags"], nx, ny, Q_COLUMN))
            out[st] = (g, abs(s.mass() - m0) / m0, float(np.mean(qs)))
    g32, _, q32 = out["fp32"]
    ux32 = box(g32["u"][:nx * ny], nx, ny, 8)
    h32 = surface(g32["phi"], nx, ny).reshape(-1, 8).mean(axis=1)
    x32 = front(g32["flags"], nx, ny)
    assert x32 > case.water[0] + 100    # the column has collapsed
    assert q32 > 1.0                    # water flows through the sampled column
    for st, (max_dx, max_mean_dh, max_du, max_dq) in {"fp16s": (4, 0.5, 3e-2, 1e-2),
                                                      "bf16s": (24, 8.0, 0.4, 0.25)}.items():
        g, drift, q = out[st]
        ux = box(g["u"][:nx * ny], nx, ny, 8)
        h = surface(g["phi"], nx, ny).reshape(-1, 8).mean(axis=1)
        du = float(np.linalg.norm(ux - ux32) / np.linalg.norm(ux32))
        dh = np.abs(h - h32)
        x = front(g["flags"], nx, ny)
        dx = abs(x - x32)
        dq = abs(q - q32) / abs(q32)
        print(f"\n{st} 8192x2048 x {BIG_STEPS} vs FP32: front {x} vs {x32}; "
              f"surface (8-column mean) max|dh| {dh.max():.3f} mean {dh.mean():.4f} cells; "
              f"box-averaged relL2 u_x {du:.2e}; discharge at x = {Q_COLUMN} {q:.5e} vs "
              f"{q32:.5e} (rel {dq:.2e}); mass drift {drift:.2e} (fp32 {out['fp32'][1]:.2e})")
        assert dx <= max_dx and dh.mean() <= max_mean_dh and du <= max_du
        assert dq <= max_dq
        assert drift <= 1e-6
