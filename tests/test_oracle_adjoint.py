"""Pins the oracle's FP64 discrete adjoint (single-phase step with Brinkman
drag) with central finite differences of the forward objective."""
import numpy as np
import pytest

from flowlb_amd import cases

INLET = 0x20


def small_case():
    c = cases.scaled(cases.ADJOINT, 8)          # 128 x 32, opening 8 at x = 50..60
    flags, rho0, u0 = c.build()
    rng = np.random.default_rng(7)
    alpha = np.zeros(c.N)
    win = [(x, y) for y in range(1, 9) for x in range(40, 48)]
    for x, y in win:
        alpha[x + y * c.nx] = rng.uniform(0, 1)
    return c, flags, rho0.astype(np.float64), u0.astype(np.float64), alpha, win


def make(orc, c, flags, rho0, u0, alpha, uin=None):
    u0 = u0.copy()
    if uin is not None:
        u0[:c.N][(flags & INLET) != 0] = uin
    return orc.Oracle(flags, c.nx, c.ny, c.tau, use_force=True, fx=c.force[0], fy=c.force[1],
                      rho0=rho0, u0=u0, alpha=alpha, threads=4)


@pytest.mark.parametrize("ckpt", [7, 1000])
def test_adjoint_matches_finite_differences(orc, ckpt):
    c, flags, rho0, u0, alpha, win = small_case()
    steps, t0, t1, col = 240, 160, 239, c.nx - 2
    o = make(orc, c, flags, rho0, u0, alpha)
    J, g, duin = o.adjoint(steps, col, t0, t1, ckpt=ckpt)
    assert abs(J - o.objective(steps, col, t0, t1)) < 1e-14 * abs(J)
    h = 1e-4
    rng = np.random.default_rng(3)
    for k in rng.choice(len(win), 5, replace=False):
        x, y = win[k]
        n = x + y * c.nx
        ap, am = alpha.copy(), alpha.copy()
        ap[n] += h
        am[n] -= h
        fd = (make(orc, c, flags, rho0, u0, ap).objective(steps, col, t0, t1)
              - make(orc, c, flags, rho0, u0, am).objective(steps, col, t0, t1)) / (2 * h)
        assert abs(g[n] - fd) <= 1e-6 * abs(fd) + 1e-12, (x, y, g[n], fd)
    uin = float(np.float32(c.u_inlet))
    hu = 1e-6
    fd = (make(orc, c, flags, rho0, u0, alpha, uin + hu).objective(steps, col, t0, t1)
          - make(orc, c, flags, rho0, u0, alpha, uin - hu).objective(steps, col, t0, t1)) / (2 * hu)
    assert abs(duin - fd) <= 1e-6 * abs(fd), (duin, fd)
    # cells outside the Brinkman window have zero alpha but still a gradient
    assert np.abs(g).max() > 0


def test_adjoint_without_brinkman_or_force(orc):
    c, flags, rho0, u0, _, _ = small_case()
    o = orc.Oracle(flags, c.nx, c.ny, c.tau, rho0=rho0, u0=u0, threads=4)
    steps, t0, t1, col = 120, 60, 119, c.nx - 2
    J, g, duin = o.adjoint(steps, col, t0, t1, ckpt=50)
    uin = float(np.float32(c.u_inlet))
    hu = 1e-6
    def J_of(v):
        u = u0.copy()
        u[:c.N][(flags & INLET) != 0] = v
        return orc.Oracle(flags, c.nx, c.ny, c.tau, rho0=rho0, u0=u, threads=4).objective(
            steps, col, t0, t1)
    fd = (J_of(uin + hu) - J_of(uin - hu)) / (2 * hu)
    assert abs(duin - fd) <= 1e-6 * abs(fd), (duin, fd)
