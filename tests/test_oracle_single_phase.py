"""Pins the oracle's single-phase step (stream-collide, slot algebra, walls,
inlet/outlet, Guo forcing, TRT, Smagorinsky) with checks that do not reuse its
code: an independent two-lattice numpy formulation, closed forms and
conservation laws."""
import numpy as np
import pytest

from lbm_numpy import CX, CY, OPP, W, TwoLattice, esoteric_view

FLUID, WALL, INLET, OUTLET = 0x01, 0x08, 0x20, 0x40


def channel(orc, nx, ny, left=INLET, right=OUTLET, up=WALL, down=WALL, obstacles=()):
    return orc.build_channel_flags(nx, ny, 1.0, left, right, up, down, obstacles=obstacles)


def inlet_u(flags, N, uin):
    u0 = np.zeros(2 * N)
    u0[:N][(flags & INLET) != 0] = uin
    return u0


@pytest.mark.parametrize("use_force", [False, True])
def test_matches_two_lattice_formulation(orc, use_force):
    nx, ny = 24, 10
    flags = channel(orc, nx, ny, obstacles=[(11.5, 6.5, 5.0, 4.0)])
    N = nx * ny
    rng = np.random.default_rng(3)
    rho0 = 1.0 + rng.uniform(-1e-3, 1e-3, N)
    u0 = inlet_u(flags, N, 0.05)
    inv_tau = float(np.float32(1) / np.float32(0.56))
    kw = dict(use_force=use_force, fx=1e-5 if use_force else 0.0, fy=-2e-6 if use_force else 0.0)
    o = orc.Oracle(flags, nx, ny, 0.56, rho0=rho0, u0=u0, threads=1, **kw)
    ref = TwoLattice(flags, nx, ny, inv_tau, rho0=rho0, u0=u0, **kw)
    active = (flags & WALL) == 0
    for t in range(40):
        fin_ref = ref.incoming()
        fin_orc = esoteric_view(o.get()["f"], nx, ny, t)
        np.testing.assert_allclose(fin_orc[:, active], fin_ref[:, active], rtol=0, atol=1e-12)
        ref.step()
        o.step()
        g = o.get()
        np.testing.assert_allclose(g["rho"], ref.rho, rtol=0, atol=1e-12)
        np.testing.assert_allclose(g["u"], ref.u, rtol=0, atol=1e-12)


def test_esoteric_slots_are_a_permutation(orc):
    """Every slot is loaded exactly once per parity (in-place without conflicts)."""
    nx, ny = 7, 5
    N = nx * ny
    f = np.arange(9 * N, dtype=np.float64)
    for t in (0, 1):
        v = esoteric_view(f, nx, ny, t)
        assert sorted(v.ravel().astype(np.int64).tolist()) == list(range(9 * N))


def test_uniform_rest_is_fixed_point(orc):
    nx, ny = 16, 8
    flags = np.full(nx * ny, FLUID, np.uint8)
    o = orc.Oracle(flags, nx, ny, 0.7, threads=1)
    f0 = o.get()["f"]
    o.step(5)
    np.testing.assert_allclose(o.get()["f"], f0, rtol=0, atol=1e-15)


def test_equilibrium_moments(orc):
    """tau = 1: post-collision populations are feq(rho, u).  Their moments are
    rho, rho*u and rho*(u u + I/3)."""
    nx, ny = 12, 9
    N = nx * ny
    flags = np.full(N, FLUID, np.uint8)
    rng = np.random.default_rng(0)
    rho0 = 1 + rng.uniform(-0.05, 0.05, N)
    u0 = rng.uniform(-0.08, 0.08, 2 * N)
    o = orc.Oracle(flags, nx, ny, 1.0, rho0=rho0, u0=u0, threads=1)
    o.step(1)
    g = o.get()
    # populations arriving at step 1 came from x - c_k; gather them back per source cell
    fin = esoteric_view(g["f"], nx, ny, 1)
    y, x = np.divmod(np.arange(N), nx)
    post = np.empty((9, N))
    for k in range(9):
        dst = (x + CX[k]) % nx + ((y + CY[k]) % ny) * nx
        post[k] = fin[k, dst]
    rho, ux, uy = g["rho"], g["u"][:N], g["u"][N:]
    np.testing.assert_allclose(post.sum(0), rho, rtol=1e-14)
    np.testing.assert_allclose((CX[:, None] * post).sum(0), rho * ux, atol=1e-15)
    np.testing.assert_allclose((CY[:, None] * post).sum(0), rho * uy, atol=1e-15)
    np.testing.assert_allclose((CX[:, None] ** 2 * post).sum(0), rho * (ux * ux + 1 / 3), atol=1e-15)
    np.testing.assert_allclose((CY[:, None] ** 2 * post).sum(0), rho * (uy * uy + 1 / 3), atol=1e-15)
    np.testing.assert_allclose(((CX * CY)[:, None] * post).sum(0), rho * ux * uy, atol=1e-15)


def test_guo_force_moments(orc):
    """tau = 1 from rest: post = feq(rho, F/2) + F_i/2.  Sum F_i = 0 and
    sum c_i F_i = F * 3 * 0.33333334f (the reference's literal)."""
    nx, ny = 6, 6
    N = nx * ny
    flags = np.full(N, FLUID, np.uint8)
    fx, fy = 3e-4, -7e-4
    o = orc.Oracle(flags, nx, ny, 1.0, use_force=True, fx=fx, fy=fy, threads=1)
    o.step(1)
    g = o.get()
    fin = esoteric_view(g["f"], nx, ny, 1)
    jx = (CX[:, None] * fin).sum(0)
    jy = (CY[:, None] * fin).sum(0)
    third = float(np.float32(0.33333334))
    fx32, fy32 = float(np.float32(fx)), float(np.float32(fy))
    # sum F_i = 9 (1/3 - 0.33333334f) (u . F), not exactly zero with the literal; the
    # residual ~1e-14 comes from c_i . F being formed in float
    ux, uy = fx32 / 2, fy32 / 2
    np.testing.assert_allclose(fin.sum(0) - 1.0, 0.5 * 9 * (1 / 3 - third) * (ux * fx32 + uy * fy32),
                               rtol=0, atol=1e-14)
    np.testing.assert_allclose(jx, fx32 / 2 + fx32 * 3 * third / 2, rtol=1e-7)  # c.F in float
    np.testing.assert_allclose(jy, fy32 / 2 + fy32 * 3 * third / 2, rtol=1e-7)
    np.testing.assert_allclose(g["u"][:N], fx32 / 2, rtol=1e-12)


def test_full_way_bounce_back(orc):
    """A population sent into a wall returns reversed to its sender two steps
    later: f_opp,in(n, t+2) = f_i*(n, t)."""
    nx, ny = 8, 6
    flags = channel(orc, nx, ny, left=FLUID, right=FLUID)
    N = nx * ny

    def make():
        rng = np.random.default_rng(1)
        return orc.Oracle(flags, nx, ny, 0.8, rho0=1 + rng.uniform(-1e-2, 1e-2, N),
                          u0=rng.uniform(-0.02, 0.02, 2 * N), threads=1)

    x, y = 3, 1          # fluid cell whose lower neighbours are wall
    n = x + y * nx
    for t in (3, 4):
        for i in (4, 6, 7):  # directions pointing into the bottom wall
            o = make()
            o.step(t)
            o.step(1)        # step t: its post-collision populations are in the slots
            f_after = o.get()["f"].reshape(9, N)
            nb = (x + CX[i]) % nx + ((y + CY[i]) % ny) * nx
            # storef slot of f_i*: odd i -> (i+1 if t odd else i, n + c_i);
            # even i -> (i-1 if t odd else i, n)
            if i % 2 == 1:
                sent = f_after[i + 1 if t % 2 else i, nb]
            else:
                sent = f_after[i - 1 if t % 2 else i, n]
            o.step(1)        # step t + 1
            fin = esoteric_view(o.get()["f"], nx, ny, t + 2)
            assert fin[OPP[i], n] == sent


def test_mass_conserved_periodic(orc):
    nx, ny = 20, 14
    N = nx * ny
    flags = np.full(N, FLUID, np.uint8)
    flags[(np.arange(N) % nx == 9) & (np.arange(N) // nx > 4)] = WALL
    rng = np.random.default_rng(2)
    o = orc.Oracle(flags, nx, ny, 0.6, rho0=1 + rng.uniform(-1e-2, 1e-2, N),
                   u0=rng.uniform(-0.05, 0.05, 2 * N), use_force=True, fx=1e-5, threads=1)
    m0 = o.get()["f"].sum()
    o.step(300)
    assert abs(o.get()["f"].sum() - m0) < 1e-11 * m0


def test_uniform_inflow_stays_uniform(orc):
    """Inlet and outlet at the same velocity on a channel periodic in y keep a
    uniform flow uniform (inlet/outlet rules and previous-step semantics)."""
    nx, ny = 16, 5
    N = nx * ny
    flags = channel(orc, nx, ny, up=FLUID, down=FLUID)
    u0 = np.zeros(2 * N)
    u0[:N] = 0.04
    o = orc.Oracle(flags, nx, ny, 0.7, u0=u0, threads=1)
    o.step(50)
    g = o.get()
    np.testing.assert_allclose(g["u"][:N], 0.04, rtol=1e-13)
    np.testing.assert_allclose(g["u"][N:], 0.0, atol=1e-15)
    np.testing.assert_allclose(g["rho"], 1.0, rtol=1e-14)


def test_poiseuille_profile(orc):
    """Body-force Poiseuille between full-way bounce-back walls: the wall sits
    half-way between the wall row and the first fluid row, u_max = F H^2/(8 rho nu)."""
    nx, ny = 4, 18
    N = nx * ny
    flags = channel(orc, nx, ny, left=FLUID, right=FLUID)
    tau = 0.8
    nu = (tau - 0.5) / 3
    F = 1e-6
    o = orc.Oracle(flags, nx, ny, tau, use_force=True, fx=F, threads=1)
    o.step(20000)
    ux = o.get()["u"][:N].reshape(ny, nx)[:, 0]
    H = ny - 2
    y = np.arange(ny) - 0.5
    exact = float(np.float32(F)) / (2 * nu) * y * (H - y)
    fluid = slice(1, ny - 1)
    err = np.linalg.norm(ux[fluid] - exact[fluid]) / np.linalg.norm(exact[fluid])
    assert err < 5e-3, err
    assert abs(ux[fluid].max() - float(np.float32(F)) * H * H / (8 * nu)) / ux[fluid].max() < 1e-2


def test_trt_equals_srt_at_unit_tau(orc):
    """At tau = 1 with magic 1/4 both relaxation rates are 1, so TRT and SRT
    produce the same populations (lbmKernel2D.cu:478)."""
    nx, ny = 10, 8
    N = nx * ny
    flags = channel(orc, nx, ny)
    rng = np.random.default_rng(5)
    kw = dict(rho0=1 + rng.uniform(-1e-3, 1e-3, N), u0=inlet_u(flags, N, 0.03), threads=1)
    a = orc.Oracle(flags, nx, ny, 1.0, collision=0, **kw)
    b = orc.Oracle(flags, nx, ny, 1.0, collision=1, **kw)
    a.step(30)
    b.step(30)
    # the reference's TRT temporaries are float even in double mode (lbmKernel2D.cu:495-497)
    np.testing.assert_allclose(a.get()["f"], b.get()["f"], rtol=0, atol=1e-8)


def test_trt_with_force_streams_only(orc):
    """The reference's TRT branch has no forced variant: populations are only
    streamed (lbmKernel2D.cu:485-488)."""
    nx, ny = 8, 6
    N = nx * ny
    flags = np.full(N, FLUID, np.uint8)
    rng = np.random.default_rng(6)
    o = orc.Oracle(flags, nx, ny, 0.7, collision=1, use_force=True, fx=1e-4, threads=1,
                   rho0=1 + rng.uniform(-1e-2, 1e-2, N))
    f0 = o.get()["f"]
    o.step(1)
    pre = esoteric_view(f0, nx, ny, 0)
    y, x = np.divmod(np.arange(N), nx)
    post = esoteric_view(o.get()["f"], nx, ny, 1)
    for k in range(9):
        dst = (x + CX[k]) % nx + ((y + CY[k]) % ny) * nx
        np.testing.assert_array_equal(post[k, dst], pre[k])


def test_trt_guo_equals_srt_guo_at_unit_tau(orc):
    """The defined forced TRT branch (collision 2): at tau = 1 both rates are 1,
    so the symmetric/antisymmetric split of the Guo terms recombines to SRT's
    (1 - w/2) F_i.  Without a force collision 2 is plain TRT."""
    nx, ny = 10, 8
    N = nx * ny
    flags = channel(orc, nx, ny)
    rng = np.random.default_rng(9)
    kw = dict(rho0=1 + rng.uniform(-1e-3, 1e-3, N), u0=inlet_u(flags, N, 0.03), threads=1)
    a = orc.Oracle(flags, nx, ny, 1.0, collision=0, use_force=True, fx=2e-5, fy=-1e-5, **kw)
    b = orc.Oracle(flags, nx, ny, 1.0, collision=2, use_force=True, fx=2e-5, fy=-1e-5, **kw)
    a.step(30)
    b.step(30)
    np.testing.assert_allclose(a.get()["f"], b.get()["f"], rtol=0, atol=1e-8)
    c = orc.Oracle(flags, nx, ny, 0.7, collision=1, **kw)
    d = orc.Oracle(flags, nx, ny, 0.7, collision=2, **kw)
    c.step(30)
    d.step(30)
    np.testing.assert_array_equal(c.get()["f"], d.get()["f"])


def test_trt_guo_poiseuille_profile(orc):
    """Forced TRT (collision 2) reproduces the body-force Poiseuille parabola
    at tau = 0.8, where the as-written TRT would not collide at all."""
    nx, ny = 4, 18
    N = nx * ny
    flags = channel(orc, nx, ny, left=FLUID, right=FLUID)
    tau = 0.8
    nu = (tau - 0.5) / 3
    F = 1e-6
    o = orc.Oracle(flags, nx, ny, tau, use_force=True, fx=F, collision=2, threads=1)
    o.step(20000)
    ux = o.get()["u"][:N].reshape(ny, nx)[:, 0]
    H = ny - 2
    y = np.arange(ny) - 0.5
    exact = float(np.float32(F)) / (2 * nu) * y * (H - y)
    fluid = slice(1, ny - 1)
    err = np.linalg.norm(ux[fluid] - exact[fluid]) / np.linalg.norm(exact[fluid])
    assert err < 5e-3, err


def test_smagorinsky_effective_tau(orc):
    """tau_eff = (tau + sqrt(tau^2 + 18 (C Delta)^2 sqrt(Q)/rho))/2 with
    18 (C Delta)^2 = 0.54037965; at rest Q = 0 and tau_eff = tau."""
    nx, ny = 6, 6
    N = nx * ny
    flags = np.full(N, FLUID, np.uint8)
    a = orc.Oracle(flags, nx, ny, 0.6, smagorinsky=True, threads=1)
    a.step(3)
    b = orc.Oracle(flags, nx, ny, 0.6, threads=1)
    b.step(3)
    np.testing.assert_array_equal(a.get()["f"], b.get()["f"])
    # one non-equilibrium cell: compare one step against the closed form
    rng = np.random.default_rng(9)
    rho0 = 1 + rng.uniform(-1e-2, 1e-2, N)
    u0 = rng.uniform(-0.05, 0.05, 2 * N)
    s = orc.Oracle(flags, nx, ny, 0.6, smagorinsky=True, rho0=rho0, u0=u0, threads=1)
    s.step(1)  # now populations are not at equilibrium
    f1 = s.get()["f"]
    fin = esoteric_view(f1, nx, ny, 1)
    s.step(1)
    fin2 = esoteric_view(s.get()["f"], nx, ny, 2)
    n = 14
    f = fin[:, n]
    rho = f.sum()
    ux = (CX * f).sum() / rho
    uy = (CY * f).sum() / rho
    cu = 3 * (CX * ux + CY * uy)
    fe = W * rho * (1 + cu + 0.5 * cu * cu - 1.5 * (ux * ux + uy * uy))
    fne = f - fe
    Q = (((CX * CX) * fne).sum()) ** 2 + (((CY * CY) * fne).sum()) ** 2 + 2 * (((CX * CY) * fne).sum()) ** 2
    tau = float(np.float32(1) / (np.float32(1) / np.float32(0.6)))
    teff = 0.5 * (tau + np.sqrt(tau * tau + 0.54037965 * np.sqrt(Q) / rho))
    w = 1 / teff
    post = (1 - w) * f + w * fe
    y, x = divmod(n, nx)
    for k in range(9):
        dst = (x + CX[k]) % nx + ((y + CY[k]) % ny) * nx
        assert abs(fin2[k, dst] - post[k]) < 2e-8
