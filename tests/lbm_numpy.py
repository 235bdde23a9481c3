"""Independent textbook formulation of the single-phase step, in numpy (FP64).

Used only to pin the oracle.  It keeps two lattices (post-collision populations
of the last two steps, P[t-1] and P[t-2]) and streams by gathering from
x - c_k, instead of the in-place esoteric slot algebra.  Walls never compute; a
population sent into a wall comes back reversed two steps later (full-way
bounce-back), and at t = 0 every cell, walls included, emits its initial
equilibrium.  Inlet and outlet rules follow src/cuda/lbmKernel2D.cu:334-346
with the density/velocity fields of the previous step.
"""
import numpy as np

CX = np.array([0, 1, -1, 0, 0, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 1, -1, 1, -1, -1, 1])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7])
W = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])
F13 = float(np.float32(0.33333334))
WALL, INLET, OUTLET = 0x08, 0x20, 0x40


def feq(rho, ux, uy):
    cu = 3.0 * (CX[:, None] * ux[None] + CY[:, None] * uy[None])
    usq = 1.5 * (ux * ux + uy * uy)
    return W[:, None] * rho[None] * (1.0 + cu + 0.5 * cu * cu - usq[None])


def guo(ux, uy, fx, fy):
    cF = CX * fx + CY * fy
    uF = -F13 * (ux * fx + uy * fy)
    return 9.0 * W[:, None] * (cF[:, None] * (CX[:, None] * ux[None] + CY[:, None] * uy[None] + F13)
                              + uF[None])


class TwoLattice:
    def __init__(self, flags, nx, ny, inv_tau, fx=0.0, fy=0.0, use_force=False, rho0=None,
                 u0=None):
        self.nx, self.ny, self.N = nx, ny, nx * ny
        self.flags = flags.astype(np.uint8)
        self.w = float(inv_tau)
        self.fx, self.fy = (float(np.float32(fx)), float(np.float32(fy))) if use_force else (0.0, 0.0)
        self.use_force = use_force
        N = self.N
        self.rho = np.ones(N) if rho0 is None else rho0.astype(np.float64).copy()
        self.u = np.zeros(2 * N) if u0 is None else u0.astype(np.float64).copy()
        p0 = feq(self.rho, self.u[:N], self.u[N:])
        self.P1 = p0.copy()   # P[t-1]
        self.P2 = p0.copy()   # P[t-2] (only read for t >= 1)
        self.t = 0
        y, x = np.divmod(np.arange(N), nx)
        self.src = np.empty((9, N), dtype=np.int64)   # x - c_k, periodic
        for k in range(9):
            self.src[k] = (x - CX[k]) % nx + ((y - CY[k]) % ny) * nx
        self.wall = (self.flags & WALL) != 0

    def incoming(self):
        """Populations each cell loads at step t."""
        fin = np.empty((9, self.N))
        for k in range(9):
            s = self.src[k]
            v = self.P1[k, s]
            if self.t >= 1:
                bounced = self.wall[s]
                v = np.where(bounced, self.P2[OPP[k]], v)
            fin[k] = v
        return fin

    def step(self):
        N = self.N
        fin = self.incoming()
        fl = self.flags
        active = ~self.wall
        rho = fin.sum(axis=0)
        jx = (CX[:, None] * fin).sum(axis=0)
        jy = (CY[:, None] * fin).sum(axis=0)
        ux = jx / rho
        uy = jy / rho
        inl = (fl & INLET) != 0
        out = ((fl & OUTLET) != 0) & ~inl
        idx = np.arange(N)
        rho_prev, u_prev = self.rho.copy(), self.u.copy()
        rho = np.where(inl, rho_prev[np.minimum(idx + 1, N - 1)], rho)
        ux = np.where(inl, u_prev[:N], ux)
        uy = np.where(inl, u_prev[N:], uy)
        rho = np.where(out, rho_prev, rho)
        ux = np.where(out, u_prev[np.maximum(idx - 1, 0)], ux)
        uy = np.where(out, u_prev[N + np.maximum(idx - 1, 0)], uy)
        Fi = np.zeros((9, N))
        if self.use_force:
            ux = ux + 0.5 / rho * self.fx
            uy = uy + 0.5 / rho * self.fy
            Fi = guo(ux, uy, self.fx, self.fy)
        io = inl | out
        new_rho = np.where(active, rho, rho_prev)
        new_ux = np.where(active & ~inl, ux, u_prev[:N])
        new_uy = np.where(active & ~inl, uy, u_prev[N:])
        fe = feq(rho, ux, uy)
        w = self.w
        post = (1.0 - w) * fin + w * fe + (1.0 - 0.5 * w) * Fi
        post = np.where(io[None], fe, post)
        post = np.where(active[None], post, self.P1)   # walls keep their populations
        self.P2, self.P1 = self.P1, post
        self.rho = new_rho
        self.u = np.concatenate([new_ux, new_uy])
        self.t += 1
        return fin


def esoteric_view(f, nx, ny, t):
    """Populations the in-place esoteric array hands to each cell at step t
    (the slot table of src/cuda/cellCalculations.cuh:185-196)."""
    N = nx * ny
    f = f.reshape(9, N)
    y, x = np.divmod(np.arange(N), nx)
    out = np.empty((9, N))
    out[0] = f[0]
    odd = t % 2 == 1
    for i in (1, 3, 5, 7):
        nb = (x + CX[i]) % nx + ((y + CY[i]) % ny) * nx
        out[i] = f[i] if odd else f[i + 1]
        out[i + 1] = f[i + 1, nb] if odd else f[i, nb]
    return out
