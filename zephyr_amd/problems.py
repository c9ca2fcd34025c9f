"""Synthetic initial data for the benchmark workloads (seeded, host-side setup).

Draw order and quadrant numbering follow the reference's Riemann2D test
(phys/tests/test_2D.cpp:72-106); the stream is SplitMix64 with
u = (x >> 11) * 2^-53.
"""
import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.x = seed & _M64

    def next_u64(self):
        self.x = (self.x + 0x9E3779B97F4A7C15) & _M64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def energy_rP(rho, p, gamma, P0=0.0, e0=0.0):
    """StiffenedGas::energy_rP (phys/matter/eos/stiffened_gas.cpp:121-122)."""
    return e0 + (p + gamma * P0) / ((gamma - 1.0) * rho)


def quadrant_states(seed):
    """Per-quadrant (rho, u, v, p) for quadrants 1..4: rho~U[0.1,2], u,v~U[-1,1], p~U[0.1,2]."""
    g = SplitMix64(seed)
    out = []
    for _ in range(4):
        r = 0.1 + 1.9 * g.uniform()
        u = -1.0 + 2.0 * g.uniform()
        v = -1.0 + 2.0 * g.uniform()
        p = 0.1 + 1.9 * g.uniform()
        out.append((r, u, v, p))
    return out


def four_quadrant(cx, cy, seed, gamma=1.4):
    """Four-quadrant Riemann data at cell centres -> (rho, u, v, p, e)."""
    q = quadrant_states(seed)
    cx = np.asarray(cx)
    cy = np.asarray(cy)
    idx = np.full(cx.shape, 3)
    idx[(cx >= 0.5) & (cy >= 0.5)] = 0
    idx[(cx <= 0.5) & (cy >= 0.5) & (idx == 3)] = 1
    idx[(cx <= 0.5) & (cy <= 0.5) & (idx == 3)] = 2
    tab = np.array(q)
    rho, u, v, p = (tab[idx, k].copy() for k in range(4))
    e = energy_rP(rho, p, gamma)
    return rho, u, v, p, e


def gaussian_random_field(nx, ny, seed, corr_len=0.05, Lx=1.0, Ly=1.0):
    """Unit-variance Gaussian random field on the cell grid (index i*ny + j):
    white noise filtered by a Gaussian kernel of correlation length corr_len
    in Fourier space, then normalised to mean 0, std 1."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nx, ny))
    kx = np.fft.fftfreq(nx, d=Lx / nx) * 2.0 * np.pi
    ky = np.fft.fftfreq(ny, d=Ly / ny) * 2.0 * np.pi
    K2 = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-0.25 * K2 * corr_len ** 2)))
    f = (f - f.mean()) / f.std()
    return f.reshape(-1)


def smooth_density_field(nx, ny, seed, gamma=1.4):
    """Adjoint / uniform-mesh workload: rho = 1 + 0.1 g, p = 1, u = v = 0."""
    rho = 1.0 + 0.1 * gaussian_random_field(nx, ny, seed)
    p = np.ones_like(rho)
    z = np.zeros_like(rho)
    return rho, z, z.copy(), p, energy_rP(rho, p, gamma)


def smooth_flow_field(nx, ny, seed, gamma=1.4, corr_len=0.05):
    """Adjoint workload: rho = 1 + 0.1 g0, p = 1 + 0.1 g1, u = 0.1 g2,
    v = 0.1 g3, with g_k the seeded unit Gaussian random fields of seeds
    seed + k.  Pressure and velocity vary, so the final pressure depends on
    the initial density through the acoustics (with p and u constant it
    would not: a resting contact keeps p = const)."""
    g = [gaussian_random_field(nx, ny, seed + k, corr_len=corr_len) for k in range(4)]
    rho = 1.0 + 0.1 * g[0]
    p = 1.0 + 0.1 * g[1]
    return rho, 0.1 * g[2], 0.1 * g[3], p, energy_rP(rho, p, gamma)


def circle_blast_params(seed):
    """Seeded parameters of the AMR circle workload: centre ~U[0.3,0.7]^2
    (x then y), perturbation wavenumber k ~ U{4..16}."""
    g = SplitMix64(seed)
    x0 = 0.3 + 0.4 * g.uniform()
    y0 = 0.3 + 0.4 * g.uniform()
    k = 4 + int(13.0 * g.uniform())
    return x0, y0, k


def circle_blast(cx, cy, seed, gamma=1.4, radius=0.2):
    """Circular density/pressure jump 10:1 (rho = p = 1 inside radius 0.2,
    0.1 outside) with rho *= 1 + 0.01 sin(2 pi k (x + y)); u = v = 0."""
    x0, y0, k = circle_blast_params(seed)
    cx = np.asarray(cx, dtype=np.float64)
    cy = np.asarray(cy, dtype=np.float64)
    inside = (cx - x0) ** 2 + (cy - y0) ** 2 < radius * radius
    rho = np.where(inside, 1.0, 0.1) * (1.0 + 0.01 * np.sin(2.0 * np.pi * k * (cx + cy)))
    p = np.where(inside, 1.0, 0.1)
    z = np.zeros_like(rho)
    return rho, z, z.copy(), p, energy_rP(rho, p, gamma)


# ---- two-material shock-bubble (stiffened-gas water with an air bubble) ----
WATER = (4.4, 6.0e8, 0.0, 4160.0, 230.7, 1000.0)   # StiffenedGas("Water") (stiffened_gas.cpp:33-40)
AIR = (1.4, 0.0, 0.0, 718.0, 0.0, 1.25)            # StiffenedGas("Air") (stiffened_gas.cpp:22-25)


def mm_bubble_shock_params(seed):
    """Seeded shock Mach number M ~ U[1.5, 3]."""
    g = SplitMix64(seed)
    return 1.5 + 1.5 * g.uniform()


def mm_bubble_shock(cx, cy, h, seed, radius=0.15, centre=(0.5, 0.5), x_shock=0.1, p1=1.0e5, sub=4):
    """Water at rest (rho 1000, P = p1) with an air bubble (rho 1.25) of the
    given radius; a planar shock of seeded Mach number at x = x_shock moves
    into the water along +x, the post-shock water state from the
    stiffened-gas Rankine-Hugoniot relations (P + P0 behaves as an ideal-gas
    pressure).  Cells cut by the bubble get the volume fraction of a sub x sub
    sampling, mass fractions beta_i = alpha_i rho_i / rho.  Returns rho, u, v,
    P and beta (2, n): material 0 water, 1 air."""
    M = mm_bubble_shock_params(seed)
    gw, P0w = WATER[0], WATER[1]
    rw, ra = WATER[5], AIR[5]
    pt1 = p1 + P0w
    c1 = np.sqrt(gw * pt1 / rw)
    D = M * c1
    r2 = rw * (gw + 1.0) * M * M / ((gw - 1.0) * M * M + 2.0)
    p2 = pt1 * (2.0 * gw * M * M - (gw - 1.0)) / (gw + 1.0) - P0w
    u2 = D * (1.0 - rw / r2)
    n = len(cx)
    offs = (np.arange(sub) + 0.5) / sub - 0.5
    alpha = np.zeros(n)
    for ox in offs:
        for oy in offs:
            alpha += ((cx + ox * h - centre[0]) ** 2 + (cy + oy * h - centre[1]) ** 2 < radius * radius)
    alpha /= sub * sub
    shocked = cx < x_shock
    rho_w = np.where(shocked, r2, rw)
    rho = alpha * ra + (1.0 - alpha) * rho_w
    b_air = alpha * ra / rho
    u = np.where(shocked, u2, 0.0)
    P = np.where(shocked, p2, p1)
    return rho, u, np.zeros(n), P, np.stack([1.0 - b_air, b_air])
