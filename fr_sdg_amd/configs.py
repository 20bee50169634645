"""Problem descriptions of the benchmark configurations (plain dicts whose keys
are the frsdg_desc fields), and their initial-condition parameters.

* sod1d        — 1D Sod tube, [0,1], 200 elements, p=3 (factory.hpp:175-229)
* vortex       — 2D isentropic vortex, periodic [0,10]^2, 128^2, p=4, HLLC
                 (factory.hpp:311-395; the reference's box with the vortex at its centre)
* bump         — channel [-1.5,1.5]x[0,0.8] with a sin^2 bump on the lower wall,
                 256x64, p=5, HLLC, M=0.5 inflow (tests/bump.cpp pattern)
* batch64      — 64 radial shock tubes, N in {32,48,...,192}, p in {2..6}, half
                 ideal / half van der Waals, Rusanov, filter s=16 alpha=36 n_c=0
                 (factory.hpp:400-465 CircularExplosion on [0,2]^2)
* warped       — [-1,1]^2 sine-warped, 1024^2, p=7, van der Waals, HLLC + filter (order 8)
"""
import math

SEED = 20261203
GAMMA = 1.4

VORTEX_PRM = [1.0, 1.0, 5.0, 1.0, 10.0]            # u0, v0, beta, r_v, L
SOD_PRM = [1.0, 0.0, 1.0, 0.125, 0.0, 0.1]         # rhoL, uL, pL, rhoR, uR, pR


class MT19937_64:
    """64-bit Mersenne Twister (Matsumoto & Nishimura 2000), the generator of
    std::mt19937_64: same seeding (init_genrand64) and output sequence."""

    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        self.idx = 312

    def _twist(self):
        um, lm = 0xFFFFFFFF80000000, 0x7FFFFFFF
        for i in range(312):
            x = (self.mt[i] & um) | (self.mt[(i + 1) % 312] & lm)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            self.mt[i] = self.mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & 0xFFFFFFFFFFFFFFFF


def sod1d(riemann=0, apply_filter=0):
    return dict(dims=1, nx=200, ny=1, p=3, lx=1.0, riemann=riemann, bc_type=[1, 1, 1, 1],
                apply_filter=apply_filter, filt_cutoff=0, filt_strength=8, filt_alpha=36.0)


def vortex(nx=128, ny=128, p=4):
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=10.0, ly=10.0, riemann=1, bc_type=[5, 5, 5, 5])


def bump(nx=256, ny=64, p=5, mach=0.5, tangent_bc_mode=0):
    rho, pr = 1.0, 1.0
    u = mach * math.sqrt(GAMMA * pr / rho)
    inflow = [rho, rho * u, 0.0, pr / (GAMMA - 1) + 0.5 * rho * u * u]
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=3.0, ly=0.8, riemann=1, geometry=3,
                geom_param=[0.0625, -1.5, 0.0, 1.0],
                bc_type=[2, 1, 2, 6],  # inflow Dirichlet, outflow, top free-stream Dirichlet, slip wall
                bc_state=[inflow, [0.0] * 4, inflow, [0.0] * 4], tangent_bc_mode=tangent_bc_mode)


def bump_freestream(mach=0.5):
    return [1.0, mach * math.sqrt(GAMMA), 0.0, 1.0]  # rho, u, v, p


def batch64_draws(count=64, seed=SEED):
    """(N, p) per instance: two raw std::mt19937_64 draws per instance, in
    instance order: N = 32 + 16 * (x0 mod 11), p = 2 + (x1 mod 5)."""
    g = MT19937_64(seed)
    out = []
    for _ in range(count):
        N = 32 + 16 * (g() % 11)
        p = 2 + (g() % 5)
        out.append((N, p))
    return out


def explosion(N, p, vdw, riemann=0, filt=1):
    c = dict(dims=2, nx=N, ny=N, p=p, lx=2.0, ly=2.0, riemann=riemann, bc_type=[1, 1, 1, 1],
             apply_filter=filt, filt_cutoff=0, filt_strength=16, filt_alpha=36.0)
    if vdw:
        c.update(gas_kind=1, vdw_a=0.1, vdw_b=0.02)
    return c


EXPLOSION_PRM = [1.0, 1.0, 0.125, 0.1, 0.4, 1.0, 1.0]  # rhoL, pL, rhoR, pR, radius, xc, yc
# The sharp jump at element-interior GLL nodes loses positivity under the
# post-step exponential filter at once (the filter of the IC alone gives
# rho < 0); the sweeps run the same radial shock with the jump spread over a
# tanh layer of width 0.1.
EXPLOSION_SMOOTH_PRM = EXPLOSION_PRM + [0.1]


def explicit_cfl(p):
    """CFL (on the element size) used for explicit runs: 0.4 / (2p + 1)
    (0.044 at p = 4, where 0.1 is unstable)."""
    return 0.4 / (2 * p + 1)


def batch64(count=64, scale=1.0):
    """Config of the batched sweep; `scale` < 1 shrinks every mesh (tests)."""
    cfgs = []
    for k, (N, p) in enumerate(batch64_draws(count)):
        Ns = max(2, int(round(N * scale)))
        cfgs.append(explosion(Ns, p, vdw=(k >= count // 2)))
    return cfgs


def warped(n=1024, p=7, filt=1):
    """Filter order 8 (alpha 36): with order 16 the p=7 run on this map loses
    positivity within 200 steps at CFL 0.4/(2p+1), also from a smoother
    initial jump; order 8 keeps the 2000-step run positive."""
    return dict(dims=2, nx=n, ny=n, p=p, lx=2.0, ly=2.0, riemann=1, geometry=4,
                geom_param=[0.05, -1.0, -1.0, 0.0], bc_type=[1, 1, 1, 1], gas_kind=1, vdw_a=0.1,
                vdw_b=0.02, apply_filter=filt, filt_cutoff=0, filt_strength=8, filt_alpha=36.0)


WARPED_PRM = [1.0, 1.0, 0.125, 0.1, 0.4, 0.0, 0.0, 0.1]


def sound_speed_max(u, cfg):
    """max(|v| + c) over a state tensor (torch), for CFL time steps."""
    import torch
    g = cfg.get("gamma", GAMMA)
    if cfg.get("dims", 2) == 2:
        st = u.view(-1, 4)
        rho, mx, my, E = st[:, 0], st[:, 1], st[:, 2], st[:, 3]
        vel2 = (mx * mx + my * my) / (rho * rho)
    else:
        n = cfg["p"] + 1
        st = u.view(-1, 3, n)
        rho, mx, E = st[:, 0], st[:, 1], st[:, 2]
        vel2 = mx * mx / (rho * rho)
    e = E / rho - 0.5 * vel2
    if cfg.get("gas_kind", 0) == 0:
        c2 = g * (g - 1) * e
    else:
        a, b = cfg["vdw_a"], cfg["vdw_b"]
        cov = 1 - b * rho
        c2 = (g * (g - 1) * (e + a * rho)) / cov ** 2 + ((g - 1) * a * rho * (2 - b * rho)) / cov ** 2 - 2 * a * rho
    return float(torch.max(torch.sqrt(vel2) + torch.sqrt(c2)))
