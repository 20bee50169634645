"""Shared problem descriptions and seeded inputs for the parity tests.

Configs are plain dicts whose keys are the frsdg_desc / fro_desc fields, so the
same dict builds the HIP context and the oracle.
"""
import numpy as np

SEED = 20261203

PER = [5, 5, 5, 5]
OUT = [1, 1, 1, 1]


def vortex_cfg(nx=8, ny=6, p=4):
    # config "2D isentropic vortex": [0,10]^2 periodic (the reference's box, vortex at its centre)
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=10.0, ly=10.0, riemann=1, bc_type=PER)


def explosion_cfg(nx=6, ny=5, p=3, gas=0, riemann=0, filt=0):
    c = dict(dims=2, nx=nx, ny=ny, p=p, lx=2.0, ly=2.0, riemann=riemann, bc_type=OUT,
             apply_filter=filt, filt_cutoff=0, filt_strength=16, filt_alpha=36.0)
    if gas:
        c.update(gas_kind=1, vdw_a=0.1, vdw_b=0.02)
    return c


def bump_cfg(nx=8, ny=4, p=5):
    g = 1.4
    rho, p_, u = 1.0, 1.0, 0.5 * np.sqrt(g)
    inflow = [rho, rho * u, 0.0, p_ / (g - 1) + 0.5 * rho * u * u]
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=3.0, ly=0.8, riemann=1, geometry=3,
                geom_param=[0.0625, -1.5, 0.0, 1.0],
                bc_type=[2, 1, 2, 6],  # L Dirichlet inflow, R outflow, T Dirichlet, B slip wall
                bc_state=[inflow, [0.0] * 4, inflow, [0.0] * 4])


def warped_cfg(nx=5, ny=4, p=7, filt=1):
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=2.0, ly=2.0, riemann=1, geometry=4,
                geom_param=[0.05, -1.0, -1.0, 0.0], bc_type=OUT, gas_kind=1, vdw_a=0.1, vdw_b=0.02,
                apply_filter=filt, filt_cutoff=0, filt_strength=16, filt_alpha=36.0)


def wall_cfg(nx=4, ny=3, p=2, riemann=0):
    return dict(dims=2, nx=nx, ny=ny, p=p, lx=1.0, ly=1.0, riemann=riemann, bc_type=[0, 0, 0, 0])


def sod_cfg(nx=200, p=3, riemann=0):
    return dict(dims=1, nx=nx, ny=1, p=p, lx=1.0, riemann=riemann, bc_type=[1, 1, 1, 1])


def rhoE_of(cfg, rho, vel2, p):
    g = cfg.get("gamma", 1.4)
    if cfg.get("gas_kind", 0) == 0:
        return p / (g - 1.0) + 0.5 * rho * vel2
    a, b = cfg.get("vdw_a", 0.0), cfg.get("vdw_b", 0.0)
    return ((p + a * rho * rho) * (1.0 - b * rho)) / (g - 1.0) - a * rho * rho + 0.5 * rho * vel2


def random_state(cfg, seed=SEED):
    """Seeded physical state: rho in U[0.5,1.5], velocities in U[-0.5,0.5], p in U[0.5,1.5]."""
    rng = np.random.default_rng(seed)
    n = cfg["p"] + 1
    if cfg.get("dims", 2) == 2:
        E = cfg["nx"] * cfg["ny"]
        rho = rng.uniform(0.5, 1.5, (E, n * n))
        u = rng.uniform(-0.5, 0.5, (E, n * n))
        v = rng.uniform(-0.5, 0.5, (E, n * n))
        p = rng.uniform(0.5, 1.5, (E, n * n))
        st = np.stack([rho, rho * u, rho * v, rhoE_of(cfg, rho, u * u + v * v, p)], axis=-1)
        return st.reshape(-1)
    E = cfg["nx"]
    rho = rng.uniform(0.5, 1.5, (E, n))
    u = rng.uniform(-0.5, 0.5, (E, n))
    p = rng.uniform(0.5, 1.5, (E, n))
    st = np.stack([rho, rho * u, rhoE_of(cfg, rho, u * u, p)], axis=1)  # [E][eq][node]
    return st.reshape(-1)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
