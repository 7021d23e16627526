"""Workload definitions: parameter sets and initial conditions.

Initial conditions are generated on the host with a committed PRNG
(splitmix64, seed 1234 by default) and uploaded, so the GPU path and the CPU
oracle start from identical bits. The reference instead draws with cuRAND on
the GPU and std::mt19937_64 on the CPU (example_Reactive_ModelB/kernel.cu:1032-1042,
example_Reactive_ModelB/CPU_version/base_CPU.cpp:541-551), which differ.

Named workloads:
  dendrite_k93   2D Kobayashi dendrite, 512^2, dx=0.03, dt=1e-4, tau=3e-4,
                 eps_bar=0.01, delta=0.02, j=4, theta0=0, alpha=0.9, gamma=10,
                 K=1.6, T_eq=1; phi=1 in a disk of radius 5 cells, T=0; zero-flux
                 boundaries; 16x16 activation tiles.
  ch_reactive    2D Cahn-Hilliard + linear reaction, 1024^2 periodic, M=1,
                 kappa=1, dx=1, dt=0.01, Gamma=1e-3, cbar=0, c = cbar + U[-0.05,0.05];
                 32x32 tiles; 9-point isotropic Laplacian.
"""
import numpy as np

from . import abi

MASK64 = (1 << 64) - 1


def splitmix64(seed, n):
    """n uint64 values of the splitmix64 sequence started at `seed` (vectorised)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n):
    """Doubles in [0, 1): the top 53 bits of splitmix64, times 2^-53."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def base_params(model, nx, ny, nz=1, tile=16, mode=abi.MODE_DBA):
    p = abi.Params()
    p.model = model
    p.mode = mode
    p.nx, p.ny, p.nz = nx, ny, nz
    p.tile_x = p.tile_y = tile
    p.tile_z = 1 if nz == 1 else tile
    p.history_cap = 0
    return p


def dendrite_k93_params(n=512, tile=16, mode=abi.MODE_DBA, delta=0.02, aniso_j=4, latent_k=1.6,
                        theta0=0.0, threads=0, history_cap=0, bc=abi.BC_NOFLUX, ny=None):
    p = base_params(abi.MODEL_DENDRITE_K93, n, ny or n, 1, tile, mode)
    p.bc_x = p.bc_y = bc
    p.dx = p.dy = p.dz = 0.03
    p.dt = 1e-4
    p.tau = 3e-4
    p.eps_bar = 0.01
    p.delta = delta
    p.aniso_j = aniso_j
    p.theta0 = theta0
    p.alpha = 0.9
    p.gamma = 10.0
    p.latent_k = latent_k
    p.t_eq = 1.0
    p.thresh_a = 0.01   # Metric_p (example_Dendritic_growth/Input.txt:13)
    p.thresh_b = 1e-5   # Metric_T (example_Dendritic_growth/Input.txt:15)
    p.threads = threads
    p.history_cap = history_cap
    return p


def dendrite_k93_ic(nx, ny, seeds=None, radius=5.0):
    """phi = 1 inside disks of `radius` cells (default: one at the centre), T = 0."""
    if seeds is None:
        seeds = [(nx // 2, ny // 2)]
    y, x = np.mgrid[0:ny, 0:nx]
    phi = np.zeros((ny, nx))
    for (cx, cy) in seeds:
        phi[(x - cx) ** 2 + (y - cy) ** 2 <= radius * radius] = 1.0
    return phi, np.zeros((ny, nx))


def ch_params(n=1024, tile=32, mode=abi.MODE_DBA, stencil=9, react_gamma=1e-3, cbar=0.0, eps=1e-6,
              threads=0, history_cap=0, ny=None):
    p = base_params(abi.MODEL_CH, n, ny or n, 1, tile, mode)
    p.bc_x = p.bc_y = abi.BC_PERIODIC
    p.dx = p.dy = p.dz = 1.0
    p.dt = 0.01
    p.mobility = 1.0
    p.kappa = 1.0
    p.react_gamma = react_gamma
    p.cbar = cbar
    p.stencil = stencil
    p.thresh_a = eps    # Metric_eps (example_Reactive_ModelB/Input.txt:25)
    p.threads = threads
    p.history_cap = history_cap
    return p


def ch_ic(nx, ny, cbar=0.0, amp=0.05, seed=1234):
    return cbar + amp * (2.0 * uniform01(seed, nx * ny) - 1.0).reshape(ny, nx)


def kh_params(n=2048, tile=16, mode=abi.MODE_DBA, gas_gamma=1.4, cfl=0.4, walls=True, metric=abi.KH_METRIC_DENSITY,
              eps=0.1, dt=2e-4, dt_mode=abi.DT_CFL, threads=0, history_cap=0, ny=None):
    """Compressible Euler on the unit square (dx = dy = 1/n)."""
    p = base_params(abi.MODEL_KH, n, ny or n, 1, tile, mode)
    p.bc_x = abi.BC_PERIODIC
    p.bc_y = abi.BC_WALL if walls else abi.BC_PERIODIC
    p.dx = p.dy = p.dz = 1.0 / n
    p.dt = dt
    p.dt_mode = dt_mode
    p.gas_gamma = gas_gamma
    p.cfl = cfl
    p.metric_kind = metric
    p.thresh_a = eps
    p.threads = threads
    p.history_cap = history_cap
    return p


def kh_reference_params(n=512, tile=8, mode=abi.MODE_DBA, dt=2e-4, dx=0.002, eps=0.0, history_cap=0):
    """The reference KH model (example_Kelvin_Helmholtz/Input.txt): gamma = 5/3,
    periodic in x and y, fixed dt, 8x8 blocks, sum-of-8-gradients detector."""
    p = base_params(abi.MODEL_KH, n, n, 1, tile, mode)
    p.bc_x = p.bc_y = abi.BC_PERIODIC
    p.dx = p.dy = p.dz = dx
    p.dt = dt
    p.dt_mode = abi.DT_FIXED
    p.gas_gamma = 5.0 / 3.0
    p.metric_kind = abi.KH_METRIC_ALLGRAD
    p.thresh_a = eps
    p.history_cap = history_cap
    return p


def kh_conserved(rho, vx, vy, pr, gas_gamma, vol):
    """getConserved (example_Kelvin_Helmholtz/KH_DBA.cu:363-367), same operation order."""
    mass = rho * vol
    momx = rho * vx * vol
    momy = rho * vy * vol
    energy = (pr / (gas_gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy)) * vol
    return [mass, momx, momy, energy]


def kh_primitive(U, gas_gamma, vol):
    """getPrimitive (example_Kelvin_Helmholtz/KH_DBA.cu:401-405)."""
    mass, momx, momy, energy = U
    rho = mass / vol
    vx = momx / rho / vol
    vy = momy / rho / vol
    pr = (energy / vol - 0.5 * rho * (vx * vx + vy * vy)) * (gas_gamma - 1.0)
    return rho, vx, vy, pr


def kh_ic(n, gas_gamma=1.4, ny=None):
    """Sharp shear layers: rho=2, vx=+0.5 for |y-1/2|<1/4, else rho=1, vx=-0.5; p=2.5;
    vy = 0.01 sin(4 pi x) (exp(-(y-1/4)^2/0.005) + exp(-(y-3/4)^2/0.005)), cell centres
    (the sharp form of example_Kelvin_Helmholtz/CPU_version/KH.py:222-225, the two
    Gaussians summed). Returns conserved fields."""
    ny = ny or n
    x = (np.arange(n) + 0.5) / n
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    inner = np.abs(Y - 0.5) < 0.25
    rho = 1.0 + inner
    vx = -0.5 + inner
    vy = 0.01 * np.sin(4 * np.pi * X) * (np.exp(-(Y - 0.25) ** 2 / 0.005) + np.exp(-(Y - 0.75) ** 2 / 0.005))
    pr = 2.5 * np.ones_like(X)
    return kh_conserved(rho, vx, vy, pr, gas_gamma, (1.0 / n) * (1.0 / n))


def kh_reference_ic(n, dx=0.002):
    """Smooth profile of example_Kelvin_Helmholtz/KH_DBA.cu:1102-1125 (integer y / Ny)."""
    y = np.arange(n, dtype=np.float64)[:, None] * np.ones((1, n))
    x = np.ones((n, 1)) * np.arange(n, dtype=np.float64)[None, :]
    rho1, rho2, vx1, vx2 = 1.0, 2.0, 0.5, -0.5
    rho_m, vx_m = 0.5 * (rho1 - rho2), 0.5 * (vx1 - vx2)
    ys = 1.0 * y / n
    rho = np.empty_like(y)
    vx = np.empty_like(y)
    a = y < n // 4
    b = (y >= n // 4) & (y < n // 2)
    c = (y >= n // 2) & (y < 3 * n // 4)
    d = y >= 3 * n // 4
    rho[a] = rho1 - rho_m * np.exp((ys[a] - 0.25) / 0.025)
    vx[a] = vx1 - vx_m * np.exp((ys[a] - 0.25) / 0.025)
    rho[b] = rho2 + rho_m * np.exp((-ys[b] + 0.25) / 0.025)
    vx[b] = vx2 + vx_m * np.exp((-ys[b] + 0.25) / 0.025)
    rho[c] = rho2 + rho_m * np.exp((ys[c] - 0.75) / 0.025)
    vx[c] = vx2 + vx_m * np.exp((ys[c] - 0.75) / 0.025)
    rho[d] = rho1 - rho_m * np.exp((-ys[d] + 0.75) / 0.025)
    vx[d] = vx1 - vx_m * np.exp((-ys[d] + 0.75) / 0.025)
    vy = 0.01 * np.sin(4.0 * np.pi * x / n)
    pr = 2.5 * np.ones_like(y)
    return kh_conserved(rho, vx, vy, pr, 5.0 / 3.0, dx * dx)


def modelb_ic(nx, ny, nz, x0=0.2, y0=0.2, z0=0.01, seed=1234):
    """Initialize of example_Reactive_ModelB/kernel.cu:209-211 with host uniforms:
    x = x0 + 0.05 (u - 1/2), y = y0 + 0.05 (u - 1/2), z = z0 + 0.001 (u - 1/2);
    three independent streams (seeds seed, seed+1, seed+2) of splitmix64."""
    n = nx * ny * nz
    shp = (nz, ny, nx)
    x = x0 + 0.05 * (uniform01(seed, n) - 0.5)
    y = y0 + 0.05 * (uniform01(seed + 1, n) - 0.5)
    z = z0 + 0.001 * (uniform01(seed + 2, n) - 0.5)
    return x.reshape(shp), y.reshape(shp), z.reshape(shp)


def modelb_params(n=256, nz=None, tile=4, mode=abi.MODE_DBA, eps=1e-6, dt=0.0005, history_cap=0, **kw):
    """The reference ternary Model B (example_Reactive_ModelB/Input.txt)."""
    nz = n if nz is None else nz
    p = base_params(abi.MODEL_MODELB, n, n, nz, tile, mode)
    p.tile_z = tile if nz > 1 else 1
    p.bc_x = p.bc_y = p.bc_z = abi.BC_PERIODIC
    p.dx = p.dy = p.dz = 1.0
    p.dt = dt
    p.chi_xy = p.chi_xz = p.chi_yz = p.chi_xb = p.chi_yb = 0.0
    p.chi_zb = 4.0
    p.v_x, p.v_y, p.v_z = 1.0, 1.0, 2.0
    p.r_x, p.r_y, p.r_z = 1.0, 1.0, 2.0
    p.n_exp = p.m_exp = 1.0
    p.epsx_sq = p.epsy_sq = p.epsz_sq = 4.0
    p.epsphi_sq = 0.01
    p.mob_x = p.mob_y = p.mob_z = p.mob_phi = 1.0
    p.k_eq, p.z_crit, p.p_gel, p.k_0 = 100.0, 0.4, 1.0, 1.0
    p.thresh_a = eps
    p.history_cap = history_cap
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def dendrite_k94_params(n=256, nz=None, tile=4, mode=abi.MODE_DBA, history_cap=0, **kw):
    """The reference 3D dendrite (example_Dendritic_growth/Input.txt)."""
    nz = n if nz is None else nz
    p = base_params(abi.MODEL_DENDRITE_K94, n, n, nz, tile, mode)
    p.tile_z = tile if nz > 1 else 1
    p.bc_x = p.bc_y = p.bc_z = abi.BC_PERIODIC
    p.dx = p.dy = p.dz = 0.02
    p.dt = 0.00005
    p.delta, p.tau, p.eps_bar, p.alpha, p.gamma = 0.5, 0.0003, 0.01, 0.9, 40.0
    p.thresh_a = 0.01
    p.thresh_b = 1e-5
    p.history_cap = history_cap
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def dendrite_k94_ic(nx, ny, nz, Delta=0.25):
    """Initialize of example_Dendritic_growth/kernel.cu:167-176: p = 1 - 0.2 dist
    within dist <= 5 of (Nx/2, Ny/2, Nz/2), T = -Delta."""
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    dist = np.sqrt(((x - nx // 2) ** 2 + (y - ny // 2) ** 2 + (z - nz // 2) ** 2).astype(np.float64))
    p = np.where(dist <= 5, 1.0 - 0.2 * dist, 0.0)
    return p, np.full((nz, ny, nx), -Delta)
