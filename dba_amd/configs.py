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
                        theta0=0.0, threads=0, history_cap=0, bc=abi.BC_NOFLUX):
    p = base_params(abi.MODEL_DENDRITE_K93, n, n, 1, tile, mode)
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
