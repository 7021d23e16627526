"""Helpers shared by the GPU parity tests (host re-evaluation of the DBA rule)."""
import numpy as np

from oracle.oracle import dilate


def tile_any(mask_cells, tile):
    ny, nx = mask_cells.shape
    return mask_cells.reshape(ny // tile, tile, nx // tile, tile).any(axis=(1, 3)).reshape(-1)


def lap_periodic(c, stencil, dx):
    """Same operation order as the step kernels' LDS stencils."""
    E, W = np.roll(c, -1, 1), np.roll(c, 1, 1)
    N, S = np.roll(c, -1, 0), np.roll(c, 1, 0)
    s1 = (E + W) + (N + S)
    if stencil == 5:
        return (s1 - 4.0 * c) * (1.0 / (dx * dx))
    NE, NW = np.roll(E, -1, 0), np.roll(W, -1, 0)
    SE, SW = np.roll(E, 1, 0), np.roll(W, 1, 0)
    s2 = (NE + NW) + (SE + SW)
    return ((4.0 * s1 + s2) - 20.0 * c) * (1.0 / (6.0 * dx * dx))


def dilate_faces(N, ntx, nty, periodic):
    w = 1 if periodic else 0
    return dilate(N.astype(np.uint8), ntx, nty, 1, 0, (w, w, w))
