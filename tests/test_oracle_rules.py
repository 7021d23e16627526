"""Pins of the activation rule pieces of the oracle by brute force.

Wakeup_Neighbor (example_Dendritic_growth/kernel.cu:577-634: face neighbours,
periodic wrap of the block grid) and the KH 3x3 dilation
(example_Kelvin_Helmholtz/KH_DBA.cu:988-1044) are checked against a direct
set construction over every flag pattern of small block grids; the compaction
against sorted index lists (std::copy_if order).
"""
import itertools

import numpy as np

from oracle.oracle import compact, dilate


def brute_dilate(N, ntx, nty, kind, wrap):
    A = np.zeros(ntx * nty, dtype=np.uint8)
    offs = [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)]
    if kind == 1:
        offs += [(1, 1), (1, -1), (-1, 1), (-1, -1)]
    for j in range(nty):
        for i in range(ntx):
            if not N[i + j * ntx]:
                continue
            for di, dj in offs:
                ii, jj = i + di, j + dj
                if wrap:
                    ii %= ntx
                    jj %= nty
                elif not (0 <= ii < ntx and 0 <= jj < nty):
                    continue
                A[ii + jj * ntx] = 1
    return A


def test_dilation_all_patterns_3x3():
    for kind in (0, 1):
        for wrap in (0, 1):
            for bits in itertools.product((0, 1), repeat=9):
                N = np.array(bits, dtype=np.uint8)
                got = dilate(N, 3, 3, 1, kind, (wrap, wrap, wrap))
                assert np.array_equal(got, brute_dilate(N, 3, 3, kind, wrap)), (kind, wrap, bits)


def test_dilation_random_rectangles():
    rng = np.random.default_rng(7)
    for _ in range(200):
        ntx, nty = rng.integers(1, 9, size=2)
        N = (rng.random(ntx * nty) < 0.15).astype(np.uint8)
        for kind in (0, 1):
            for wrap in (0, 1):
                got = dilate(N, ntx, nty, 1, kind, (wrap, wrap, wrap))
                assert np.array_equal(got, brute_dilate(N, ntx, nty, kind, wrap))


def test_dilation_3d_faces():
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = rng.integers(1, 5, size=3)
        N = (rng.random(n.prod()) < 0.1).astype(np.uint8)
        got = dilate(N, n[0], n[1], n[2], 0, (1, 1, 1)).reshape(n[2], n[1], n[0])
        ref = np.zeros((n[2], n[1], n[0]), dtype=np.uint8)
        for k, j, i in zip(*np.nonzero(N.reshape(n[2], n[1], n[0]))):
            for d in [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
                ref[(k + d[2]) % n[2], (j + d[1]) % n[1], (i + d[0]) % n[0]] = 1
        assert np.array_equal(got, ref)


def test_compact_equals_copy_if():
    rng = np.random.default_rng(11)
    for n in (0, 1, 63, 64, 65, 1000):
        A = (rng.random(n) < 0.3).astype(np.uint8)
        assert np.array_equal(compact(A), np.nonzero(A)[0].astype(np.int32))
