"""Host reference of the droplet (coarsening) analysis used by the GPU and driver
tests: scipy.ndimage.label of !(v < avg), unions of the labels that touch across
each periodic face, droplets = components holding a cell v > avg, ordered by
their first such cell in index order (x fastest), as
example_Reactive_ModelB/coarsening_analysis.cpp:48-170 scans."""
import numpy as np
from scipy import ndimage


def host_droplets(a):
    """(avg, volume, ordered sizes) of the array a (shape (nz, ny, nx) or (ny, nx))."""
    avg = (max(0.5, a.max()) + min(0.5, a.min())) / 2.0
    member = ~(a < avg)
    lab, n = ndimage.label(member)
    parent = list(range(n + 1))

    def root(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    for ax in range(a.ndim):
        if a.shape[ax] == 1:
            continue
        lo = np.take(lab, 0, axis=ax).ravel()
        hi = np.take(lab, -1, axis=ax).ravel()
        for u, v in zip(lo, hi):
            if u and v:
                ru, rv = root(u), root(v)
                if ru != rv:
                    parent[ru] = rv
    flat = lab.ravel()
    roots = np.array([root(l) if l else 0 for l in range(n + 1)])
    rl = roots[flat]
    sizes = np.bincount(rl, minlength=n + 1)
    seed = (a.ravel() > avg) & (rl > 0)
    first = {}
    for idx in np.flatnonzero(seed):
        r = rl[idx]
        if r not in first:
            first[r] = idx
    order = sorted(first, key=lambda r: first[r])
    return avg, int((a > avg).sum()), [int(sizes[r]) for r in order]


def radius(volume, count, three_d):
    """coarsening_analysis.cpp:170 (3D) and its 2D counterpart."""
    if three_d:
        return float(np.cbrt(volume * 0.75 / np.pi / count))
    return float(np.sqrt(volume / np.pi / count))
