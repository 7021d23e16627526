"""GPU: the reference's 3D models stepped as a batch (one merged worklist, the
multi-instance path of the pipelined step kernels and a multi-segment
finalize) give bit-identical fields and awake sets to the same instances
stepped alone (the single-instance path). The alone runs are pinned to the
oracle by tests/test_gpu_ref3d.py."""
import numpy as np
import pytest

from dba_amd import abi, configs

pytestmark = pytest.mark.gpu


def modelb_case(n, nz, mode, seed_shift):
    p = configs.modelb_params(n=n, nz=nz, mode=mode, eps=1e-4)
    X = list(configs.modelb_ic(n, n, nz)) + [np.zeros((nz, n, n))]
    # a uniform slab that sleeps, so the batch mixes awake and sleeping blocks
    X[0][..., n // 2:] = 0.2
    X[1][..., n // 2:] = 0.2
    X[2][..., n // 2:] = 0.01 + 0.001 * seed_shift
    return p, X


def run(ctx, cases, steps, together):
    """Fields after all steps, and every instance's awake set after each chunk."""
    insts = []
    for p, X in cases:
        inst = ctx.create(p)
        for f, a in enumerate(X):
            inst.upload(f, a)
        insts.append(inst)
    masks = [[] for _ in insts]
    for chunk in steps:
        if together:
            ctx.step(insts, chunk)
        else:
            for inst in insts:
                inst.step(chunk)
        for k, inst in enumerate(insts):
            masks[k].append(inst.active(0)[1])
    out = [([inst.download(f) for f in range(len(X))], m) for inst, (_, X), m in zip(insts, cases, masks)]
    for inst in insts:
        inst.destroy()
    return out


def same(alone, batch):
    for (fa, ma), (fb, mb) in zip(alone, batch):
        for a, b in zip(fa, fb):
            assert np.array_equal(a, b)
        for a, b in zip(ma, mb):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [abi.MODE_DBA, abi.MODE_CONSERVED])
def test_modelb_batched_equals_alone(ctx, mode):
    # 48^3 has 1728 blocks: two finalize segments
    cases = [modelb_case(16, 16, mode, 0), modelb_case(48, 48, mode, 1), modelb_case(24, 8, mode, 2)]
    steps = (1, 5, 14)
    alone = run(ctx, cases, steps, together=False)
    batch = run(ctx, cases, steps, together=True)
    same(alone, batch)
    # after the first step the uniform half of the large instance sleeps
    m = alone[1][1][0]
    assert 0 < m.sum() < m.size


def test_k94_batched_equals_alone(ctx):
    cases = []
    for n in (32, 48):
        p = configs.dendrite_k94_params(n=n, nz=n)
        phi, T = configs.dendrite_k94_ic(n, n, n)
        cases.append((p, [phi, T]))
    steps = (3, 20)
    alone = run(ctx, cases, steps, together=False)
    batch = run(ctx, cases, steps, together=True)
    same(alone, batch)
    m = alone[1][1][-1]
    assert 0 < m.sum() < m.size   # the dendrite wakes only blocks near the seed
