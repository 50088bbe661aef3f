"""The debug build (`make debug`, libtesserae_hip_debug.so; the counterpart of the reference's debug_mode
preference, src/Core/debug.jl:5-15): every binning is read back and its invariants checked on the host
(permutation, sorted keys, each tile's run, the active-tile list).  Workloads that exercise each binning
path (full sort, kept layout, merged movers, layered reorder, batches eager and in graph mode, the
adjoint's replays) run through it without a check failing, and give the release build's results bit for
bit (the checks only read)."""
import os

import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import particles_np, to_gpu
from tesserae_hip import _abi
from tesserae_hip.scenes import batched_scenes, block_drop, inverse_cube, random_box

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(_abi.DEBUG_LIB_PATH), reason="debug build not made")]

FIELDS = ("x", "v", "F", "sigma", "V")


def _equal(a, b):
    for name in FIELDS:
        np.testing.assert_array_equal(a[name], b[name], err_msg=name)


def test_build_flags():
    rel, dbg = th.Context(), th.Context(library=_abi.DEBUG_LIB_PATH)
    assert rel.lib.tess_build_flags() == 0
    assert dbg.lib.tess_build_flags() == _abi.TESS_BUILD_DEBUG_CHECKS
    dbg.close()
    rel.close()


def _steps(sc, library, nsteps, reorder):
    ctx = th.Context(library=library)
    mesh, P, G = to_gpu(sc)
    th.step(G, P, sc.basis, sc.params, nsteps=nsteps, reorder_threshold=reorder, ctx=ctx)
    ctx.sync()
    out = particles_np(P)
    ctx.close()
    return out


@pytest.mark.parametrize("case", ["block_drop_reorder", "fast_movers_2d", "fast_movers_3d"])
def test_step_binnings_pass_the_checks(case):
    if case == "block_drop_reorder":
        sc, n, thr = block_drop(dtype=torch.float64, dx=1 / 48), 8, 0.85
    elif case == "fast_movers_2d":
        sc, n, thr = random_box(2, 3000, 7, th.BSpline(th.Quadratic()), h=0.05, vscale=3.0, stress=0.5), 6, 0.999
    else:
        sc, n, thr = random_box(3, 4000, 8, th.BSpline(th.Cubic()), h=0.1, vscale=3.0, stress=0.5), 6, 0.0
    _equal(_steps(sc, _abi.DEBUG_LIB_PATH, n, thr), _steps(sc, None, n, thr))


@pytest.mark.parametrize("variant", [2, 10], ids=["eager", "graph"])
def test_batch_binnings_pass_the_checks(variant):
    scenes = batched_scenes(5, seed=41, dtype=torch.float64, cells=(16, 48), particles=(200, 2500))
    res = []
    for library in (_abi.DEBUG_LIB_PATH, None):
        ctx = th.Context(library=library)
        members, gpu = [], []
        for sc in scenes:
            mesh, P, G = to_gpu(sc)
            members.append((G, P, sc.params))
            gpu.append(P)
        batch = th.Batch(members, th.BSpline(th.Quadratic()), ctx=ctx)
        batch.set_variant(variant)
        batch.set_reorder(0.999)
        batch.step(2)
        batch.step(3)
        ctx.sync()
        res.append([particles_np(P) for P in gpu])
        batch.close()
        ctx.close()
    for a, b in zip(*res):
        _equal(a, b)


def test_adjoint_replays_pass_the_checks():
    sc = inverse_cube(dtype=torch.float64, dx=1 / 16)
    target = torch.from_numpy(sc.x + np.array([0.1, 0.0, 0.0])).cuda()
    res = []
    for library in (_abi.DEBUG_LIB_PATH, None):
        ctx = th.Context(library=library)
        mesh, P, G = to_gpu(sc)
        th.rollout(G, P, sc.basis, sc.params, 9, checkpoint_every=4, ctx=ctx)
        adj = th.Adjoint(sc.n, 3)
        adj.x.copy_(2.0 * (P.get("x").double() - target) / sc.n)
        th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)
        ctx.sync()
        res.append((adj.v.cpu().numpy(), adj.x.cpu().numpy()))
        ctx.close()
    for a, b in zip(*res):
        np.testing.assert_array_equal(a, b)
