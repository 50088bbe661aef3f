"""Batches of uneven scenes: one launch per kernel for all scenes, each scene
equal to the oracle's run of that scene alone (Float64, 1e-10)."""
import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import particles_np, rel_err, to_gpu
from oracle_bind import HostState
from tesserae_hip.scenes import batched_scenes, random_box

pytestmark = pytest.mark.gpu


def _run_and_check(oracle, scenes, basis, nsteps, tune=0, variant=None, calls=None):
    members, gpu = [], []
    for sc in scenes:
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
        gpu.append(P)
    ctx = th.Context()
    batch = th.Batch(members, basis, ctx=ctx)
    extra = 0
    if tune:
        v = batch.autotune(tune)
        eager = list(range(6)) if scenes[0].mesh.dim == 2 else [0, 2]
        assert v in eager + [e | 8 for e in eager] + [e | 24 for e in eager]
        extra = 3 * len(eager) * (tune + 1)
    if variant is not None:
        batch.set_variant(variant)
    for k in calls or (nsteps,):
        batch.step(k)
    ctx.sync()
    if variant is not None and variant & 8:
        st = batch.stats()
        print(f"[batch graph] variant={variant} {st}", flush=True)
        # each call's first step sorts every particle again outside the graphs
        assert st["captures"] == 1 and st["graph_replays"] == nsteps - len(calls or (nsteps,))
    for sc, P in zip(scenes, gpu):
        st = HostState(sc)
        oracle.step(st, basis, sc.params, nsteps + extra)
        got = particles_np(P)
        for name in ("x", "v", "F", "sigma", "V"):
            assert rel_err(got[name], getattr(st, name)) <= 1e-10, (sc.name, name)
    batch.close()
    ctx.close()


def test_batch_2d_uneven_scenes_match_oracle(oracle):
    scenes = batched_scenes(12, seed=3, dtype=torch.float64, cells=(16, 64), particles=(100, 3000))
    print(f"[batch2d] n={[s.n for s in scenes]}", flush=True)
    _run_and_check(oracle, scenes, th.BSpline(th.Quadratic()), 5)


def test_batch_autotune_then_steps(oracle):
    scenes = batched_scenes(5, seed=9, dtype=torch.float64, cells=(16, 48), particles=(100, 2000))
    _run_and_check(oracle, scenes, th.BSpline(th.Quadratic()), 2, tune=1)


@pytest.mark.parametrize("variant", range(6))
def test_batch_2d_every_launch_variant_matches_oracle(oracle, variant):
    """P2G with one wave or one per last-axis offset, G2P with 4, 2 or 1 waves per tile."""
    scenes = batched_scenes(6, seed=11, dtype=torch.float64, cells=(16, 64), particles=(100, 3000))
    _run_and_check(oracle, scenes, th.BSpline(th.Quadratic()), 3, variant=variant)


def test_batch_set_variant_rejects_unknown_numbers():
    sc = random_box(3, 200, 1, th.BSpline(th.Quadratic()), dtype=torch.float64)
    mesh, P, G = to_gpu(sc)
    batch = th.Batch([(G, P, sc.params)], th.BSpline(th.Quadratic()))
    for bad in (1, 3, 4, 6, 9, 16, 18, 32, -1):
        with pytest.raises(th.ArgumentError):
            batch.set_variant(bad)
    batch.set_variant(2)
    batch.set_variant(10)
    batch.set_variant(26)
    batch.close()


@pytest.mark.parametrize("basis", [th.BSpline(th.Linear()), th.BSpline(th.Cubic())], ids=["linear", "cubic"])
def test_batch_3d_matches_oracle(oracle, basis):
    scenes = [random_box(3, 800 + 300 * k, 40 + k, basis, dtype=torch.float64, h=0.1 / (k + 1), lo=0.1, hi=0.8,
                         stress=0.5) for k in range(3)]
    _run_and_check(oracle, scenes, basis, 3)


def test_batch_rejects_mixed_dimensions():
    a = random_box(2, 100, 1, th.BSpline(th.Quadratic()), dtype=torch.float64)
    b = random_box(3, 100, 2, th.BSpline(th.Quadratic()), dtype=torch.float64)
    members = []
    for sc in (a, b):
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
    with pytest.raises(th.ArgumentError):
        th.Batch(members, th.BSpline(th.Quadratic()))


@pytest.mark.parametrize("threshold", [1.0, 0.999], ids=["every_step", "by_contiguity"])
def test_batch_reorder_matches_oracle(oracle, threshold):
    """With the layered reorder on (threshold 1: every step; below 1: when the contiguity score counted on a
    call's first step by the layered-order pass, later by the previous step's P2G, is below it), every scene
    still equals the oracle's run of it alone, up to the order of its particles; particles outside the mesh
    stay in their scene; the arrays were permuted.  Two calls, of two and four steps."""
    basis = th.BSpline(th.Quadratic())
    # fast particles (|v| up to 3, a few percent of a cell per step) break the layered order within a call
    scenes = [random_box(2, 300 + 400 * k, 70 + k, basis, dtype=torch.float64, h=0.05 * (1 + k % 3), lo=-0.05,
                         hi=1.05, stress=0.5, vscale=3.0) for k in range(6)]
    members, gpu = [], []
    for sc in scenes:
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
        gpu.append(P)
    ctx = th.Context()
    batch = th.Batch(members, basis, ctx=ctx)
    batch.set_reorder(threshold)
    batch.step(2)
    batch.step(4)  # a P2G count reflects the previous G2P's moves from the call's third step on
    ctx.sync()
    stats = batch.stats()
    print(f"[batch reorder] threshold={threshold} {stats}", flush=True)
    assert stats["reorders"] >= 1
    if threshold < 1.0:
        # within a call, steps after the first decide from the previous step's P2G count of broken pairs
        assert stats["reorders_by_count"] >= 1
    permuted = 0
    for sc, P in zip(scenes, gpu):
        st = HostState(sc)
        oracle.step(st, basis, sc.params, 6)
        got = particles_np(P)
        ig, io = np.lexsort(got["x"].T[::-1]), np.lexsort(st.x.T[::-1])
        for name in ("x", "v", "F", "sigma", "V"):
            assert rel_err(got[name][ig], getattr(st, name)[io]) <= 1e-10, (sc.name, name)
        permuted += int(np.any(ig != io))
    assert permuted > 0
    batch.close()
    ctx.close()


@pytest.mark.parametrize("variant", [8, 9, 10, 13, 24, 26, 29])
def test_batch_graph_mode_matches_oracle(oracle, variant):
    """Graph mode: each call's first step sorts every particle; the rest are replays of the captured hipGraphs
    (the binning merges the particles whose key changed into the previous order; launches over device-side
    counts: one workgroup per possible tile, or persistent grids with bit 4).  Two calls, of two and three
    steps: the second call replays the graphs captured by the first."""
    scenes = batched_scenes(7, seed=21, dtype=torch.float64, cells=(16, 64), particles=(100, 3000))
    _run_and_check(oracle, scenes, th.BSpline(th.Quadratic()), 5, variant=variant, calls=(2, 3))


def test_batch_graph_mode_3d_matches_oracle(oracle):
    basis = th.BSpline(th.Quadratic())
    scenes = [random_box(3, 700 + 400 * k, 80 + k, basis, dtype=torch.float64, h=0.1 / (k + 1), lo=0.05, hi=0.9,
                         stress=0.5) for k in range(3)]
    _run_and_check(oracle, scenes, basis, 4, variant=8)
    _run_and_check(oracle, scenes, basis, 4, variant=26)


def test_batch_switches_between_eager_and_graph_mode(oracle):
    """Eager, graph, eager again on the same batch: each mode zeroes the node blocks the other left behind."""
    basis = th.BSpline(th.Quadratic())
    scenes = batched_scenes(5, seed=23, dtype=torch.float64, cells=(16, 48), particles=(100, 2500))
    members, gpu = [], []
    for sc in scenes:
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
        gpu.append(P)
    ctx = th.Context()
    batch = th.Batch(members, basis, ctx=ctx)
    batch.set_variant(2)
    batch.step(2)
    batch.set_variant(10)
    batch.step(3)
    batch.set_variant(2)
    batch.step(2)
    ctx.sync()
    for sc, P in zip(scenes, gpu):
        st = HostState(sc)
        oracle.step(st, basis, sc.params, 7)
        got = particles_np(P)
        for name in ("x", "v", "F", "sigma", "V"):
            assert rel_err(got[name], getattr(st, name)) <= 1e-10, (sc.name, name)
    batch.close()
    ctx.close()


@pytest.mark.parametrize("threshold", [1.0, 0.999], ids=["every_chunk", "by_contiguity"])
def test_batch_graph_mode_reorder_matches_oracle(oracle, threshold):
    """Graph mode with the layered reorder: the test runs before the call's replays, on the batch's private
    stream; every scene equals the oracle's run of it up to the order of its particles."""
    basis = th.BSpline(th.Quadratic())
    scenes = [random_box(2, 300 + 400 * k, 90 + k, basis, dtype=torch.float64, h=0.05 * (1 + k % 3), lo=-0.05,
                         hi=1.05, stress=0.5, vscale=3.0) for k in range(6)]
    members, gpu = [], []
    for sc in scenes:
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
        gpu.append(P)
    ctx = th.Context()
    batch = th.Batch(members, basis, ctx=ctx)
    batch.set_variant(8)
    batch.set_reorder(threshold)
    batch.step(3)
    batch.step(3)
    ctx.sync()
    stats = batch.stats()
    print(f"[batch graph reorder] threshold={threshold} {stats}", flush=True)
    if threshold >= 1.0:
        assert stats["reorders"] >= 2
    for sc, P in zip(scenes, gpu):
        st = HostState(sc)
        oracle.step(st, basis, sc.params, 6)
        got = particles_np(P)
        ig, io = np.lexsort(got["x"].T[::-1]), np.lexsort(st.x.T[::-1])
        for name in ("x", "v", "F", "sigma", "V"):
            assert rel_err(got[name][ig], getattr(st, name)[io]) <= 1e-10, (sc.name, name)
    batch.close()
    ctx.close()
