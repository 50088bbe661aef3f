"""Failure injection through the C ABI (tess_debug_fail_after), the counterpart of the reference's
tests that inject error("boom") into an equation or a phase (test/transfer.jl:587-604,
test/parallel_region.jl:76-170).

For each entry point below, the k-th HIP call the library checks inside one call is made to report a
failure, for every k until the call finishes before reaching it.  Each injected failure must come back
as TESS_ERR_DEVICE (DeviceError), with the message naming it, instead of aborting or hanging.  After
each one, the same context (and batch) runs the call again from the initial state.  Its result must
equal that of a context that never failed (1e-12 relative, Float64).

The injected failure is reported after the HIP call itself went through.  Every launched kernel still
runs to its end on valid data, so no test here makes the GPU fault.
"""
import numpy as np
import pytest
import torch

import tesserae_hip as th
from gpu_util import particles_np, rel_err, to_gpu
from tesserae_hip.scenes import batched_scenes, inverse_cube, two_disks

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "sigma", "V")


def _restore(P, sc):
    for name in ("x", "v", "V", "F", "sigma", "grad_v", "l", "logE"):
        a = getattr(sc, name)
        if a is not None:
            P.set(name, a)


def _inject_everywhere(lib, run, check, limit=2000):
    """run() with a failure injected at its k-th checked HIP call, k = 1, 2, ..., until run() ends before
    its k-th call; check() after each injected failure.  Returns the number of failures injected."""
    fired = 0
    for k in range(1, limit):
        assert lib.tess_debug_fail_after(k) == 0
        try:
            run()
        except th.DeviceError as e:
            assert "injected failure" in str(e), str(e)
            fired += 1
        else:
            return fired
        finally:
            lib.tess_debug_fail_after(0)
        check()
    raise AssertionError(f"more than {limit} checked calls")


def test_fail_after_rejects_negative():
    ctx = th.Context()
    assert ctx.lib.tess_debug_fail_after(-1) != 0
    assert ctx.lib.tess_debug_fail_after(0) == 0
    ctx.close()


def test_step_recovers_from_every_injected_failure():
    sc = two_disks(dtype=torch.float64, h=0.04, steps=3)
    ref_ctx = th.Context()
    mesh, Pr, Gr = to_gpu(sc)
    th.step(Gr, Pr, sc.basis, sc.params, nsteps=3, ctx=ref_ctx)
    ref_ctx.sync()
    ref = particles_np(Pr)
    ref_m = Gr.get("m").double().cpu().numpy()
    ref_ctx.close()

    ctx = th.Context()
    mesh, P, G = to_gpu(sc)

    def run():
        _restore(P, sc)
        th.step(G, P, sc.basis, sc.params, nsteps=3, ctx=ctx)
        ctx.sync()

    def check():
        run()
        got = particles_np(P)
        for name in FIELDS:
            assert rel_err(got[name], ref[name]) <= 1e-12, name
        assert rel_err(G.get("m").double().cpu().numpy(), ref_m) <= 1e-12

    fired = _inject_everywhere(ctx.lib, run, check)
    print(f"[failure] step: {fired} injected failures recovered", flush=True)
    assert fired >= 10
    check()
    ctx.close()


def test_rollout_and_pullback_recover_from_every_injected_failure():
    sc = inverse_cube(dtype=torch.float64, dx=1 / 16)
    target = sc.x + np.array([0.1, 0.0, 0.0])

    def grads(ctx, P, G):
        _restore(P, sc)
        th.rollout(G, P, sc.basis, sc.params, 8, checkpoint_every=3, ctx=ctx)
        x = P.get("x").double()
        adj = th.Adjoint(sc.n, 3)
        adj.x.copy_(2.0 * (x - torch.from_numpy(target).cuda()) / sc.n)
        th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)
        ctx.sync()
        return adj.v.cpu().numpy(), adj.x.cpu().numpy()

    ref_ctx = th.Context()
    mesh, Pr, Gr = to_gpu(sc)
    ref = grads(ref_ctx, Pr, Gr)
    ref_ctx.close()

    ctx = th.Context()
    mesh, P, G = to_gpu(sc)
    adj_box = {}

    def run():
        adj_box["g"] = grads(ctx, P, G)

    def check():
        got = grads(ctx, P, G)
        for a, b in zip(got, ref):
            assert rel_err(a, b) <= 1e-12

    fired = _inject_everywhere(ctx.lib, run, check)
    print(f"[failure] rollout + pullback: {fired} injected failures recovered", flush=True)
    assert fired >= 20
    check()
    ctx.close()


@pytest.mark.parametrize("variant", [2, 10], ids=["eager", "graph"])
def test_batch_recovers_from_every_injected_failure(variant):
    scenes = batched_scenes(4, seed=31, dtype=torch.float64, cells=(16, 40), particles=(100, 1500))
    basis = th.BSpline(th.Quadratic())

    ref_ctx = th.Context()
    members, ref_P = [], []
    for sc in scenes:
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
        ref_P.append(P)
    rb = th.Batch(members, basis, ctx=ref_ctx)
    rb.set_variant(variant)
    rb.step(3)
    ref_ctx.sync()
    ref = [particles_np(P) for P in ref_P]
    rb.close()
    ref_ctx.close()

    ctx = th.Context()
    members, gpu = [], []
    for sc in scenes:
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
        gpu.append(P)
    batch = th.Batch(members, basis, ctx=ctx)
    batch.set_variant(variant)

    def run():
        for sc, P in zip(scenes, gpu):
            _restore(P, sc)
        batch.step(3)
        ctx.sync()

    def check():
        run()
        for P, r in zip(gpu, ref):
            got = particles_np(P)
            for name in FIELDS:
                assert rel_err(got[name], r[name]) <= 1e-12, name

    fired = _inject_everywhere(ctx.lib, run, check)
    print(f"[failure] batch variant {variant}: {fired} injected failures recovered", flush=True)
    assert fired >= 10
    check()
    batch.close()
    ctx.close()


@pytest.mark.parametrize("sweep", ["0", "1"], ids=["composed", "sweep"])
def test_g2p2g_steps_recover_from_every_injected_failure(sweep, monkeypatch):
    """tess_g2p2g_steps (three passes, reorder at 0.85), with the composed kernels and with the single-sweep kernel
    (TESS_G2P2G_SWEEP=1)."""
    monkeypatch.setenv("TESS_G2P2G_SWEEP", sweep)
    sc = two_disks(dtype=torch.float64, h=0.04, steps=3)

    def fresh_run(ctx, P, G):
        _restore(P, sc)
        th.P2G(G, P, sc.basis, sc.params, ctx=ctx)
        th.g2p2g_steps(G, P, sc.basis, sc.params, nsteps=3, reorder_threshold=0.85, ctx=ctx)
        ctx.sync()

    ref_ctx = th.Context()
    mesh, Pr, Gr = to_gpu(sc)
    fresh_run(ref_ctx, Pr, Gr)
    ref = particles_np(Pr)
    ref_ctx.close()

    ctx = th.Context()
    mesh, P, G = to_gpu(sc)

    def run():
        fresh_run(ctx, P, G)

    def check():
        run()
        got = particles_np(P)
        key = lambda p: np.lexsort(p["x"].T[::-1])  # noqa: E731  (the reorder may permute the particles)
        ia, ib = key(got), key(ref)
        for name in FIELDS:
            assert rel_err(got[name][ia], ref[name][ib]) <= 1e-12, name

    fired = _inject_everywhere(ctx.lib, run, check)
    print(f"[failure] g2p2g_steps sweep={sweep}: {fired} injected failures recovered", flush=True)
    assert fired >= 10
    check()
    ctx.close()
