"""FP32 engine runs against the FP64 oracle, per config: relative L2 and L∞ error on x, v and F.

For each config the same seeded scene is stepped by the engine in Float32 (tess_step / tess_batch_step /
tess_step_sparse, deferred weights, no reordering) and by the oracle's threaded restatement of the
reference's Float64 CPU path (stored weights, CPU block partition, colour-phased P2G).  Errors are
||a - b||_2 / ||b||_2 and max|a - b| / max|b| over all particles and components.

Configs (the sizes the oracle can run in minutes; reductions are stated in the output):
  1  2D two-disk collision, dx=0.01, 1000 steps (full size)
  2  3D block drop, dx=1/256, 13,434,880 particles, 100 steps (full size)
  3  3D sphere on a sparse grid, cubic and uGIMP, 50 steps, at dx=1/128 instead of 1/512
  4  256 uneven 2D scenes in one batch, 200 steps; the oracle checks the first 16 scenes
  5  3D inverse-problem cube, dx=1/64, 200 forward steps (full size)

Run on a GPU:  python tests/fp32_error_report.py [--configs 1,2,3,4,5] [--threads 16] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tesserae.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import tesserae_hip as th  # noqa: E402
from gpu_util import particles_np, to_gpu  # noqa: E402
from oracle_bind import HostState, Oracle  # noqa: E402
from tesserae_hip import scenes  # noqa: E402


def log(msg):
    print(msg, flush=True)


def errors(got: dict, st: HostState) -> dict:
    out = {}
    for name in ("x", "v", "F"):
        a = np.asarray(got[name], np.float64).ravel()
        b = np.asarray(getattr(st, name), np.float64).ravel()
        d = a - b
        out[name] = {"L2": float(np.linalg.norm(d) / max(np.linalg.norm(b), 1e-300)),
                     "Linf": float(np.abs(d).max() / max(np.abs(b).max(), 1e-300))}
    return out


def oracle_steps(oracle, st, basis, params, nsteps, threads, chunk=10):
    W = None
    done = 0
    while done < nsteps:
        k = min(chunk, nsteps - done)
        W = oracle.step_threaded(st, basis, params, threads, k, W)
        done += k
        log(f"  oracle {done}/{nsteps} steps")


def single(oracle, sc, steps, threads, sparse=False):
    mesh, P, G = to_gpu(sc, dtype=torch.float32)
    if sparse:
        mesh = th.CartesianMesh(sc.mesh.h, *[(a, b) for a, b in zip(sc.mesh.xmin, sc.mesh.xmax())],
                                dtype=torch.float32, block_size_log2=sc.mesh.block_size_log2)
        G = th.generate_grid(th.SpArray, mesh)
    th.step(G, P, sc.basis, sc.params, nsteps=steps)
    torch.cuda.synchronize()
    got = particles_np(P)
    del P, G
    torch.cuda.empty_cache()
    log(f"  engine FP32 {steps} steps done")
    st = HostState(sc)
    oracle_steps(oracle, st, sc.basis, sc.params, steps, threads)
    return errors(got, st)


def run(cfg, oracle, threads):
    t0 = time.time()
    if cfg == 1:
        sc = scenes.two_disks(dtype=torch.float64)
        res = {"workload": f"two disks dx=0.01, {sc.n} particles, quadratic, 1000 steps",
               "errors": single(oracle, sc, 1000, threads)}
    elif cfg == 2:
        sc = scenes.block_drop(dtype=torch.float64)
        res = {"workload": f"block drop dx=1/256, {sc.n} particles, quadratic, 100 steps",
               "errors": single(oracle, sc, 100, threads)}
    elif cfg == 3:
        res = {"workload": "sparse-grid sphere at dx=1/128 (reduced from 1/512), SpArray 8^3 blocks, 50 steps",
               "by_basis": {}}
        for basis in (th.BSpline(th.Cubic()), th.uGIMP()):
            sc = scenes.sphere_spin(basis, dtype=torch.float64, dx=1 / 128)
            log(f"  sphere {basis!r}: {sc.n} particles")
            res["by_basis"][repr(basis)] = {"particles": sc.n, "errors": single(oracle, sc, 50, threads, sparse=True)}
    elif cfg == 4:
        scs = scenes.batched_scenes(256, dtype=torch.float64)
        members, gpu = [], []
        for sc in scs:
            mesh, P, G = to_gpu(sc, dtype=torch.float32)
            members.append((G, P, sc.params))
            gpu.append(P)
        ctx = th.Context()
        batch = th.Batch(members, scs[0].basis, ctx=ctx)
        batch.step(200)
        ctx.sync()
        log("  engine FP32 batch 200 steps done")
        per = []
        for k in range(16):
            st = HostState(scs[k])
            oracle_steps(oracle, st, scs[k].basis, scs[k].params, 200, threads, chunk=200)
            per.append({"scene": k, "particles": scs[k].n, "errors": errors(particles_np(gpu[k]), st)})
        batch.close()
        ctx.close()
        worst = {name: {m: max(p["errors"][name][m] for p in per) for m in ("L2", "Linf")} for name in ("x", "v", "F")}
        res = {"workload": f"256 uneven 2D scenes ({sum(s.n for s in scs)} particles), one batch, 200 steps; "
                           "oracle on scenes 0-15", "errors_worst_of_checked": worst, "scenes": per}
    elif cfg == 5:
        sc = scenes.inverse_cube(dtype=torch.float64)
        res = {"workload": f"inverse-problem cube dx=1/64, {sc.n} particles, 200 forward steps",
               "errors": single(oracle, sc, 200, threads)}
    else:
        raise SystemExit(f"unknown config {cfg}")
    res["seconds"] = time.time() - t0
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="1,2,3,4,5")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--out", default=os.path.join(ROOT, "gpurun_out", "fp32_errors.json"))
    args = ap.parse_args()
    torch.cuda.set_device(0)
    oracle = Oracle()
    out = {"threads": args.threads}
    for c in [int(s) for s in args.configs.split(",") if s]:
        log(f"config {c}")
        out[f"config{c}"] = run(c, oracle, args.threads)
        log(json.dumps(out[f"config{c}"])[:600])
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
