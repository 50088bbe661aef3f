"""Throughput of the explicit MPM step on one MI355X.

Workload: 3D elastic block drop, dx=1/256 on [0,1]^3, quadratic B-spline,
2x2x2 particles per cell in [0.25,0.75]^2 x [0.5,0.9] (13,434,880 particles),
jittered +-0.1dx, gravity, slip walls, Neo-Hookean E=1e6 nu=0.3 rho0=1e3,
dt=1e-4, Float32.  One step = update!(weights) (deferred) + @P2G (with grid
update) + @G2P, exactly the README loop body, with the reference's
reorder_particles!(threshold=0.85) maintenance check every step.  The warm-up
steps and the timed steps are each one call of the step loop (tess_step with
nsteps = W, then K): inside a call every binning after the first starts from
the keys the previous G2P wrote.

Prints one JSON line (metric particle-steps/s) with the roofline fraction of
the dominant kernel and the CPU baseline (the oracle's threaded restatement of
the reference CPU algorithm) timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "tesserae.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the block-drop kernels from two rocprofv3 --pmc runs
# (FETCH_SIZE, then WRITE_SIZE) scaled by the same counters' readings for the
# library's plain copy kernel; written by `bench.py --traffic-from` with a digest
# of the sources it measured, and used only while the sources still match
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "block_drop_hbm_traffic.json")
SOURCE_GLOBS = ("tesserae.jl_amd/csrc/*.hip", "tesserae.jl_amd/csrc/*.hpp", "tesserae.jl_amd/csrc/Makefile",
                "include/*.h")


def sources_digest():
    """sha256 over the library's sources (path and content, sorted by path)."""
    import glob
    import hashlib

    h = hashlib.sha256()
    for pat in SOURCE_GLOBS:
        for f in sorted(glob.glob(os.path.join(ROOT, pat))):
            h.update(os.path.relpath(f, ROOT).encode())
            with open(f, "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def profiled_traffic(kernel):
    """(bytes per launch, note): None with the reason when the file is missing or measured other sources."""
    try:
        with open(TRAFFIC_FILE) as f:
            doc = json.load(f)
        rec = doc["kernels"][kernel]
    except (OSError, KeyError, ValueError, TypeError):
        return None, "no counter record for this kernel"
    if doc.get("sources_digest") != sources_digest():
        return None, f"counter record measured sources {doc.get('sources_digest')}, these are {sources_digest()}"
    return float(rec["bytes"]), os.path.relpath(TRAFFIC_FILE, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dx", type=float, default=1 / 256)
    ap.add_argument("--layout", default="julia", choices=["julia", "soa"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--reorder", type=float, default=0.85, help="reorder_particles! threshold per step")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward ratio leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--no-sphere", action="store_true", help="skip the sparse-grid sphere leg")
    ap.add_argument("--no-taylor", action="store_true", help="skip the Taylor-impact rod (KernelCorrection) leg")
    ap.add_argument("--no-g2p2g", action="store_true", help="skip the G2P2G leg")
    ap.add_argument("--sphere-dx", type=float, default=1 / 512)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run the library's plain copy of MIB MiB three times (the reference that scales "
                         "HBM counters to bytes)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write the HBM traffic record from two rocprofv3 --pmc output folders and exit")
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed steps of each CPU baseline run")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--checkpoint-every", type=int, default=20)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_bandwidth(ctx, mib=1024, reps=5):
    """The library's plain device copy (tess_device_copy) of `mib` MiB timed with HIP events on its stream: the
    box's measured copy rate in GB/s (bytes read + bytes written per second)."""
    import torch

    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ctx.device_copy(src, dst)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        ctx.device_copy(src, dst)
    e1.record()
    e1.synchronize()
    gbs = 2 * src.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    assert bool(torch.equal(src, dst))
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads, dtype, steps):
    """The oracle's threaded restatement of the reference CPU algorithm (stored weights, the CPU block
    partition, colour-phased P2G, threaded G2P) on the headline workload itself (13,434,880 particles):
    one warm-up step, then `steps` steps timed one by one."""
    import numpy as np
    import torch

    from oracle_bind import HostState, Oracle
    from tesserae_hip.scenes import block_drop

    sc = block_drop(dtype=dtype)
    st = HostState(sc)
    orc = Oracle()
    W = orc.step_threaded(st, sc.basis, sc.params, threads, nsteps=1)  # warm-up (allocations, page faults)
    per = []
    for _ in range(steps):
        t0 = time.perf_counter()
        orc.step_threaded(st, sc.basis, sc.params, threads, nsteps=1, weights=W)
        per.append(time.perf_counter() - t0)
    med, mean = float(np.median(per)), float(np.mean(per))
    tag = "Float32" if dtype == torch.float32 else "Float64"
    return {"value": sc.n / med, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "median_s_per_step": med, "mean_s_per_step": mean, "mean_value": sc.n / mean,
            "sample": f"block drop dx=1/256 ({sc.n} particles), {tag}, {steps} steps timed one by one after 1 "
                      f"warm-up step, {threads} threads, {sum(per):.1f} s; value from the median step"}


def traffic_record(fetch_dir, write_dir):
    """Per-kernel HBM bytes per launch from the counter CSVs of two rocprofv3 runs of
    `bench.py --steps 3 --warmup 2 --no-... --copy-probe 1024`: each kernel's first dispatch (the cold first step
    after the scene set-up) is left out; FETCH_SIZE and WRITE_SIZE (KiB) are scaled by their readings for the
    library's copy kernel, which reads and writes exactly 1 GiB per dispatch."""
    import csv
    import glob

    def load(d, counter):
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    if r.get("Counter_Name") == counter:
                        rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
        rows.sort()
        per = {}
        for _, name, v in rows:
            per.setdefault(name, []).append(v)
        return per

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    copy_kib = 1 << 20
    cf = [v for k, vs in fetch.items() if "k_stream_copy" in k for v in vs]
    cw = [v for k, vs in write.items() if "k_stream_copy" in k for v in vs]
    fscale, wscale = copy_kib / (sum(cf) / len(cf)), copy_kib / (sum(cw) / len(cw))
    alg = {"k_g2p_fast": 196 * 13434880, "k_p2g_fast": 64 * 13434880, "k_gather_grid": None}
    out = {}
    for key, label in (("k_g2p_fast", "k_g2p_tile"), ("k_p2g_fast", "k_p2g_tile"), ("k_gather_grid", "k_gather_grid")):
        name = next((k for k in fetch if key + "<float, 3" in k), None)
        if name is None or name not in write:
            continue
        fr, wr = fetch[name][1:], write[name][1:]
        rd = sum(fr) / len(fr) * 1024 * fscale
        wb = sum(wr) / len(wr) * 1024 * wscale
        out[label] = {"kernel": name, "dispatches": len(fr), "bytes_read": rd, "bytes_written": wb,
                      "bytes": rd + wb}
    return {"workload": "3D block drop dx=1/256 FP32, steady-state dispatches (first dispatch left out)",
            "sources_digest": sources_digest(), "fetch_scale": fscale, "write_scale": wscale,
            "commands": [f"rocprofv3 --pmc {c} -d {d} -o run --output-format csv -- python3 bench.py --steps 3 "
                         "--warmup 2 --no-cpu --no-adjoint --no-batch --no-sphere --no-taylor --no-g2p2g "
                         "--copy-probe 1024" for c, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir))],
            "kernels": out}


def adjoint_ratio(nsteps, every):
    """Differentiable inverse cube (dx=1/64, 132,651 particles, quadratic), nsteps steps.
    ratio: forward and backward, checkpoint recompute included (the checkpointed forward
    that records the checkpoints, then the pullback of the loss mean ||x(T) - (x(0) +
    (0.1,0,0))||^2: replay of every segment from its checkpoint plus the reversed steps,
    FP64 adjoint) over the plain FP32 forward of the same steps, each the median of three
    rounds.  pullback_over_forward leaves the checkpointed forward out."""
    import numpy as np
    import torch

    import tesserae_hip as th
    from gpu_util import to_gpu
    from tesserae_hip.scenes import inverse_cube

    sc = inverse_cube(dtype=torch.float32)
    ctx = th.Context()
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def timed(fn):
        a, b = ev(), ev()
        torch.cuda.synchronize()
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    target = torch.from_numpy(sc.x + np.array([0.1, 0.0, 0.0])).cuda()
    # warm-up: one short forward and pullback (allocations, code loading)
    mesh, P, G = to_gpu(sc)
    th.rollout(G, P, sc.basis, sc.params, 2, checkpoint_every=every, ctx=ctx)
    adj = th.Adjoint(sc.n, 3)
    th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)
    # three rounds of (plain forward, checkpointed forward, pullback), each from the initial state; the
    # figures are each column's median (the forward's launch-bound steps vary by a few percent per run)
    fwd, roll, pull = [], [], []
    for _ in range(3):
        mesh, P, G = to_gpu(sc)
        fwd.append(timed(lambda: th.step(G, P, sc.basis, sc.params, nsteps=nsteps, ctx=ctx)))
        mesh, P, G = to_gpu(sc)
        roll.append(timed(lambda: th.rollout(G, P, sc.basis, sc.params, nsteps, checkpoint_every=every, ctx=ctx)))
        x = P.get("x").double()
        loss = float(((x - target) ** 2).sum(dim=1).mean())
        adj = th.Adjoint(sc.n, 3)
        adj.x.copy_(2.0 * (x - target) / sc.n)
        pull.append(timed(lambda: th.rollout_adjoint(G, sc.basis, sc.params, adj, ctx=ctx)))
        gnorm = float(adj.v.norm())
    ctx.close()
    t_fwd, t_roll, t_adj = (float(np.median(v)) for v in (fwd, roll, pull))
    return {"workload": f"inverse cube dx=1/64, {sc.n} particles, quadratic, {nsteps} steps, checkpoint every "
                        f"{every}, FP32 forward / FP64 adjoint",
            "t_forward_ms": t_fwd, "t_rollout_ms": t_roll, "t_pullback_ms": t_adj,
            "ratio": (t_roll + t_adj) / t_fwd, "pullback_over_forward": t_adj / t_fwd, "loss": loss,
            "grad_v0_norm": gnorm,
            "rounds_ms": {"forward": fwd, "rollout": roll, "pullback": pull},
            "ratio_per_round": [(r + p) / f for f, r, p in zip(fwd, roll, pull)]}


def batch_leg(steps, reorder):
    """256 uneven 2D scenes (quadratic, FP32) stepped as one batch, after the autotuner
    picks the launch variant; each scene's particles are reordered into the tile order at
    the given threshold."""
    import torch

    import tesserae_hip as th
    from gpu_util import to_gpu
    from tesserae_hip.scenes import batched_scenes

    scenes = batched_scenes(256, dtype=torch.float32)
    members = []
    for sc in scenes:
        mesh, P, G = to_gpu(sc)
        members.append((G, P, sc.params))
    ntot = sum(sc.n for sc in scenes)
    ctx = th.Context()
    batch = th.Batch(members, th.BSpline(th.Quadratic()), ctx=ctx)
    batch.set_reorder(reorder)  # the reference loop's reorder_particles! threshold, per scene
    variant = batch.autotune(4)  # four timed steps per variant: eager and graph mode are close
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times, prof, active_nodes = {}, {}, -1
    # the tuned launch shape both eagerly (one launch per kernel, host-sized) and in graph mode (replays of
    # the captured step of persistent launches); the tuned mode's time is the batch's figure
    for mode, v in (("eager", variant & ~8), ("graph", variant | 8)):
        batch.set_variant(v)
        batch.step(2)  # graph mode: one ordinary step, the capture, one replay
        torch.cuda.synchronize()
        if mode == "eager":
            ctx.profile(True)
        a.record()
        batch.step(steps)
        b.record()
        torch.cuda.synchronize()
        times[mode] = a.elapsed_time(b) / steps
        if mode == "eager":
            prof = ctx.profile_read()
            ctx.profile(False)
            active_nodes = ctx.stats()["active_nodes"]
    batch.set_variant(variant)
    ms = times["graph" if variant & 8 else "eager"]
    bstats = batch.stats()
    bytes_step = 4 * (5 * 4 + 6 * 2 + 2) * ntot + 4 * (8 * 2 + 2) * max(active_nodes, 0)
    batch.close()
    ctx.close()
    return {"workload": f"256 uneven 2D scenes (grid 64-512 cells, 2e3-2e5 particles each, {ntot} particles), "
                        "quadratic, FP32, per-scene dt, one launch per kernel for the batch",
            "particles": ntot, "ms_per_step": ms, "particle_steps_per_s": ntot / (ms * 1e-3),
            "hbm_frac": bytes_step / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), "autotuned_variant": variant,
            "eager_ms_per_step": times["eager"], "graph_ms_per_step": times["graph"],
            "graph_replays": bstats["graph_replays"], "reorders": bstats["reorders"],
            "kernels_avg_us": {k: v[0] / max(v[1], 1) * 1e3 for k, v in prof.items()}}


def sphere_leg(dx, steps):
    """Sparse-grid sphere (SpArray 8^3 blocks): 2x2x2 particles per cell inside r=0.3 at
    dx, rigid rotation plus noise, cubic B-spline and uGIMP, FP32.  Positions are generated
    on the device.  Every step runs the reference loop's reorder_particles!(threshold=0.85);
    the first reorder (out of the generation order) falls in the two warm-up steps."""
    import math

    import torch

    import tesserae_hip as th

    out = {}
    s = dx / 2
    k0, k1 = int(math.floor(0.2 / s - 0.5)), int(math.ceil(0.8 / s - 0.5))
    ax = (torch.arange(k0, k1 + 1, device="cuda", dtype=torch.float64) + 0.5) * s
    parts = []
    for z in ax.split(64):
        X, Y, Z = torch.meshgrid(ax, ax, z, indexing="ij")
        keep = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2 < 0.09
        parts.append(torch.stack([X[keep], Y[keep], Z[keep]], dim=1).float())
    x = torch.cat(parts)
    del parts
    n = x.shape[0]
    for basis in (th.BSpline(th.Cubic()), th.uGIMP()):
        mesh = th.CartesianMesh(dx, (0, 1), (0, 1), (0, 1), dtype=torch.float32, block_size_log2=3)
        P = th.Particles(n, 3, torch.float32, with_l=basis.kind == 16)
        P.x.copy_(x)
        g = torch.Generator(device="cuda")
        g.manual_seed(7)
        v = torch.stack([-(x[:, 1] - 0.5), x[:, 0] - 0.5, torch.zeros_like(x[:, 0])], dim=1)  # ω = 2, times 0.5
        P.v.copy_(v + 0.05 * torch.randn(n, 3, device="cuda", generator=g))
        P.V.fill_(s ** 3)
        P.F.copy_(torch.eye(3, device="cuda").reshape(1, 9).expand(n, 9))
        if basis.kind == 16:
            P.l.fill_(dx / 2)
        params = th.MPMParams(dt=5e-5, E=5e5, nu=0.35, rho0=1e3, V0=s ** 3)
        G = th.generate_grid(th.SpArray, mesh)
        ctx = th.Context()
        # warm-up: storage, code loading, and the first reorder (decided from step 1's count)
        th.step(G, P, basis, params, nsteps=2, reorder_threshold=0.85, ctx=ctx)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        th.step(G, P, basis, params, nsteps=steps, reorder_threshold=0.85, ctx=ctx)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        st = ctx.stats()
        # 260 B per particle (3D FP32), plus l read by both transfers for uGIMP; 104 B per active node
        bytes_step = (260 + (8 if basis.kind == 16 else 0)) * n + 104 * max(st["active_nodes"], 0)
        out[repr(basis)] = {"ms_per_step": ms, "particle_steps_per_s": n / (ms * 1e-3), "reorders": st["reorders"],
                            "hbm_frac": bytes_step / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
                            "storage_nodes": G.spinds.storage_len, "dense_nodes": mesh.nnodes()}
        ctx.close()
        del P, G
        torch.cuda.empty_cache()
    return {"workload": f"sparse-grid sphere dx={dx:g}, {n} particles, SpArray 8^3 blocks, FP32", "particles": n,
            "by_basis": out}


def g2p2g_leg(sc, passes, layout, reorder):
    """@G2P2G of the README program on the block drop, as the time loop of a program written as one @G2P2G
    block (tess_g2p2g_steps, deferred weights, the step loop's reorder threshold): one P2G, warm-up passes, then
    `passes` timed passes in one call, each the G2P half, then the P2G and grid-node halves with the weights at
    the positions before the pass.  Timed as the library runs it (the G2P kernel, then the P2G kernels) and
    as one particle sweep per pass (TESS_G2P2G_SWEEP=1)."""
    import torch

    import tesserae_hip as th
    from gpu_util import to_gpu

    def run(sweep):
        old = os.environ.get("TESS_G2P2G_SWEEP")
        os.environ["TESS_G2P2G_SWEEP"] = "1" if sweep else "0"
        try:
            mesh, P, G = to_gpu(sc, layout=layout)
            ctx = th.Context()
            th.P2G(G, P, sc.basis, sc.params, ctx=ctx)
            th.g2p2g_steps(G, P, sc.basis, sc.params, nsteps=3, reorder_threshold=reorder, ctx=ctx)
            ctx.profile(True)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record()
            th.g2p2g_steps(G, P, sc.basis, sc.params, nsteps=passes, reorder_threshold=reorder, ctx=ctx)
            b.record()
            torch.cuda.synchronize()
            prof = ctx.profile_read()
            ctx.close()
            del P, G
            torch.cuda.empty_cache()
        finally:
            if old is None:
                del os.environ["TESS_G2P2G_SWEEP"]
            else:
                os.environ["TESS_G2P2G_SWEEP"] = old
        return a.elapsed_time(b) / passes, {k: v[0] / max(v[1], 1) * 1e3 for k, v in prof.items()}

    ms, kernels = run(False)
    ms_s, kernels_s = run(True)
    return {"workload": f"block drop dx=1/256, {sc.n} particles, quadratic, FP32, G2P2G passes", "passes": passes,
            "ms_per_pass": ms, "particle_steps_per_s": sc.n / (ms * 1e-3), "kernels_avg_us": kernels,
            "sweep_ms_per_pass": ms_s, "sweep_kernels_avg_us": kernels_s}


def taylor_leg(steps):
    """Taylor-impact rod of the reference's GPU example (docs/src/gpu.md:202-393): 1.48M particles,
    KernelCorrection(BSpline(Quadratic())) with deferred weights, the rod's bottom layer on the floor (truncated
    supports: the corrected branch), FP32 and FP64, next to the plain quadratic B-spline on the same rod.  The
    example's published whole-run times on an RTX 5090 (FP32 8 s, FP64 31 s for ~1.8k iterations) include its
    von Mises return mapping; this engine runs the README's Neo-Hookean solid, so the comparison covers the
    transfers and not the material update."""
    import torch

    import tesserae_hip as th
    from gpu_util import to_gpu
    from tesserae_hip.scenes import taylor_rod

    out = {}
    n = iters = 0
    for dtype, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        for basis in (th.KernelCorrection(th.BSpline(th.Quadratic())), th.BSpline(th.Quadratic())):
            sc = taylor_rod(dtype=dtype, basis=basis)
            n, iters = sc.n, int(round(80e-6 / sc.params.dt))
            mesh, P, G = to_gpu(sc)
            ctx = th.Context()
            th.step(G, P, basis, sc.params, nsteps=3, reorder_threshold=0.85, ctx=ctx)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            a.record()
            th.step(G, P, basis, sc.params, nsteps=steps, reorder_threshold=0.85, ctx=ctx)
            b.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / steps
            out[f"{tag} {basis!r}"] = {"ms_per_step": ms, "particle_steps_per_s": n / (ms * 1e-3),
                                       "projected_run_s": ms * 1e-3 * iters}
            ctx.close()
            del P, G
            torch.cuda.empty_cache()
    return {"workload": f"Taylor-impact rod, {n} particles, mesh spacing R/12, {iters} iterations to 80 us, "
                        "Neo-Hookean, deferred weights", "particles": n, "iterations": iters,
            "reference_rtx5090_run_s": {"f32": 8.0, "f64": 31.0}, "by_config": out}


def main():
    args = parse()
    if args.traffic_from:
        rec = traffic_record(*args.traffic_from)
        with open(TRAFFIC_FILE, "w") as f:
            json.dump(rec, f, indent=1)
        print(json.dumps(rec, indent=1))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import tesserae_hip as th
    from gpu_util import to_gpu
    from tesserae_hip.scenes import block_drop

    torch.cuda.set_device(0)
    if args.copy_probe:
        # the library's plain copy of known size: its FETCH_SIZE / WRITE_SIZE counter readings
        # scale the engine's kernels' readings to bytes
        src = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            th.default_context().device_copy(src, dst)
        torch.cuda.synchronize()
        del src, dst
        log(f"bench: copy probe {args.copy_probe} MiB x3")
    log("bench: building scene")
    sc = block_drop(dtype=torch.float32, dx=args.dx)
    mesh, P, G = to_gpu(sc, layout=args.layout)
    n = sc.n
    ctx = th.default_context()
    torch.cuda.synchronize()

    # the reference loop's maintenance: reorder_particles!(...; threshold=0.85) every step
    def steps(k):
        th.step(G, P, sc.basis, sc.params, nsteps=k, reorder_threshold=args.reorder, ctx=ctx)

    log(f"bench: n={n} warmup={args.warmup}")
    if args.warmup > 0:
        steps(args.warmup)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"bench: timing {args.steps} steps")
    ctx.profile(True)  # per-kernel HIP events on the library's stream
    torch.cuda.synchronize()
    start.record()
    t0 = time.perf_counter()
    steps(args.steps)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    stats = ctx.stats()
    ms = start.elapsed_time(stop) / args.steps
    value = n / (ms * 1e-3)
    # algorithmic bytes per step (README layout, weights evaluated in-kernel):
    # particles s*(5d^2+6d+2), active grid nodes s*(8d+2)
    s, d = 4, 3
    bytes_p = s * (5 * d * d + 6 * d + 2)
    bytes_node = s * (8 * d + 2)
    step_bytes = bytes_p * n + bytes_node * max(stats["active_nodes"], 0)
    kernels = {k: {"avg_us": v[0] / max(v[1], 1) * 1e3, "calls": v[1],
                   "achieved_GBs": (v[2] / v[1]) / (v[0] / v[1] * 1e-3) / 1e9 if v[0] > 0 and v[2] > 0 else None}
               for k, v in prof.items()}
    dom = max((k for k in prof if prof[k][2] > 0), key=lambda k: prof[k][0])
    dom_ms = prof[dom][0] / prof[dom][1]
    dom_bytes = prof[dom][2] / prof[dom][1]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic, traffic_src = profiled_traffic(dom)
    out = {
        "metric": "3D MPM particle-steps/s at 13M pcls, % of HBM peak; adjoint/forward time ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3D elastic block drop, dx=1/256, quadratic B-spline, 13,434,880 particles, FP32",
                   "particles": n, "layout": args.layout},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms},
        "step_hbm": {"algorithmic_bytes_per_step": step_bytes, "achieved_GBs": step_bytes / (ms * 1e-3) / 1e9,
                     "frac": step_bytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), "active_nodes": stats["active_nodes"]},
        "kernels": kernels,
        # binnings of the context so far that kept the previous tile layout (no particle
        # changed its key), and the movers merged at the last incremental re-binning
        "binning": {"kept_layout": stats["kept_binnings"], "movers_last": stats["movers"],
                    "reorders": stats["reorders"]},
        "wall_s": wall,
    }
    if not args.no_g2p2g:
        log("bench: G2P2G passes")
        out["g2p2g"] = g2p2g_leg(sc, 10, args.layout, args.reorder)
    if not args.no_batch:
        log("bench: batched scenes")
        out["batch"] = batch_leg(20, args.reorder)
    if not args.no_sphere:
        log("bench: sparse-grid sphere")
        out["sphere"] = sphere_leg(args.sphere_dx, 3)
    if not args.no_taylor:
        log("bench: Taylor-impact rod (KernelCorrection)")
        out["taylor"] = taylor_leg(20)
    if not args.no_adjoint:
        log("bench: adjoint/forward ratio")
        out["adjoint"] = adjoint_ratio(args.adjoint_steps, args.checkpoint_every)
    out["copy_GBs"] = copy_bandwidth(ctx)
    log(f"bench: device copy {out['copy_GBs']:.0f} GB/s")
    if not args.no_cpu:
        log("bench: CPU baseline")
        del P, G
        torch.cuda.empty_cache()
        nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        # the box's CPU share for one GPU is 16 threads, whatever nproc shows
        threads = args.cpu_threads or min(16, nproc)
        out["cpu_baseline"] = cpu_baseline(threads, torch.float64, args.cpu_steps)
        out["cpu_baseline"]["host"] = {"nproc": nproc, "cpu_count": os.cpu_count(), "cpu_model": cpu_model()}
        out["cpu_baseline_f32"] = cpu_baseline(threads, torch.float32, args.cpu_steps)
        if threads != 8:  # the thread count of the reference's published CPU numbers
            out["cpu_baseline_8_threads"] = cpu_baseline(8, torch.float64, max(2, args.cpu_steps // 2))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
