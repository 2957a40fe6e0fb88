#!/usr/bin/env python3
"""Benchmark: points/s end-to-end at depth 10 on one MI355X.

Workload: 2,000,000 float32 points on the union of 3 seeded superquadrics
(noise sigma 0.001, analytic normals), octree depth 10, FP32 CG to 1e-6
relative residual with FP64 dot products.  A step is the reference's whole
main() minus file I/O: the points are uploaded from (pinned) host memory,
normalised, the octree, divergence, per-depth solves, isovalue, leaf-level
marching cubes and the coarse-leaf subdivision batches run, and the mesh lands
in (pinned) host memory.  Beside it the line reports the same step from
device-resident points, the leaf-level mesh alone, and one run with file I/O
(a .bnpts read and the ASCII PLY write).

Prints one JSON line (rank 0).  One GPU only: --gpus N with N > 1 exits.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
SPMV_ROW_BYTES = 116.0  # per row, reference representation (int32 neighbour indices, FP32 values)


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipHostFree.argtypes = [C.c_void_p]
    return hip


class PinnedMesh:
    """Pinned host buffers reused by every step for the mesh copy."""

    def __init__(self, hip, nverts, ntris):
        self.hip = hip
        self.vcap, self.tcap = 3 * nverts, 3 * ntris
        self.v, self.t = C.c_void_p(), C.c_void_p()
        assert hip.hipHostMalloc(C.byref(self.v), 4 * self.vcap, 0) == 0
        assert hip.hipHostMalloc(C.byref(self.t), 4 * self.tcap, 0) == 0

    def free(self):
        self.hip.hipHostFree(self.v)
        self.hip.hipHostFree(self.t)


def workload(args):
    import poissonrecon_gpu_amd as prg

    if args.workload == "superquadrics_2M_d10":
        p, n = prg.synth_superquadrics(args.points, 3, 0.001, 3)
        return p, n, 10
    if args.workload == "superquadrics_150k_d10":  # sparse: large coarse leaves the surface crosses
        p, n = prg.synth_superquadrics(args.points if args.points != 2_000_000 else 150_000, 3, 0.001, 21)
        return p, n, 10
    if args.workload == "torus_500k_d9":
        p, n = prg.synth_torus(args.points, 2)
        return p, n, 9
    if args.workload == "sphere_36k_d8":
        p, n = prg.synth_sphere(args.points, 1)
        return p, n, 8
    raise SystemExit(f"unknown workload {args.workload}")


def run_step(g, p, nrm, n, on_device, pinned=None, subdivide=1):
    """One pass of the pipeline: points (device or host pointers) to the mesh in host memory."""
    import poissonrecon_gpu_amd as prg

    g._call("prg_set_points", p, nrm, n, int(on_device))
    g.build_octree()
    g.compute_divergence()
    stats = g.solve()
    g._call("prg_isovalue", None)  # the value stays on the device for the mesh
    if pinned is None:
        mesh = g.extract_mesh(subdivide=bool(subdivide))
        return stats, (len(mesh.vertices), len(mesh.triangles))
    m = prg.PrgMesh()
    g._call("prg_extract_mesh_into", int(subdivide), pinned.v, pinned.vcap, pinned.t, pinned.tcap, C.byref(m))
    return stats, (m.n_vertices, m.n_triangles)


def timed_steps(hip, g, p, nrm, n, on_device, pinned, subdivide, steps):
    """Seconds per step over `steps` steps (device synchronised on both sides)."""
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = run_step(g, p, nrm, n, on_device, pinned, subdivide)
    hip.hipDeviceSynchronize()
    return (time.perf_counter() - t0) / steps, out


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, n, D, threads, min_root_depth):
    """The CPU oracle (restatement of the reference stages) on the whole cloud:
    normalisation through the leaf-level mesh and the subdivision batches of the
    roots at depth >= min_root_depth (the shallower groups would need the
    reference's explicit virtual trees, hundreds of GB at depth 10)."""
    import oracle

    oracle.set_num_threads(threads)
    o = oracle.Oracle(D)
    t = {}
    t0 = time.perf_counter()
    o.set_points(p, n)
    o.build_octree()
    t["octree"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    o.compute_divergence()
    t["divergence"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    o.solve(mode=0)
    t["solve"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    o.isovalue()
    vl, tl = o.extract_leaf_mesh()
    t["leaf_mesh"] = time.perf_counter() - t1
    t1 = time.perf_counter()
    v, tr = o.extract_subdivision(min_root_depth)
    t["subdivision"] = time.perf_counter() - t1
    dt = time.perf_counter() - t0
    return o, dt, {k: round(x, 2) for k, x in t.items()}, (len(v), len(tr))


def fp64_errors(o, gt, counts, bases):
    """Per-depth error of the GPU solve against the oracle's FP64 solve of the
    same system (the oracle's divergence injected into the GPU context)."""
    import poissonrecon_gpu_amd as prg

    b = o.read_buffer(prg.BUF_DIVERGENCE)
    gt.write_buffer(prg.BUF_DIVERGENCE, b)
    stats = gt.solve()
    x = gt.read_buffer(prg.BUF_SOLUTION).astype(np.float64)
    x64 = o.solve_fp64(tol=1e-12)
    out = []
    for d, s in enumerate(stats):
        sl = slice(int(bases[d]), int(bases[d] + counts[d]))
        den = np.linalg.norm(x64[sl])
        out.append({"depth": d, "iters": s["iters"], "true_rel_residual": float(f"{s['true_rel_residual']:.3e}"),
                    "rel_l2_vs_fp64": float(f"{np.linalg.norm(x[sl] - x64[sl]) / den:.3e}") if den > 0 else 0.0})
    return out


def learned_leg(args):
    """Learned warm start vs plain CG at equal residual (1e-6 relative at every depth).

    Training and held-out shapes: unions of 1-4 seeded superquadrics
    (seeds 1000 + k), 200,000 points each, depth 9.  The network works on the
    64^3 grid of depth 6; its training pairs (divergence, solution) come from
    the solver itself.  Solve ms is the solve stage's wall time, including the
    network's inference for the learned runs."""
    import poissonrecon_gpu_amd as prg

    D, gd = 9, 6
    ntr, nte = args.learn_train, args.learn_test
    g = prg.Reconstructor(max_depth=D)
    B, M, X, S = [], [], [], []
    t0 = time.perf_counter()
    for k in range(ntr + nte):
        p, n = prg.synth_superquadrics(args.learn_points, 1 + k % 4, 0.001, 1000 + k)
        g.set_points(p, n)
        g.build_octree()
        g.compute_divergence()
        g.solve()
        b, rms = g.dense_grid(gd, 0)
        B.append(b)
        M.append(g.dense_grid(gd, 2)[0])
        X.append(g.dense_grid(gd, 1)[0])
        S.append(rms)
    g.close()
    t_data = time.perf_counter() - t0
    B, M, X, S = np.stack(B), np.stack(M), np.stack(X), np.array(S, np.float32)
    net = prg.UNet(grid=1 << gd, seed=1)
    t0 = time.perf_counter()
    rng = np.random.default_rng(0)
    losses = []
    for ep in range(args.learn_epochs):
        order = rng.permutation(ntr)
        lr = args.learn_lr * (0.5 ** (ep // max(1, args.learn_epochs // 3)))
        el = []
        for i in range(0, ntr, args.learn_batch):
            idx = order[i:i + args.learn_batch]
            el.append(net.train_step(B[idx], M[idx], X[idx], S[idx], lr))
        losses.append(float(np.mean(el)))
    t_train = time.perf_counter() - t0
    # held-out: plain and learned solves of the same systems
    plain = prg.Reconstructor(max_depth=D)
    warm = prg.Reconstructor(max_depth=D, warm_start=prg.WARM_LEARNED)
    warm.set_warm_start_net(net)
    res = {"plain": [], "learned": []}
    for k in range(ntr, ntr + nte):
        p, n = prg.synth_superquadrics(args.learn_points, 1 + k % 4, 0.001, 1000 + k)
        for name, r in (("plain", plain), ("learned", warm)):
            r.set_points(p, n)
            r.build_octree()
            r.compute_divergence()
            r.solve()  # warm-up of this shape's buffers
            t_before = r.stage_times()["solve"]  # stage times add up from set_points on
            st = r.solve()
            res[name].append((r.stage_times()["solve"] - t_before, st[gd]["iters"], sum(s["iters"] for s in st),
                              max(s["true_rel_residual"] for s in st)))
    plain.close()
    warm.close()
    # held-out prediction error at depth 6 (relative L2 on the nodes)
    errs = []
    for k in range(ntr, ntr + nte):
        pred = net.predict(B[k], M[k], float(S[k]))
        m = M[k] > 0
        errs.append(float(np.linalg.norm((pred - X[k])[m]) / max(np.linalg.norm(X[k][m]), 1e-30)))
    net.close()
    mp = np.array(res["plain"])
    ml = np.array(res["learned"])
    return {
        "protocol": f"{ntr} training + {nte} held-out unions of 1-4 superquadrics, {args.learn_points} points, "
                    f"depth {D}, network on the 64^3 grid of depth {gd}, {args.learn_epochs} epochs, batch "
                    f"{args.learn_batch}, Adam lr {args.learn_lr}",
        "solve_ms_plain": round(float(mp[:, 0].mean()), 3),
        "solve_ms_learned": round(float(ml[:, 0].mean()), 3),
        "speedup": round(float(mp[:, 0].mean() / ml[:, 0].mean()), 4),
        "depth6_iters_plain": round(float(mp[:, 1].mean()), 2),
        "depth6_iters_learned": round(float(ml[:, 1].mean()), 2),
        "total_iters_plain": round(float(mp[:, 2].mean()), 2),
        "total_iters_learned": round(float(ml[:, 2].mean()), 2),
        "max_true_rel_residual_plain": float(mp[:, 3].max()),
        "max_true_rel_residual_learned": float(ml[:, 3].max()),
        "heldout_depth6_prediction_rel_l2": round(float(np.mean(errs)), 4),
        "train_loss_by_epoch": [round(x, 5) for x in losses],
        "data_s": round(t_data, 1),
        "train_s": round(t_train, 1),
    }


def batch_leg(args):
    """Many clouds in one pass: 64 clouds, sizes log-uniform in [10k, 4M]
    (seed 5), unions of 1-4 superquadrics, depth clamp(7 + round(log2(N/10k)/2),
    7, 11) capped at 10 (the deepest this build supports)."""
    import poissonrecon_gpu_amd as prg

    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(4e6), args.batch_clouds)).astype(np.int64)
    clouds, depths, capped = [], [], 0
    for k, n in enumerate(sizes):
        clouds.append(prg.synth_superquadrics(int(n), 1 + k % 4, 0.001, 5000 + k))
        d = int(np.clip(7 + round(np.log2(n / 1e4) / 2), 7, 11))
        capped += d > 10
        depths.append(min(d, 10))
    prg.reconstruct_batch(clouds[:2], depths[:2], concurrency=2)  # warm-up
    t0 = time.perf_counter()
    st, res = prg.reconstruct_batch(clouds, depths, concurrency=args.batch_workers)
    dt = time.perf_counter() - t0
    total = int(sizes.sum())
    return {"clouds": len(sizes), "points": total, "workers": args.batch_workers, "wall_s": round(dt, 3),
            "points_per_s": round(total / dt, 1), "status": int(st),
            "depths": {str(d): depths.count(d) for d in sorted(set(depths))}, "depth11_capped_to_10": int(capped),
            "max_true_rel_residual": max(r["max_true_rel_residual"] for r in res),
            "note": "host points (the PCIe copy of each cloud is inside the timed region)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="superquadrics_2M_d10")
    ap.add_argument("--points", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline and the FP64 error leg")
    ap.add_argument("--cpu-min-root-depth", type=int, default=8,
                    help="CPU baseline: subdivision batches of roots at this depth or deeper")
    ap.add_argument("--no-io", action="store_true", help="skip the run with file I/O")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--learn-train", type=int, default=512)
    ap.add_argument("--learn-test", type=int, default=64)
    ap.add_argument("--learn-points", type=int, default=200_000)
    ap.add_argument("--learn-epochs", type=int, default=6)
    ap.add_argument("--learn-batch", type=int, default=8)
    ap.add_argument("--learn-lr", type=float, default=1e-3)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-cloud batch workload")
    ap.add_argument("--batch-clouds", type=int, default=64)
    ap.add_argument("--batch-workers", type=int, default=4)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with a 512 MiB device-to-device copy (a copy kernel of known size for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import poissonrecon_gpu_amd as prg

    p, nrm, D = workload(args)
    N = len(p)
    hip = _hip()
    # inputs: pinned host copies (the headline step uploads them) and device copies
    hp, hn, dp, dn = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipHostMalloc(C.byref(hp), p.nbytes, 0) == 0 and hip.hipHostMalloc(C.byref(hn), nrm.nbytes, 0) == 0
    C.memmove(hp, p.ctypes.data, p.nbytes)
    C.memmove(hn, nrm.ctypes.data, nrm.nbytes)
    assert hip.hipMalloc(C.byref(dp), p.nbytes) == 0 and hip.hipMalloc(C.byref(dn), nrm.nbytes) == 0
    assert hip.hipMemcpy(dp, p.ctypes.data, p.nbytes, 1) == 0
    assert hip.hipMemcpy(dn, nrm.ctypes.data, nrm.nbytes, 1) == 0

    g = prg.Reconstructor(max_depth=D)
    _, (nv, nt) = run_step(g, hp, hn, N, False)  # sizes the pinned mesh buffers
    pinned = PinnedMesh(hip, int(nv * 1.25) + 1024, int(nt * 1.25) + 1024)
    for _ in range(max(0, args.warmup - 1)):
        run_step(g, hp, hn, N, False, pinned)
    # headline: host points -> full mesh (leaf level + subdivision) in host memory
    dt, (stats, mesh) = timed_steps(hip, g, hp, hn, N, False, pinned, 1, args.steps)
    stage_ms = {k: round(v, 3) for k, v in g.stage_times().items()}
    roots, vleaves = C.c_int64(), C.c_int64()
    g._call("prg_subdivision_info", C.byref(roots), C.byref(vleaves))
    bricks = g.subdivision_bricks()
    counts, bases = g.level_info()
    # the same step from device-resident points, and the leaf-level mesh alone
    dt_dev, _ = timed_steps(hip, g, dp, dn, N, True, pinned, 1, args.steps)
    run_step(g, dp, dn, N, True, pinned, 0)
    dt_leaf, (_, mesh_leaf) = timed_steps(hip, g, dp, dn, N, True, pinned, 0, args.steps)
    variants = {
        "device_points_full_mesh": {"ms_per_step": round(1000 * dt_dev, 3), "points_per_s": round(N / dt_dev, 1)},
        "device_points_leaf_mesh_only": {"ms_per_step": round(1000 * dt_leaf, 3), "points_per_s": round(N / dt_leaf, 1),
                                         "mesh": {"vertices": int(mesh_leaf[0]), "triangles": int(mesh_leaf[1])}},
    }
    g.close()
    if not args.no_io:
        import tempfile

        with tempfile.TemporaryDirectory() as td:
            src, dst = Path(td) / "cloud.bnpts", Path(td) / "mesh.ply"
            np.concatenate([p, nrm], axis=1).astype(np.float32).tofile(src)
            with prg.Reconstructor(max_depth=D) as gi:
                t0 = time.perf_counter()
                q, qn = prg.read_points(str(src))
                gi.set_points(q, qn)
                gi.build_octree()
                gi.compute_divergence()
                gi.solve()
                gi.isovalue()
                m = gi.extract_mesh(subdivide=True)
                prg.write_mesh_ply(str(dst), m)
                t_io = time.perf_counter() - t0
            variants["with_file_io"] = {"s": round(t_io, 3), "points_per_s": round(N / t_io, 1),
                                        "ply_bytes": dst.stat().st_size,
                                        "note": "one run: .bnpts read, pipeline, ASCII PLY write"}

    # one more pass with per-kernel hipEvents (on the context stream) for the roofline
    gt = prg.Reconstructor(max_depth=D, timing=True)
    run_step(gt, dp, dn, N, True)  # sizes this context's buffers; its first-touch costs are not recorded
    gt.kernel_stats_reset()
    run_step(gt, dp, dn, N, True)
    kern = {}
    for name in ["stencil_spmv_dot_finest", "stencil_spmv_dot", "cg_update", "cg_idle", "cg_true_residual", "cg_refresh",
                 "cg_small", "cg_init", "vector_field", "div_subtree_sums", "div_subtree_finish",
                 "div_from_subtree_sums", "div_leaf", "parent_neighbor_table", "radix_downsweep", "radix_upsweep",
                 "point_values", "corner_value", "neighbors", "levels_fill", "edge_owner", "corner_owner",
                 "scan_exclusive", "gather_points", "morton", "point_weights", "spmv_tile_plan", "div_parent_sums",
                 "face_marks", "coarse_corner_value", "subdivide_screen", "subdivide_count", "subdivide_vertices",
                 "subdivide_triangles"]:
        try:
            cnt, ms, bpl = gt.kernel_stats(name)
            kern[name] = {"launches": cnt, "total_ms": round(ms, 4), "avg_us": round(1000 * ms / cnt, 2),
                          "alg_bytes_per_launch": bpl}
        except prg.PrgError:
            pass
    hip.hipFree(dp)
    hip.hipFree(dn)
    if args.calibrate_copy:
        a, b = C.c_void_p(), C.c_void_p()
        size = 512 << 20
        assert hip.hipMalloc(C.byref(a), size) == 0 and hip.hipMalloc(C.byref(b), size) == 0
        assert hip.hipMemcpy(b, a, size, 3) == 0  # device to device
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)

    # Dominant kernel of the solve: the CG SpMV at the finest depth (k_cg_spmv;
    # one launch per CG iteration over all M_D rows).  Its algorithmic bytes in
    # this build's layout (the tile plan: 27 int16 row ids per sibling block, a
    # count per tile and 4 B per listed block; r_i and p_{k-1,i} read, p_{k,i}
    # and q_i written) give `achieved`; the reference's CSR representation
    # (116 B/row) is reported beside it as an equivalent rate, and `traffic` is
    # the measured HBM bytes per launch (FETCH_SIZE x2 + WRITE_SIZE, calibrated
    # on a copy kernel).
    dom = "stencil_spmv_dot_finest"
    k = kern[dom]
    rows = int(counts[D])
    sec = k["avg_us"] * 1e-6
    achieved = k["alg_bytes_per_launch"] / sec / 1e9
    roofline = {"kernel": dom, "hip_kernel": "k_cg_spmv", "rows_per_launch": rows, "bound": "hbm",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "alg_bytes_per_launch": k["alg_bytes_per_launch"],
                "alg_bytes_per_row": round(k["alg_bytes_per_launch"] / rows, 3),
                "reference_representation_rate_gbs": round(rows * SPMV_ROW_BYTES / sec / 1e9, 1),
                "avg_us": k["avg_us"]}
    traffic_file = ROOT / "profiles" / f"traffic_k_cg_spmv_{args.workload}.json"
    if traffic_file.exists():
        t = json.loads(traffic_file.read_text())
        roofline["traffic"] = t.get("hbm_bytes_per_launch")
        roofline["traffic_source"] = str(traffic_file.relative_to(ROOT))
        if roofline["traffic"]:
            roofline["measured_traffic_rate_gbs"] = round(roofline["traffic"] / sec / 1e9, 1)
            roofline["measured_traffic_frac"] = round(roofline["traffic"] / sec / 1e9 / HBM_PEAK_GBS, 4)

    base = None
    solve_fp64 = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        o, t_cpu, cpu_stages, cpu_mesh = cpu_baseline(p, nrm, D, threads, args.cpu_min_root_depth)
        solve_fp64 = fp64_errors(o, gt, counts, bases)
        o.close()
        base = {"value": round(N / t_cpu, 1), "unit": "points/s", "cores": threads, "kind": "port",
                "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                "sample": f"the whole {N}-point cloud at depth {D}: octree, divergence, FP32 CG, isovalue, leaf-level "
                          f"mesh and the subdivision batches of roots at depth >= {args.cpu_min_root_depth} "
                          f"({t_cpu:.1f} s on {threads} OpenMP threads)",
                "stage_s": cpu_stages, "mesh": {"vertices": cpu_mesh[0], "triangles": cpu_mesh[1]}}
        # one-thread run of the 35,947-point sphere at depth 8, whole mesh
        ps, ns = prg.synth_sphere(35947, 1)
        o1, t1, st1, m1 = cpu_baseline(ps, ns, 8, 1, 0)
        o1.close()
        base["one_thread_sphere_36k_d8"] = {"s": round(t1, 2), "points_per_s": round(35947 / t1, 1),
                                            "stage_s": st1, "mesh": {"vertices": m1[0], "triangles": m1[1]}}
    gt.close()
    pinned.free()
    hip.hipHostFree(hp)
    hip.hipHostFree(hn)

    learned = None if args.no_learned else learned_leg(args)
    batch = None if args.no_batch else batch_leg(args)

    out = {
        "metric": "points/s end-to-end at depth 10; solve ms learned vs plain CG at equal residual",
        "value": round(N / dt, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * dt, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 reductions)",
        "data": "synthetic",
        "config": {"workload": args.workload, "points": N, "max_depth": D, "nodes": int(counts.sum()),
                   "leaf_nodes": int(counts[-1]), "cg_tol": 1e-6, "parallelism": "single GPU",
                   "step": "host points uploaded, leaf-level mesh and coarse-leaf subdivision, mesh in host memory"},
        "stage_ms": stage_ms,
        "solve_ms": stage_ms["solve"],
        "cg_iters": [s["iters"] for s in stats],
        "solve_vs_fp64_oracle": solve_fp64,
        "mesh": {"vertices": int(mesh[0]), "triangles": int(mesh[1]), "subdivision_roots": roots.value,
                 "virtual_leaves": vleaves.value,
                 "subdivision_bricks": {"bricks": bricks[0], "left_by_interval_certificate": bricks[1],
                                        "evaluated": bricks[2]}},
        "variants": variants,
        "kernels": kern,
        "roofline": roofline,
        "cpu_baseline": base,
        "learned_warm_start": learned,
        "batch": batch,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
