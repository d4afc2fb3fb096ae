"""Benchmark: H-HIGNN mobility matvec u = M(x) f on one MI355X.

Workload (default): 1,000,000 spheres of radius 1 in 200 Gaussian blobs
(sigma = 15), blob centres uniform in the cube sized for a 10% mean volume
fraction, minimum centre distance 2.05, seed 2; gravity forcing (0,0,-1);
leaf 80 (the workload does not fix it; the chosen form measured at N=1M, s
per matvec: leaf 48 1.453, 64 1.407, 80 1.383, 96 1.405, 128 1.491, 160 1.598,
192 1.716; the FP32 path runs on the same tree), ACA epsilon
1e-4, at most 60 ACA iterations, max_relative_coord 500000 (the reference's
default); pair-function weights: the smooth seeded set of the 3-128-512-9
MLP (hignn_amd.weights.smooth(0)).

A step is one matvec (HignnModel.dot) on a built tree with forces and
velocities resident in device memory.  ``value`` is particle-updates/s of
the chosen form: BF16 MLP on MFMA with FP32 accumulation for the near field,
far-field ACA entries from the learned surrogate (trained on the GPU on 10M
samples labelled by the full MLP before the timed region) for every far
block that passes the full MLP's per-block judge and from the BF16 MLP for
the rest, FP32 ACA factors and arithmetic, FP64 norms and velocity sums.  The same form without the surrogate and the plain FP32 path
are timed in the same run, and all are checked against an FP64 evaluation
of the same block structure (dense far blocks) on a random subsample of
targets.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MLP_FLOP = 141_056  # 2*(3*128 + 128*512 + 512*9) per pair-function evaluation
# HBM bytes per pair-function evaluation (read + written), measured with
# rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this benchmark at N=1M, each
# scaled by the counter's reading for a 1 GiB device copy (x2.0 and x1.0).
# The near kernel's writes are its FP64 velocity atomics as the L2 issues
# them; the source-side ones come once per MLP call of T targets and the
# target sums once per task and block, so they depend on the tile shape and
# the leaf size: at leaf 80, shape 0 (4 tiles per call, 128 workgroups) reads
# 0.75 B and writes 8.35 B per evaluation, shape 1 (2 tiles, 256 workgroups)
# 0.76 B and 14.53 B, shape 2 (2 tiles, 96 workgroups) 0.81 B and 14.88 B
# (at leaf 128: 2 tiles 0.46 B and 13.81 B, 4 tiles 0.63 B and 7.49 B).
# Shape 3 (4 tiles) is taken as shape 0.  The ACA passes of the full MLP
# (leaf 256, FP16, earlier setting) wrote 36 B of entries and read 9.5 B.
NEAR_TRAFFIC_PER_PAIR = {0: 9.10, 1: 15.29, 2: 15.69, 3: 9.10}  # by near shape, leaf 80
NEAR_SHAPE_TILES = {0: 4, 1: 2, 2: 2, 3: 4}  # near kernel shapes (tiles, threads): 0 (4, 512),
                                             # 1 (2, 512), 2 (2, 1024), 3 (4, 256)
TRAFFIC_PER_PAIR = {"far_mlp": 45.6, "near": NEAR_TRAFFIC_PER_PAIR[0]}
PEAK = {"bf16": 2500.0, "fp16": 2500.0, "fp32": 157.3}  # dense MFMA TFLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def work_counts(exp, ranks):
    tree = exp["tree"].astype(np.int64)
    sz = tree[:, 3] - tree[:, 2]
    c, f = exp["close"], exp["far"]
    e_near = int(np.sum(sz[c[:, 0]] * sz[c[:, 1]]))
    e_far = int(np.sum(ranks.astype(np.int64) * (sz[f[:, 0]] + sz[f[:, 1]])))
    return e_near, e_far


def cpu_baseline(x, w, f, leaf, eps, max_iter, mrc, exp, budget_s):
    """Oracle (C++ restatement, all host threads) timed on a bounded sample of
    the same workload: CloseDot on every k-th close block and the full ACA of
    every k-th far block, extrapolated to the whole matvec by pair-function
    evaluation counts."""
    from oracle import oracle as O
    om = O.OracleModel(x, leaf, epsilon=eps, max_iter=max_iter, max_relative_coord=mrc)
    om.update_coord()
    nclose, nfar = len(exp["close"]), len(exp["far"])
    # start small, grow the sample until it costs a few seconds
    cs, fs = max(1, nclose // 8), max(1, nfar // 8)
    t, wn, wf = om.sample_time(w, f, cs, fs)
    target = budget_s
    while t < 0.2 * target and (cs > 1 or fs > 1):
        scale = max(2, int(target / max(t, 1e-3) * 0.5))
        cs, fs = max(1, cs // scale), max(1, fs // scale)
        t, wn, wf = om.sample_time(w, f, cs, fs)
    return t, wn, wf, cs, fs


THREE_BODY_FLOP = {"three_body": 2 * (6 * 256 + 256 * 1024 + 1024 * 512 + 512 * 9),
                   "self": 2 * (3 * 256 + 256 * 1024 + 1024 * 512 + 512 * 9)}


def three_body_bench(args):
    """velocity_update_without_2body (python/gravity_field.py:34-59) on its
    own workload: a 50^3 lattice of spacing 3 (python/gravity_field.py:85-95),
    three-body radius 5, force (0,0,-1); seeded Xavier nets of the shapes of
    HIGNN_nn_3body.pkl / HIGNN_nn_self.pkl.  A step is one evaluation: lists
    (neighbours, triplets) and both nets over every edge, velocities summed."""
    from hignn_amd import hignn, weights

    hignn.Init()
    m = args.lattice
    g = np.arange(m, dtype=np.float64) * 3.0
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3).astype(np.float32)
    n = x.shape[0]
    f = np.zeros((n, 3), np.float32)
    f[:, 2] = -1.0
    w3, ws = weights.xavier(1, weights.net_shapes(6)), weights.xavier(2, weights.net_shapes(3))
    out = {}
    for prec in ("fp16", "fp32"):
        tb = hignn.ThreeBodyModel()
        tb.load_weights(w3, ws)
        tb.set_three_body_epsilon(5.0)
        tb.set_precision(prec)
        steps = args.steps if prec == args.prec else 1
        for _ in range(args.warmup):
            tb.velocity(x, f)
        t0 = time.perf_counter()
        for _ in range(steps):
            v = tb.velocity(x, f)
        el = (time.perf_counter() - t0) / steps
        prof = tb.last_profile()
        out[prec] = (el, prof, v, steps)
        log(f"[three-body] {prec}: {1e3 * el:.1f} ms per evaluation, {prof}")
    el, prof, v, steps = out[args.prec]
    e3, es = prof["three_body_edges"], prof["self_edges"]
    fl3, fls = THREE_BODY_FLOP["three_body"] * e3, THREE_BODY_FLOP["self"] * es
    mlp_ms = prof["three_body_ms"] + prof["self_ms"]
    achieved = (fl3 + fls) / (mlp_ms * 1e-3) / 1e12
    v32 = out["fp32"][2]
    rel = float(np.linalg.norm(v - v32) / np.linalg.norm(v32))
    cpu = None
    if not args.no_cpu:
        from oracle import three_body as OT
        k = 12  # a 12^3 sub-lattice: same neighbourhoods, ~1/70 of the edges
        sub = x.reshape(m, m, m, 3)[:k, :k, :k].reshape(-1, 3)
        t0 = time.perf_counter()
        OT.velocity(sub, f[:sub.shape[0]], w3, ws, 5.0)
        ct = time.perf_counter() - t0
        lists, _ = OT.neighbour_lists(sub, 5.0)
        trip, self_e = OT.three_body_edges(lists)
        ce = trip.shape[1] + self_e.shape[1]
        cpu = {"value": round(ce / ct, 1), "unit": "edges/s", "cores": os.cpu_count() or 1,
               "kind": "port",
               "sample": f"oracle (numpy float64, BLAS threads) on the 12^3 corner of the lattice: "
                         f"{ce} edges in {ct:.1f}s"}
    line = {
        "metric": "three-body + self corrections: edges/s (gravity_field lattice)",
        "value": round((e3 + es) / el, 1),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.prec,
        "data": "synthetic",
        "config": {
            "workload": f"gravity_field lattice {m}^3 (N={n}), spacing 3, three-body radius 5, "
                        "force (0,0,-1); Xavier nets (seeds 1, 2)",
            "three_body_edges": e3, "self_edges": es,
            "mlp_ms": round(mlp_ms, 3),
            "lists_and_copies_ms": round(1e3 * el - mlp_ms, 3),
            "fp32_ms_per_step": round(1e3 * out["fp32"][0], 3),
            "rel_l2_vs_fp32_path": rel,
        },
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2),
                     "peak": PEAK[args.prec], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK[args.prec], 4), "traffic": None,
                     "kernel": "layer1 + gemm_act (x2) + layer4 over all edges"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def stepper_bench(args):
    """Integrator steps on the reference's filament_cloud configuration
    (python/config/filament_cloud.json: 1024 filaments x 21 beads, rest length
    3, k_t 20, k_b 50, body force (0,0,-1), Explicit Euler dt 0.005; model
    settings of python/simulate.py:41-51).  Each step is one full velocity
    evaluation (forcing, update_coord with a tree rebuild, matvec) plus the
    position update.  Each model first times its far-field launch choices on
    the forces at the initial positions and keeps the fastest (untimed;
    --no-tune-far keeps the automatic choice).  Timed four ways on the same
    model settings: the Python
    velocity callback over host arrays (as the reference's scripts drive the
    integrator), the device-resident integrator with per-stage HIP graphs,
    the same without graphs, and the device-resident integrator with the tree
    kept across steps (opt-in)."""
    from hignn_amd import abi, hignn, weights, workloads

    hignn.Init()
    nchain, rest, k_t, k_b, dt = 21, 3.0, 20.0, 50.0, 0.005
    x = workloads.filament_cloud(args.filaments, nchain, rest, 240.0, seed=0)
    n = x.shape[0]
    w = weights.xavier(0)
    prec = args.prec
    tuning = [None]  # rows (mode, chunk, matvec ms) of the last model's far-field tuning

    def setup(reuse):
        m = hignn.HignnModel(x, 100 if n > 100 else n // 2)
        m.load_two_body_weights(w)
        m.set_epsilon(0.01)
        m.set_max_iter(50)
        m.set_mat_pool_size_factor(200)
        m.set_post_check_flag(False)
        m.set_use_symmetry_flag(True)
        m.set_max_far_dot_work_node_size(10000)
        m.set_max_relative_coord(100000)
        m.set_precision(prec)
        m.set_tree_reuse(reuse)
        fc = hignn.DeviceForcing()
        fc.set_body_force([0.0, 0.0, -1.0])
        fc.set_filaments(args.filaments, nchain, rest, k_t, k_b)
        if not args.no_tune_far:  # far-field launch choice, timed on the first forces
            m.update_coord(x)
            f0 = abi.DeviceArray.of(fc.eval(x))
            tuning[0] = m.tune_far_field(f0.ptr, 0)
        return m, fc

    def timed(mode, steps):
        m, fc = setup(mode == "device_reuse")
        it = hignn.ExplicitEuler()
        it.set_time_step(dt)
        it.set_num_rigid_body(n)
        if mode == "callback":
            def vel(t, pos):
                xx = np.ascontiguousarray(pos[:, :3])
                m.update_coord(xx)
                f = fc.eval(xx)
                u = np.zeros_like(f)
                m.dot(u, f)
                return u
            it.set_velocity_func(vel)
        else:
            it.set_device_velocity(m, fc)
            it.set_graphs(mode != "device_eager")
            if mode == "device_reuse":
                it.set_tree_rebuild(args.rebuild_every)
        state = np.zeros((n, 6), np.float32)
        state[:, :3] = x
        res = []
        for k, s in ((0, args.warmup), (1, steps)):
            it.set_final_time(dt * max(s, 1) * (1 - 1e-3))
            it.initialize(state)
            t0 = time.perf_counter()
            it.run()
            el = time.perf_counter() - t0
            res.append((it.func_count(), el))
        cnt, el = res[1]
        whole = it.device_stats()["whole_stage_graphs"] if mode != "callback" else 0
        log(f"[stepper] {mode}: {cnt} steps in {el:.3f}s ({1e3 * el / cnt:.2f} ms/step, "
            f"{whole} whole-stage graph launches)")
        return cnt, el, it.positions(), whole

    out = {}
    for mode in ("callback", "device_graphs", "device_eager", "device_reuse"):
        out[mode] = timed(mode, args.steps)
    xs = {k: v[2] for k, v in out.items()}
    diff = float(np.abs(xs["device_graphs"] - xs["callback"]).max())
    diff_reuse = float(np.abs(xs["device_reuse"] - xs["device_graphs"]).max())
    moved = float(np.abs(xs["device_graphs"] - x).max())
    rate = {k: v[0] / v[1] for k, v in out.items()}
    steps, el = out["device_graphs"][0], out["device_graphs"][1]
    line = {
        "metric": "Explicit Euler steps/s, filament_cloud suspension (device-resident integrator)",
        "value": round(rate["device_graphs"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic",
        "config": {
            "workload": f"filament_cloud: {args.filaments} straight filaments x {nchain} beads "
                        f"(N={n}), rest 3, k_t 20, k_b 50, body force (0,0,-1), dt 0.005; leaf "
                        "100, eps 0.01, max_iter 50, max_relative_coord 100000; Xavier(seed 0)",
            "callback_steps_per_s": round(rate["callback"], 3),
            "device_eager_steps_per_s": round(rate["device_eager"], 3),
            "device_tree_reuse_steps_per_s": round(rate["device_reuse"], 3),
            "device_tree_reuse_rebuild_every": args.rebuild_every,
            "device_tree_reuse_max_abs_position_diff_vs_rebuilt": diff_reuse,
            "max_displacement": moved,
            "device_tree_reuse_whole_stage_graph_launches": out["device_reuse"][3],
            "speedup_vs_callback": round(rate["device_graphs"] / rate["callback"], 3),
            "far_field_tuning": ("off" if args.no_tune_far else
                                 [{"mode": ("persistent", "batched")[r[0]], "chunk": r[1],
                                   "matvec_ms": round(r[2], 3)} for r in tuning[0]]),
            "max_abs_position_diff_vs_callback": diff,
        },
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--prec", default="bf16", choices=["bf16", "fp16"])
    ap.add_argument("--no-surrogate", action="store_true",
                    help="far entries from the full MLP (no learned surrogate)")
    ap.add_argument("--sur-samples", type=int, default=10_000_000)
    ap.add_argument("--sur-steps", type=int, default=60000)
    ap.add_argument("--sur-rlo", type=float, default=2.0)
    ap.add_argument("--sur-rhi", type=float, default=650.0)
    ap.add_argument("--sur-tol", type=float, default=2e-3)
    ap.add_argument("--nosur-steps", type=int, default=1,
                    help="steps of the same 16-bit form without the surrogate (0: skip)")
    ap.add_argument("--check-targets", type=int, default=2000)
    ap.add_argument("--fp32-steps", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--max-iter", type=int, default=60)
    ap.add_argument("--leaf", type=int, default=80)
    ap.add_argument("--max-relative-coord", type=int, default=500000)
    ap.add_argument("--epsilon", type=float, default=1e-4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="one device-to-device copy of this size before the timed region "
                         "(known-size reference for HBM byte counters)")
    ap.add_argument("--near-alone-steps", type=int, default=2,
                    help="matvecs with the near field alone (not beside the far field), after "
                         "the timed steps, for the near kernel's own rate")
    ap.add_argument("--surrogate-file", default="",
                    help="matvec bench: load the surrogate's parameters from this .npy if it "
                         "exists, else train and save them there")
    ap.add_argument("--no-tune-near", action="store_true",
                    help="matvec bench: keep the default tile shape and workgroup count of the "
                         "16-bit near kernel")
    ap.add_argument("--tune-far", action="store_true",
                    help="matvec bench: also time the far field's launch choices (6 matvecs)")
    ap.add_argument("--no-tune-far", action="store_true",
                    help="stepper: keep the automatic far-field choice instead of timing the "
                         "candidates on the first forces")
    ap.add_argument("--stepper", action="store_true",
                    help="time Explicit Euler steps of the filament_cloud suspension instead "
                         "(Python velocity callback vs the device-resident integrator)")
    ap.add_argument("--rebuild-every", type=int, default=10,
                    help="stepper, tree kept across stages: full rebuild every this many steps")
    ap.add_argument("--filaments", type=int, default=1024,
                    help="filaments of 21 beads in the --stepper workload")
    ap.add_argument("--three-body", action="store_true",
                    help="time the three-body and self corrections on the gravity_field lattice "
                         "instead")
    ap.add_argument("--workload", choices=["c2", "c4"], default=None,
                    help="the full-size C2 (FP32 path + DOPRI steps) or C4 (filaments + surrogate) "
                         "configuration instead of the headline matvec")
    ap.add_argument("--c2-steps", type=int, default=8,
                    help="accepted DOPRI steps of the C2 workload (toward its final time 0.2; "
                         "~45 s each at N=100k under the reference's step control)")
    ap.add_argument("--instances", type=int, default=0,
                    help="time the many-instance workload with this many suspensions (64) "
                         "instead")
    ap.add_argument("--lattice", type=int, default=50,
                    help="lattice points per side in the --three-body workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are "
              "not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.stepper:
        return stepper_bench(args)
    if args.three_body:
        return three_body_bench(args)
    if args.instances:
        return instances_bench(args)
    if args.workload:
        return workload_bench(args)

    return matvec_bench(args)


def workload_bench(args):
    """The full-size C2 and C4 configurations (hignn_amd.workloads.CONFIGS).

    c2: 100,000 spheres uniform in a cube at 10% volume fraction (seed 1),
    gravity, plain FP32 path (leaf 256, eps 1e-4, max rank 32, smooth seeded
    weights): one matvec timed after a warm-up and checked against the FP64
    evaluation at --check-targets targets, then --c2-steps steps of the
    reference's "Rk4" (Dormand-Prince 5(4), dt 0.002, error threshold 1e-3)
    with the device-resident integrator (forcing, tree rebuild and matvec per
    stage on the device); value = accepted steps per second.

    c4: 2,000 filaments of 50 beads (bond 2.2, 8%, seed 3), bonded Morse
    (D 1, alpha 1, r0 2.2) plus gravity evaluated on the device, leaf 256,
    eps 1e-4, max_iter 32; the surrogate trained on 10M displacements with
    |r| log-uniform in [4, 200]; value = particle-updates/s of the BF16 +
    surrogate matvec, with BF16 alone and the FP32 path timed beside it and
    every form checked against the FP64 evaluation.

    The reference's step control (include/TimeIntegrator.hpp:612-840) starts
    every step at dt and compares the unnormalised L2 norm of the embedded
    error over all particles with 1e-3.  At N = 100,000 the accepted steps
    are ~1e-5 long, each after several attempts (~45 s per accepted step), so
    reaching the final time 0.2 would take ~2e4 steps: the run stops after
    --c2-steps accepted steps."""
    from hignn_amd import abi, hignn, weights, workloads
    from oracle import surrogate as S

    name = {"c2": "c2_uniform_100k", "c4": "c4_filaments_100k"}[args.workload]
    cfg = workloads.CONFIGS[name]
    t0 = time.time()
    x = workloads.positions(name)
    n = x.shape[0]
    log(f"[{args.workload}] positions N={n} in {time.time() - t0:.1f}s")
    w = weights.smooth(0)
    hignn.Init()
    if args.workload == "c4":
        fc = hignn.DeviceForcing()
        fc.set_body_force([0.0, 0.0, -1.0])
        fc.set_bonds(cfg["nfil"], cfg["nbead"], 1.0, 1.0, 2.2)
        f = fc.eval(x)
    else:
        f = workloads.gravity(n)
    m = abi.Model(x, cfg["leaf"])
    m.load_weights(w)
    m.set(epsilon=cfg["epsilon"], max_iter=cfg["max_iter"])
    t_update, _ = time_update_coord(m)
    counts = m.counts()
    f_dev = abi.DeviceArray.of(f)
    u_dev = abi.DeviceArray((n, 3))
    sur = None

    def run(prec, use_sur):
        m.set(precision=prec)
        m.set_surrogate(sur if use_sur else None, tol=args.sur_tol)
        abi.synchronize()
        t = time.perf_counter()
        m.dot_device(u_dev.ptr, f_dev.ptr)
        abi.synchronize()
        return time.perf_counter() - t

    g = np.random.default_rng(7)
    tg = np.sort(g.choice(n, size=min(args.check_targets, n), replace=False))
    t0 = time.time()
    ref = m.check_fp64(f, tg, 0)
    t_chk = time.time() - t0

    def rel(u):
        return float(np.sqrt(np.sum((u[tg] - ref) ** 2) / np.sum(ref ** 2)))

    cfg_line = {"workload": "", "n": n, "leaf": cfg["leaf"], "epsilon": cfg["epsilon"],
                "max_iter": cfg["max_iter"], "tree_rows": counts[0], "close_pairs": counts[1],
                "far_blocks": counts[2], "update_coord_s": round(t_update, 3),
                "check_targets": int(tg.size), "check_s": round(t_chk, 1)}
    if args.workload == "c2":
        run("fp32", False)
        t32 = min(run("fp32", False) for _ in range(max(1, args.fp32_steps)))
        e32 = rel(u_dev.numpy())
        log(f"[c2] fp32 matvec {t32:.3f}s, error vs FP64 {e32:.2e}")
        hm = hignn.HignnModel(x, cfg["leaf"])
        hm.load_two_body_weights(w)
        hm.set_epsilon(cfg["epsilon"])
        hm.set_max_iter(cfg["max_iter"])
        fg = hignn.DeviceForcing()
        fg.set_body_force([0.0, 0.0, -1.0])
        it = hignn.ExplicitRk4()
        it.set_time_step(0.002)
        it.set_final_time(0.2)  # 100 steps of dt 0.002
        it.set_max_steps(args.c2_steps)
        it.set_num_rigid_body(n)
        it.set_device_velocity(hm, fg)
        state = np.zeros((n, 6), np.float32)
        state[:, :3] = x
        it.initialize(state)
        t_start = [0.0]
        accepted = [0]

        def progress(pos, vel):  # one line per accepted step
            accepted[0] += 1
            log(f"[c2] accepted step {accepted[0]}: {time.perf_counter() - t_start[0]:.1f}s")

        it.set_output_step(1)
        it.set_output_func(progress)
        abi.synchronize()
        t0 = t_start[0] = time.perf_counter()
        it.run()
        abi.synchronize()
        t_run = time.perf_counter() - t0
        st = it.device_stats()
        moved = float(np.abs(it.positions() - x).max())
        log(f"[c2] {it.func_count()} DOPRI steps in {t_run:.1f}s: {st}, moved up to {moved:.3f}")
        cfg_line.update({
            "workload": f"uniform cube N={n}, 10% volume fraction, seed 1, gravity; FP32 path; "
                        f"the first {args.c2_steps} accepted steps of Dormand-Prince 5(4) (the "
                        "reference's Rk4: dt 0.002 per step, final time 0.2, threshold 1e-3), "
                        "device-resident integrator with a tree rebuild per stage",
            "fp32_matvec_s": round(t32, 4), "rel_err_fp64_fp32_path": e32,
            "dopri_steps": it.func_count(), "velocity_evaluations": st["velocity_evaluations"],
            "rejected": st["rejected"], "integration_s": round(t_run, 2),
            "max_displacement": round(moved, 4)})
        line = {"metric": "C2: FP32-path DOPRI steps/s at N=100k (device-resident integrator)",
                "value": round(it.func_count() / t_run, 4), "unit": "steps/s", "n_gpus": 1,
                "steps": it.func_count(), "warmup": 0,
                "ms_per_step": round(1000.0 * t_run / max(1, it.func_count()), 2),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
                "data": "synthetic", "config": cfg_line}
    else:
        t0 = time.time()
        r_s, t_s = m.surrogate_samples(10_000_000, 4.0, 200.0, seed=1)
        sur = abi.Surrogate(4.0, 200.0)
        sur.set_params(S.init(0))
        loss = sur.train(r_s.ptr, t_s.ptr, 10_000_000, args.sur_steps, 3e-3, batch=16384,
                         lr_final=3e-5)
        abi.synchronize()
        del r_s, t_s
        t_train = time.time() - t0
        log(f"[c4] surrogate trained in {t_train:.1f}s (loss {loss:.3e})")
        near_tuning = []
        if not args.no_tune_near:  # tile shape and workgroups of the near kernel, untimed
            m.set(precision=args.prec)
            m.set_surrogate(sur, tol=args.sur_tol)
            near_tuning = m.tune_near(f_dev.ptr)
            log(f"[c4] near-kernel launches (shape, tiles, threads, workgroups, matvec ms): "
                f"{near_tuning}")
        times = {}
        errs = {}
        for form, prec, use_sur, k in (("bf16+surrogate", args.prec, True, args.steps),
                                       ("bf16", args.prec, False, 1),
                                       ("fp32", "fp32", False, max(1, args.fp32_steps))):
            run(prec, use_sur)
            times[form] = min(run(prec, use_sur) for _ in range(k))
            errs[form] = rel(u_dev.numpy())
            if use_sur:
                stats = m.surrogate_stats()
                ranks = m.far_ranks()
            log(f"[c4] {form}: {times[form]:.3f}s, error vs FP64 {errs[form]:.2e}")
        cfg_line.update({
            "workload": f"{cfg['nfil']} filaments x {cfg['nbead']} beads (N={n}), bond 2.2, 8%, "
                        "seed 3; bonded Morse D 1, alpha 1, r0 2.2 plus gravity (device forcing); "
                        "surrogate 3-32-32-6 SiLU trained on 10M samples, |r| in [4, 200]",
            "surrogate_train_s": round(t_train, 2), "surrogate_loss": loss,
            "surrogate_stats": dict(zip(["far_blocks_judged", "on_surrogate", "fell_back",
                                         "surrogate_entries", "full_mlp_far_entries"], stats)),
            "mean_far_rank": float(ranks.mean()) if ranks.size else None,
            "matvec_s": {k: round(v, 4) for k, v in times.items()},
            "near_launch": (dict(zip(["shape", "tiles", "threads", "workgroups", "ms"],
                                     min(near_tuning, key=lambda r: r[4])))
                            if near_tuning else None),
            "rel_err_fp64": errs,
            "speedup_vs_fp32": round(times["fp32"] / times["bf16+surrogate"], 3)})
        line = {"metric": "C4: BF16+surrogate matvec particle-updates/s at N=100k filaments",
                "value": round(n / times["bf16+surrogate"], 2), "unit": "particle-updates/s",
                "n_gpus": 1, "steps": args.steps, "warmup": 1,
                "ms_per_step": round(1000.0 * times["bf16+surrogate"], 2),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": args.prec, "data": "synthetic", "config": cfg_line}
    print(json.dumps(line), flush=True)


def instances_bench(args):
    """Many independent suspensions (the many-instance workload): 64 uniform
    cubes with N log-uniform in [5,000, 200,000] and volume fraction uniform
    in [5%, 15%] (seed 4), gravity, in one model: per-instance trees and
    lists, one near-field launch and one far field over all instances per
    matvec.  Chosen form as the 1M benchmark (BF16 MLP + trained surrogate);
    value = aggregate particle-updates/s; per-instance relative error against
    the FP64 evaluation of the same blocks."""
    from hignn_amd import abi, weights, workloads
    from oracle import surrogate as S

    t0 = time.time()
    x, ns = workloads.instances(count=args.instances)
    n = x.shape[0]
    log(f"[instances] {len(ns)} instances, N {ns.min()}..{ns.max()}, total {n} in "
        f"{time.time() - t0:.1f}s")
    f = workloads.gravity(n)
    w = weights.smooth(0)
    abi.lib().hh_init()
    m = abi.Model(x, args.leaf, instances=ns)
    m.load_weights(w)
    m.set(epsilon=args.epsilon, max_iter=args.max_iter,
          max_relative_coord=args.max_relative_coord)
    t_update, t_update_host = time_update_coord(m)
    counts = m.counts()
    log(f"[instances] update_coord {t_update:.3f}s (host build {t_update_host:.3f}s): {counts}")
    t0 = time.time()
    r_s, t_s = m.surrogate_samples(args.sur_samples, args.sur_rlo, args.sur_rhi, seed=1)
    sur = abi.Surrogate(args.sur_rlo, args.sur_rhi)
    sur.set_params(S.init(0))
    loss1 = sur.train(r_s.ptr, t_s.ptr, args.sur_samples, args.sur_steps, 3e-3, batch=16384,
                      lr_final=3e-5)
    abi.synchronize()
    del r_s, t_s
    t_train = time.time() - t0
    f_dev = abi.DeviceArray.of(f)
    u_dev = abi.DeviceArray((n, 3))

    def run(prec, use_sur):
        m.set(precision=prec)
        m.set_surrogate(sur if use_sur else None, tol=args.sur_tol)
        abi.synchronize()
        t = time.perf_counter()
        m.dot_device(u_dev.ptr, f_dev.ptr)
        abi.synchronize()
        return time.perf_counter() - t

    near_tuning = []
    if not args.no_tune_near:  # tile shape and workgroups of the near kernel, untimed
        m.set(precision=args.prec)
        m.set_surrogate(sur, tol=args.sur_tol)
        near_tuning = m.tune_near(f_dev.ptr)
        log(f"[instances] near-kernel launches (shape, tiles, threads, workgroups, matvec ms): "
            f"{near_tuning}")
    for i in range(args.warmup):
        log(f"[instances] warmup {i}: {run(args.prec, True):.3f}s")
    abi.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        dt = run(args.prec, True)
        log(f"[instances] step {i}: {dt:.3f}s")
    t_total = time.perf_counter() - t_start
    u = u_dev.numpy()
    stats = m.surrogate_stats()
    t32 = None
    if args.fp32_steps > 0:
        t32 = min(run("fp32", False) for _ in range(args.fp32_steps))
        log(f"[instances] fp32: {t32:.3f}s")
    off = np.concatenate([[0], np.cumsum(ns)])
    g = np.random.default_rng(5)
    tg = np.concatenate([np.sort(g.choice(np.arange(off[k], off[k + 1]), size=min(150, ns[k]),
                                          replace=False)) for k in range(len(ns))])
    t0 = time.time()
    ref = m.check_fp64(f, tg, 0)
    errs = []
    for k in range(len(ns)):
        sel = (tg >= off[k]) & (tg < off[k + 1])
        d = u[tg[sel]] - ref[sel]
        errs.append(float(np.sqrt(np.sum(d ** 2) / np.sum(ref[sel] ** 2))))
    log(f"[instances] fp64 check ({time.time() - t0:.1f}s): max error {max(errs):.2e}")
    ms = 1000.0 * t_total / args.steps
    value = n / (t_total / args.steps)
    line = {
        "metric": "many-instance matvec: aggregate particle-updates/s (64 suspensions, ~4.3M "
                  "particles), per-instance rel err <= 1e-3",
        "value": round(value, 2), "unit": "particle-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.prec,
        "data": "synthetic",
        "config": {
            "workload": f"{len(ns)} uniform cubes, N log-uniform in [5000, 200000] (total {n}), "
                        f"volume fraction uniform in [5%, 15%], seed 4, gravity; leaf {args.leaf}, "
                        f"eps {args.epsilon}, max_iter {args.max_iter}; smooth seeded weights",
            "instances": len(ns), "n_min": int(ns.min()), "n_max": int(ns.max()),
            "tree_rows": counts[0], "close_pairs": counts[1], "far_blocks": counts[2],
            "update_coord_s": round(t_update, 3),
            "update_coord_host_build_s": round(t_update_host, 3),
            "surrogate_train_s": round(t_train, 2), "surrogate_loss": loss1,
            "surrogate_stats": dict(zip(["far_blocks_judged", "on_surrogate", "fell_back",
                                         "surrogate_entries", "full_mlp_far_entries"], stats)),
            "per_instance_rel_err_fp64_max": max(errs),
            "per_instance_rel_err_fp64_median": float(np.median(errs)),
            "per_instance_rel_err_fp64": [round(e, 8) for e in errs],
            "check_targets_per_instance": 150,
            "near_launch": (dict(zip(["shape", "tiles", "threads", "workgroups", "ms"],
                                     min(near_tuning, key=lambda r: r[4])))
                            if near_tuning else None),
            "fp32_value": round(n / t32, 2) if t32 else None,
            "speedup_vs_fp32": round(value / (n / t32), 3) if t32 else None,
        },
    }
    print(json.dumps(line), flush=True)


def host_info():
    """Host cores available to this process and the CPU model."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"affinity_cores": cores, "cpu_count": os.cpu_count(), "cpu_model": model}


def time_update_coord(model):
    """update_coord (tree + lists + work descriptors) from host coordinates
    with the device build (warm: the second call) and with the host build;
    the model ends on the device build."""
    model.set(tree_build="host")
    t0 = time.time()
    model.update_coord()
    t_host = time.time() - t0
    model.set(tree_build="device")
    model.update_coord()
    t0 = time.time()
    model.update_coord()
    return time.time() - t0, t_host


def matvec_bench(args):
    from hignn_amd import abi, weights, workloads
    from oracle import surrogate as S

    t0 = time.time()
    x = workloads.positions("c3_blobs_1m", n=args.n if args.n != 1_000_000 else None)
    n = x.shape[0]
    log(f"[bench] positions N={n} in {time.time() - t0:.1f}s")
    f = workloads.gravity(n)
    w = weights.smooth(0)
    leaf, eps, max_iter, mrc = args.leaf, args.epsilon, args.max_iter, args.max_relative_coord

    abi.lib().hh_init()
    model = abi.Model(x, leaf)
    model.load_weights(w)
    model.set(epsilon=eps, max_iter=max_iter, max_relative_coord=mrc)
    t_update, t_update_host = time_update_coord(model)
    exp = model.export()
    log(f"[bench] update_coord {t_update:.3f}s (host build {t_update_host:.3f}s): "
        f"{len(exp['close'])} close, {len(exp['far'])} far")
    f_dev = abi.DeviceArray.of(f)
    u_dev = abi.DeviceArray((n, 3))

    # the learned part: a surrogate of this pair function, trained on the GPU
    # by the library's own kernels on samples labelled by the full MLP (setup,
    # once per pair function; not part of a matvec)
    sur, sur_info = None, None
    if not args.no_surrogate and args.surrogate_file and os.path.exists(args.surrogate_file):
        # parameters of an earlier run of this benchmark (profiling runs skip
        # the training's dispatches)
        sur = abi.Surrogate(args.sur_rlo, args.sur_rhi)
        sur.set_params(np.load(args.surrogate_file, allow_pickle=False))
        sur_info = {"net": "3-32-32-6 SiLU", "loaded_from": args.surrogate_file,
                    "r_range": [args.sur_rlo, args.sur_rhi], "judge_tol": args.sur_tol,
                    "judge_samples": 32}
        log(f"[bench] surrogate: {sur_info}")
    elif not args.no_surrogate:
        t0 = time.time()
        r_s, t_s = model.surrogate_samples(args.sur_samples, args.sur_rlo, args.sur_rhi, seed=1)
        sur = abi.Surrogate(args.sur_rlo, args.sur_rhi)
        sur.set_params(S.init(0))
        loss0 = sur.train(r_s.ptr, t_s.ptr, args.sur_samples, 0, 1e-3)
        loss1 = sur.train(r_s.ptr, t_s.ptr, args.sur_samples, args.sur_steps, 3e-3,
                          batch=16384, lr_final=3e-5)
        abi.synchronize()
        del r_s, t_s
        sur_info = {"net": "3-32-32-6 SiLU", "samples": args.sur_samples,
                    "r_range": [args.sur_rlo, args.sur_rhi], "adam_steps": args.sur_steps,
                    "batch": 16384, "loss_initial": loss0, "loss_final": loss1,
                    "train_s": round(time.time() - t0, 2), "judge_tol": args.sur_tol,
                    "judge_samples": 32}
        log(f"[bench] surrogate: {sur_info}")
        if args.surrogate_file:
            np.save(args.surrogate_file, sur.params())

    def run(prec, use_sur):
        model.set(precision=prec)
        model.set_surrogate(sur if use_sur else None, tol=args.sur_tol)
        abi.synchronize()
        t = time.perf_counter()
        model.dot_device(u_dev.ptr, f_dev.ptr)
        abi.synchronize()
        return time.perf_counter() - t

    if args.calibrate_copy_mib > 0:
        # a device-to-device copy of known size on the compute queue (a HIP
        # copy kernel), the reference reading for HBM byte counters
        nb = args.calibrate_copy_mib << 20
        src = abi.DeviceArray((nb // 4,), np.ones(nb // 4, np.float32))
        dst = abi.DeviceArray((nb // 4,))
        abi.lib().hh_device_copy(dst.p, src.p, nb)
        abi.synchronize()
        log(f"[bench] calibration copy: {nb} bytes read, {nb} bytes written")
        del src, dst

    use_sur = sur is not None
    # tile shape and workgroups of the 16-bit near kernel (and, with --tune-far, the far
    # field's launch choices) tuned on these forces, untimed
    model.set(precision=args.prec)
    model.set_surrogate(sur if use_sur else None, tol=args.sur_tol)
    near_tuning = model.tune_near(f_dev.ptr) if not args.no_tune_near else []
    log(f"[bench] near-kernel launches (shape, tiles, threads, workgroups, matvec ms): {near_tuning}")
    near_best = min(near_tuning, key=lambda r: r[4]) if near_tuning else None
    near_shape = (int(near_best[0]) if near_tuning else int(os.environ.get("HH_NEAR_SHAPE", "0")))
    # workgroups of the persistent near kernel (0: one per compute unit) and
    # the compute units that the tuner's first candidate covered
    near_wg = (int(near_best[3]) if near_tuning
               else int(os.environ.get("HH_NEAR_WORKGROUPS", "0")))
    n_cus = int(near_tuning[0][3]) if near_tuning else None
    TRAFFIC_PER_PAIR["near"] = NEAR_TRAFFIC_PER_PAIR[near_shape]
    far_tuning = model.tune_far_field(f_dev.ptr) if args.tune_far else []
    if far_tuning:
        log(f"[bench] far-field launch choices (mode, chunk, matvec ms): {far_tuning}")
    # chosen form: warmup, then K timed steps
    for i in range(args.warmup):
        dt = run(args.prec, use_sur)
        log(f"[bench] warmup {i} {args.prec}: {dt:.3f}s")
    times, tcl, tfar, profs = [], [], [], []
    abi.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        dt = run(args.prec, use_sur)
        c, fa, _ = model.timings()
        times.append(dt)
        tcl.append(c)
        tfar.append(fa)
        profs.append(model.profile())
        log(f"[bench] step {i} {args.prec}: {dt:.3f}s (close {c:.1f} ms, far {fa:.1f} ms)")
    abi.synchronize()
    t_total = time.perf_counter() - t_start
    u_chosen = u_dev.numpy()
    ranks = model.far_ranks()
    e_near, e_far = work_counts(exp, ranks)
    stats = model.surrogate_stats() if use_sur else None

    # the near kernel by itself: matvecs with the near field first and alone
    # (the far field after it), outside the timed steps, in which the far
    # iterations share the near kernel's compute units
    near_alone = None
    if args.near_alone_steps > 0:
        model.set(near_overlap=False)
        model.set_near_workgroups(0)  # alone, the kernel gets every compute unit
        run(args.prec, use_sur)
        pa = []
        for _ in range(args.near_alone_steps):
            run(args.prec, use_sur)
            pa.append(model.profile())
        model.set(near_overlap=True)
        model.set_near_workgroups(near_wg)
        ms_a = sum(float(p[0][0]) for p in pa) / len(pa)
        ev_a = sum(int(p[1][0]) for p in pa) / len(pa)
        if ev_a > 0:
            near_alone = (ms_a, MLP_FLOP * ev_a / (ms_a / 1000.0) / 1e12)
            log(f"[bench] near kernel alone: {ms_a:.1f} ms, {near_alone[1]:.1f} TFLOP/s")

    # the same 16-bit MLP without the surrogate (every far entry from the full
    # MLP), one step: what the learned part gains
    u_nosur = t_nosur = ranks_nosur = None
    if use_sur and args.nosur_steps > 0:
        t_nosur = min(run(args.prec, False) for _ in range(args.nosur_steps))
        u_nosur = u_dev.numpy()
        ranks_nosur = model.far_ranks()
        log(f"[bench] {args.prec} without surrogate: {t_nosur:.3f}s")

    # plain FP32 path in the same run: one warm-up, median of fp32_steps
    f32_times = []
    if args.fp32_steps > 0:
        dt = run("fp32", False)
        log(f"[bench] fp32 warmup: {dt:.3f}s")
    for i in range(args.fp32_steps):
        dt = run("fp32", False)
        f32_times.append(dt)
        c, fa, _ = model.timings()
        log(f"[bench] fp32 step {i}: {dt:.3f}s (close {c:.1f} ms, far {fa:.1f} ms)")
    u_fp32 = u_dev.numpy() if f32_times else None

    # accuracy: FP64 evaluation of the same blocks at random targets
    err = err32 = err_nosur = None
    if args.check_targets > 0:
        g = np.random.default_rng(7)
        tg = np.sort(g.choice(n, size=min(args.check_targets, n), replace=False))
        t0 = time.time()
        ref = model.check_fp64(f, tg, 0)
        t_chk = time.time() - t0

        def rel(u):
            if u is None:
                return None
            return float(np.sqrt(np.sum((u[tg] - ref) ** 2) / np.sum(ref ** 2)))

        err, err32, err_nosur = rel(u_chosen), rel(u_fp32), rel(u_nosur)
        log(f"[bench] fp64 check on {tg.size} targets ({t_chk:.1f}s): chosen err {err}, "
            f"without surrogate {err_nosur}, fp32 err {err32}")

    ms = 1000.0 * t_total / args.steps
    value = n / (t_total / args.steps)
    fp32_rate = n / float(np.median(f32_times)) if f32_times else None
    # roofline of the dominant pair-function kernel, from HIP events on the
    # model's stream around each launch (near kernel; ACA column/row passes)
    kern = {"near": ("near_h_kernel", 0), "far_mlp": ("aca_mlp_h_kernel", 1)}
    tot = {k: sum(float(p[0][i]) for p in profs) for k, (_, i) in kern.items()}
    dom = max(tot, key=tot.get)
    i_dom = kern[dom][1]
    ms_k = tot[dom]
    pairs_k = sum(int(p[1][i_dom]) for p in profs)
    launches_k = sum(int(p[2][i_dom]) for p in profs)
    achieved = MLP_FLOP * pairs_k / (ms_k / 1000.0) / 1e12

    def rate(k):
        i = kern[k][1]
        ev = sum(int(p[1][i]) for p in profs)
        return round(MLP_FLOP * ev / max(1e-9, tot[k] / 1000.0) / 1e12, 2) if ev else None

    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK[args.prec],
            "unit": "TFLOP/s", "frac": round(achieved / PEAK[args.prec], 4),
            "traffic": int(TRAFFIC_PER_PAIR[dom] * pairs_k / max(1, launches_k)),
            "kernel": f"{kern[dom][0]}<{args.prec}>"
                      + (f" shape {near_shape} ({NEAR_SHAPE_TILES[near_shape]} tiles)"
                         + (f" on {near_wg} workgroups" if near_wg else "") if dom == "near" else ""),
            "flop_per_launch": int(MLP_FLOP * pairs_k / max(1, launches_k)),
            "avg_launch_ms": round(ms_k / max(1, launches_k), 3),
            "launches_per_step": int(launches_k / args.steps),
            "near_kernel_tflops": rate("near"), "far_mlp_tflops": rate("far_mlp")}
    if dom == "near" and near_wg and n_cus:
        # the near kernel ran on near_wg of the n_cus compute units (one
        # workgroup per compute unit); the far kernels held the others
        roof.update({"near_workgroups": near_wg, "compute_units": n_cus,
                     "frac_of_its_compute_units": round(achieved / (PEAK[args.prec] * min(1.0, near_wg / n_cus)), 4)})
    if near_alone is not None:
        roof.update({"near_alone_ms": round(near_alone[0], 3),
                     "near_alone_tflops": round(near_alone[1], 2),
                     "near_alone_frac": round(near_alone[1] / PEAK[args.prec], 4)})

    cpu = None
    if not args.no_cpu:
        try:
            t, wn, wf, cs, fs = cpu_baseline(x, w, f, leaf, eps, max_iter, mrc, exp,
                                             args.cpu_budget)
            ranks32 = ranks if not f32_times else model.far_ranks()
            e_n32, e_f32 = work_counts(exp, ranks32)
            t_full = t * (e_n32 + e_f32) / max(1, wn + wf)
            hi = host_info()
            cpu = {"value": round(n / t_full, 2), "unit": "particle-updates/s",
                   "cores": hi["affinity_cores"], "kind": "port",
                   "cpu_model": hi["cpu_model"], "cpu_count": hi["cpu_count"],
                   "sample": f"oracle (FP32 MLP, FP64 ACA, all {hi['affinity_cores']} threads of "
                             f"the process's affinity set) on CloseDot of every {cs}th close "
                             f"block + the ACA of every {fs}th far block ({wn + wf} pair "
                             f"evaluations, {t:.1f}s), extrapolated by evaluation count to the "
                             f"full plain-path matvec ({e_n32 + e_f32})"}
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:  # the baseline must not sink the GPU result
            log(f"[bench] cpu baseline failed: {e}")

    far_mlp_ms = tot["far_mlp"] / args.steps
    form = (f"{args.prec} MLP on MFMA (FP32 accumulate) for the near field"
            + (" and the far blocks the judge rejects; far ACA entries of the other blocks from "
               "the FP32 surrogate (3-32-32-6 SiLU)" if use_sur else "; far entries from the full MLP")
            + ("; ACA factors stored in FP32" if use_sur else "; ACA factors stored in BF16")
            + " (FP32 arithmetic), FP64 ACA norms and velocity sums")
    line = {
        "metric": "matvec particle-updates/s @N=1M: BF16+surrogate vs plain FP32, rel err<=1e-3",
        "value": round(value, 2),
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.prec,
        "data": "synthetic",
        "config": {
            "workload": f"gaussian blobs N={n} (200 blobs, sigma 15, 10% mean volume fraction, "
                        f"seed 2), gravity; leaf {leaf}, eps {eps}, max_iter {max_iter}, "
                        f"max_relative_coord {mrc}; smooth seeded weights (Xavier seed 0, "
                        f"layer-1 weights x 0.01)",
            "chosen_form": form,
            "surrogate": sur_info,
            "surrogate_stats": (dict(zip(["far_blocks_judged", "on_surrogate", "fell_back",
                                          "surrogate_entries", "full_mlp_far_entries"], stats))
                                if stats else None),
            "fp32_value": round(fp32_rate, 2) if fp32_rate else None,
            "fp32_ms_median": round(1e3 * float(np.median(f32_times)), 1) if f32_times else None,
            "fp32_steps": len(f32_times),
            "speedup_vs_fp32": round(value / fp32_rate, 3) if fp32_rate else None,
            f"{args.prec}_without_surrogate_value": round(n / t_nosur, 2) if t_nosur else None,
            f"{args.prec}_without_surrogate_mean_far_rank": (float(ranks_nosur.mean())
                                                             if ranks_nosur is not None else None),
            "rel_err_fp64": err, "rel_err_fp64_without_surrogate": err_nosur,
            "rel_err_fp64_fp32_path": err32,
            "check_targets": args.check_targets,
            "e_near": e_near, "e_far": e_far,
            "mean_far_rank": float(ranks.mean()) if ranks.size else 0.0,
            "update_coord_s": round(t_update, 3),
            "update_coord_host_build_s": round(t_update_host, 3),
            "near_launch_tuning": [{"shape": a, "tiles": b, "threads": c, "workgroups": d,
                                    "ms": round(e, 1)} for a, b, c, d, e in near_tuning],
            "far_launch_tuning": [{"mode": a, "chunk": b, "ms": round(c, 1)} for a, b, c in far_tuning],
            "close_ms": round(float(np.median(tcl)), 2), "far_ms": round(float(np.median(tfar)), 2),
            "far_mlp_ms": round(far_mlp_ms, 2),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
