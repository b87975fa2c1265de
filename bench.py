"""Benchmark: tet-substeps/s of the XPBD substep on one MI355X.

Workload: the 96x24x24-cell Yeoh beam (331,776 linear tets, 60,625 nodes),
1/60 s frames of 30 substeps, ground plane and kinematic sphere, inputs
resident in HBM. One step = one frame (30 substeps, captured as one hipGraph).
Schedule: sweep blocks by default (4 projection launches per substep, each
workgroup sweeping one column block through its phases in LDS; the launch
shape picked by the autotuner), or per-colour launches (--schedule slabs /
plain); the per-colour XCD-slab frame time is reported beside the blocks.

Prints one JSON line: value = whole-job tet-substeps/s, the roofline of the
projection kernels (algorithmic bytes per launch / measured launch time,
HIP events on the library stream), and the reference CPU path timed on a
bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "xpbd-fem_amd"), os.path.join(ROOT, "oracle")]

import fitting  # noqa: E402
import workloads  # noqa: E402
import xpbdfem as xf  # noqa: E402

METRIC = "tet-substeps/s plain vs batched at ≤1e-4 rel pos err; adjoint:forward time"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# FP32 VALU operations of one tet projection (YeohSkinFast, simultaneous
# solve; add/sub, mul and transcendental count one, FMA two): the
# SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F32 counters of a sweep-block launch per
# working wave and phase, 182 + 260 + 2 x 28 + 5, the same for the exact and
# the contracted kernels (profiles/xf_project_blocks_valu_flops.md).
FP32_FLOP_PER_TET = 503
# Algorithmic bytes per tet projection with the reference's precision (FP64
# positions, FP32 inverse mass and constants): node ids 16 + constants (Qi 36,
# QQ 12, QR 12, material 4) 64 + gather 4 x (xyz FP64 24 + w 4) 112 +
# scatter 4 x xyz FP64 96.
BYTES_PER_TET = 16 + 64 + 112 + 96
# Per node and substep (predict + collisions + velocity): read X, V; write X, V (FP64).
BYTES_PER_NODE = 96
# Sweep-block launches: per tet its constants 64 + block-local node indices 8
# (+ comp1 / a / body 16 with per-tet data); per block node its id 4 + FP64
# position and inverse mass read 32 + written 32 (the block's nodes stay in
# LDS through its phases).
BYTES_PER_BLOCK_TET = 64 + 8
BYTES_PER_BLOCK_NODE = 4 + 32 + 32
# fused node phase: O read and written (32 + 32)
BYTES_FUSED_NODE = 32 + 32
# Scene options of the sweep schedules: sweep blocks (4 launches per substep,
# each workgroup sweeping one block through its phases), or one launch per
# colour phase with the XCD-slab order or the plain blockIdx-linear order.
SCHEDULES = {"blocks": dict(sweep_blocks=True, launch_mode=xf.LAUNCH_BLOCKS), "slabs": dict(xcd_slabs=True),
             "plain": {}}
MODE_NAMES = {xf.LAUNCH_PER_COLOUR: "per-colour", xf.LAUNCH_PERSISTENT: "persistent", xf.LAUNCH_BLOCKS: "blocks"}
KERNELS = {xf.LAUNCH_PER_COLOUR: "xf_project_colour", xf.LAUNCH_PERSISTENT: "xf_frame_persistent",
           xf.LAUNCH_BLOCKS: "xf_project_blocks"}


def fp32_valu(tet_projections, ms):
    """Algorithmic FP32 VALU rate of `tet_projections` tet projections done
    in `ms` milliseconds of projection launches, against the FP32 peak."""
    tf = FP32_FLOP_PER_TET * tet_projections / (ms * 1e-3) / 1e12
    return {"achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP32_PEAK_TFLOPS,
            "flop_per_tet": FP32_FLOP_PER_TET}


def projection_bytes(info, mode, substeps, per_tet_data=False):
    """Algorithmic bytes of the projection launches of `substeps` substeps
    (the persistent launch also holds the node phases; with the fused node
    phase, the last block launch of every substep but the call's last reads
    and writes each node's O, 64 B a node; the X of the few nodes outside its
    blocks is not counted)."""
    T, N = info["n_tets"], info["n_nodes"]
    if mode == xf.LAUNCH_BLOCKS:
        per_tet = BYTES_PER_BLOCK_TET + (16 if per_tet_data else 0)
        fused = (substeps - 1) * BYTES_FUSED_NODE * N if info.get("fused_finish") else 0
        return substeps * (per_tet * T + BYTES_PER_BLOCK_NODE * info["n_block_nodes"]) + fused
    return substeps * (BYTES_PER_TET * T + (BYTES_PER_NODE * N if mode == xf.LAUNCH_PERSISTENT else 0))


def rel_pos_err(x, x_ref, x0):
    """max_i |x_i - x_ref,i|_2 / bounding-box diagonal of the rest positions."""
    d = np.linalg.norm((x - x_ref).reshape(-1, 3), axis=1).max()
    p0 = x0.reshape(-1, 3).astype(np.float64)
    return float(d / np.linalg.norm(p0.max(axis=0) - p0.min(axis=0)))


def host_threads():
    """Host threads for the all-core baselines: the CPUs this process may
    use, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(ctx, wl, budget_s=7.0, opts=None):
    """The reference's own Geo3d::Substep (oracle/_ref, compiled from the
    reference sources) on the same mesh, pins and velocities, 1 thread,
    YeohSkinFast, its own MINSTD sweep order, in two builds: the parity build
    (-O2 -mavx2 -ffp-contract=off) and the tuned one (-O3 -march=x86-64-v3);
    the faster is the value. Falls back to the oracle's reference-precision
    restatement when the compiled reference is absent.

    Also checks one frame of the GPU path: against the compiled reference
    (the GPU's colour order as tOrder, the reference's fixed Yeoh constants,
    no contacts: the reference has neither custom constants nor contacts)
    and against the FP64 oracle with the workload's full settings. opts:
    the scene options of the GPU runs (the headline's schedule)."""
    import pyoracle as po

    opts = opts or {}
    xyz, tets = wl["xyz"], wl["tets"]
    t = tets.size // 4
    dt = np.float32(wl["dt"])
    S = wl["substeps"]
    s = wl["settings"]
    x0 = xyz.astype(np.float64)
    check = {"frames": 1}
    plain = dict(wl, settings=dict(s, collide=0, yeoh=xf.SKIN_YEOH))
    g = workloads.make_scene(ctx, plain, **opts)
    g.step(S, float(dt))
    xg, _ = g.get_state(0)
    order = g.order(0)[0]
    g.close()
    if po.have_ref():
        rb = po.RefBody(xyz, tets, wl["density"], wl["pinned"], order=order)
        rb.set_state(X=x0, V=wl["velocity"].astype(np.float64))
        rb.substeps(rb.settings(s["compliance"], nu=s["poissons_ratio"], energy=7, simultaneous=True), dt, S)
        check["rel_pos_err_vs_reference"] = rel_pos_err(xg, rb.state()[0], xyz)
        del rb
    g = workloads.make_scene(ctx, wl, **opts)
    g.step(S, float(dt))
    xg, _ = g.get_state(0)
    order = g.order(0)[0]
    g.close()
    ob = po.body_from_workload(wl, variant="f64")
    ob.substep(po.params_from_settings(dt, s), order, count=S)
    check["rel_pos_err_vs_fp64_oracle"] = rel_pos_err(xg, ob.X, xyz)
    del ob

    def timed(run):
        run(1)  # warm-up
        n = 0
        t0 = time.perf_counter()
        while True:
            run(2)
            n += 2
            el = time.perf_counter() - t0
            if el >= budget_s:
                return n, el

    builds = {}
    if po.have_ref():
        for fast in ((False, True) if po.have_ref(fast=True) else (False,)):
            body = po.RefBody(xyz, tets, wl["density"], wl["pinned"], fast=fast)
            body.set_state(X=x0, V=wl["velocity"].astype(np.float64))
            st = body.settings(s["compliance"], nu=s["poissons_ratio"], energy=7, simultaneous=True)
            n, el = timed(lambda k: body.substeps(st, dt, k))  # noqa: B023
            builds["O3 x86-64-v3" if fast else "O2 avx2 no-contract (parity build)"] = (t * n / el, n, el)
            del body
        kind, what = "reference", "reference Geo3d::Substep (no collision pass: the reference has none)"
    else:
        ob = po.body_from_workload(wl, variant="ref64")
        order = po.minstd_order(t)
        p = po.params_from_settings(dt, s)
        n, el = timed(lambda k: ob.substep(p, order, count=k))
        builds["oracle"] = (t * n / el, n, el)
        kind, what = "port", "oracle reference-precision restatement"
    best = max(builds, key=lambda k: builds[k][0])
    v, n, el = builds[best]
    return dict(value=v, unit="tet-substeps/s", cores=1, kind=kind,
                sample=f"{what}, {best} build: {n} substeps of the {t}-tet beam, single thread, {el:.1f} s",
                builds={k: b[0] for k, b in builds.items()}, check=check)


def batched_leg(ctx, frames=3, schedule="blocks"):
    """256 boxes: one batched scene (every body's blocks, or colour c of all
    bodies, per launch; autotuned launch shape) against the plain per-body
    launch loop (one per-colour scene per body, stepped one after another)."""
    b = workloads.boxes_batch()
    S, dt = b["substeps"], float(b["dt"])
    T = sum(x["tets"].size // 4 for x in b["bodies"])
    sc = workloads.make_batch_scene(ctx, b, **SCHEDULES[schedule])
    mode, size, _ = sc.autotune(S, dt, reps=2)
    sc.step(S, dt)
    sc.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        sc.step(S, dt)
    sc.sync()
    el_b = time.perf_counter() - t0
    finite = sc.count_nonfinite() == 0
    info = sc.info()
    ncol = info["n_colours"]
    sc.close()
    singles = [workloads.make_batch_scene(ctx, b, bodies=[i]) for i in range(len(b["bodies"]))]
    for s in singles:
        s.step(S, dt)
    for s in singles:
        s.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        for s in singles:
            s.step(S, dt)
    for s in singles:
        s.sync()
    el_p = time.perf_counter() - t0
    for s in singles:
        s.close()
    batched = T * S * frames / el_b
    plain = T * S * frames / el_p
    out = {"workload": f"{b['name']}: {len(b['bodies'])} bodies, {T} tets, {S} substeps/frame, {ncol} colours",
           "batched_tet_substeps_per_s": batched, "plain_loop_tet_substeps_per_s": plain,
           "batched_over_plain": batched / plain, "frames": frames,
           "launch": {"mode": MODE_NAMES[mode], "size": size, "schedule": schedule,
                      "block_lanes": info["block_lanes"], "tet_lanes": info["tet_lanes"]},
           "finite": finite}
    out["check"] = batch_check(ctx, b, schedule, mode, size, block_lanes=info["block_lanes"],
                               tet_lanes=info["tet_lanes"])
    out["cpu_baseline"] = batch_cpu(b)
    if "value" in out["cpu_baseline"]:
        out["gpu_over_cpu_all_cores"] = batched / out["cpu_baseline"]["value"]
    return out


def batch_check(ctx, b, schedule, mode, size, sample=(0, 85, 170, 255), block_lanes=0, tet_lanes=1):
    """One frame of the batched scene (the timed launch shape) against the
    oracle's serial sweep of sampled bodies in the scene's order: at the
    reference's precision (rel. position error 0 = bit-exact) and in FP64."""
    import pyoracle as po

    S, dt = b["substeps"], b["dt"]
    opts = dict(SCHEDULES[schedule], launch_mode=mode,
                **({"persistent_blocks": size} if mode == xf.LAUNCH_PERSISTENT else {"block_size": size}))
    if mode == xf.LAUNCH_BLOCKS:
        opts.update(block_lanes=block_lanes, tet_lanes=tet_lanes)
    sc = workloads.make_batch_scene(ctx, b, **opts)
    sc.step(S, float(dt))
    err_ref, err_64 = 0.0, 0.0
    for i in sample:
        wl = b["bodies"][i]
        x, _ = sc.get_state(i)
        order = sc.order(i)[0]
        for variant in ("ref64", "f64"):
            ob = po.body_from_workload(wl, variant=variant)
            ob.substep(po.params_from_settings(dt, wl["settings"]), order, count=S)
            e = rel_pos_err(x, ob.X, wl["xyz"])
            if variant == "ref64":
                err_ref = max(err_ref, e)
            else:
                err_64 = max(err_64, e)
    sc.close()
    return {"frames": 1, "bodies": list(sample), "rel_pos_err_vs_oracle_reference_precision": err_ref,
            "rel_pos_err_vs_fp64_oracle": err_64}


def batch_cpu(b, frames=2):
    """The batch on the host: every body a reference GeoLinear3d (tuned
    build), bodies dealt to all host threads (each body's Substep is the
    reference's single-threaded code), and the single-thread rate on the same
    bodies. No ground contact (the reference has none)."""
    import pyoracle as po

    if not po.have_ref():
        return {"error": "compiled reference not present"}
    fast = po.have_ref(fast=True)
    bodies = [po.RefBody(w["xyz"], w["tets"], w["density"], None, fast=fast) for w in b["bodies"]]
    sts = [body.settings(w["settings"]["compliance"], nu=w["settings"]["poissons_ratio"],
                         energy=w["settings"]["energy"], simultaneous=True) for body, w in zip(bodies, b["bodies"])]
    T = sum(w["tets"].size // 4 for w in b["bodies"])
    dt = np.float32(b["dt"])
    threads = host_threads()
    po.substeps_many(bodies, sts, dt, 1, threads)  # warm-up
    t0 = time.perf_counter()
    po.substeps_many(bodies, sts, dt, frames * b["substeps"], threads)
    el = time.perf_counter() - t0
    t1 = time.perf_counter()
    po.substeps_many(bodies, sts, dt, 2, 1)
    el1 = time.perf_counter() - t1
    del bodies
    return {"value": T * frames * b["substeps"] / el, "unit": "tet-substeps/s", "cores": threads,
            "kind": "reference", "single_thread": T * 2 / el1,
            "sample": f"reference Geo3d::Substep ({'O3 x86-64-v3' if fast else 'parity'} build), {len(b['bodies'])} "
                      f"bodies over {threads} host threads: {frames * b['substeps']} substeps, {el:.1f} s; "
                      f"single thread 2 substeps, {el1:.1f} s"}


def rings_leg(ctx, frames=3, budget_s=8.0, xcd_slabs=True):
    """The 96x24x24 beam with the vertex-ring energy ContinuousSkin (10): the
    sweep runs over its 60,625 vertex rings (each gathers the ~24 tets around
    a vertex and solves two constraints over its ~15 nodes).

    Reports ring-substeps/s and tet-substeps/s, the ring kernel's achieved
    bandwidth over algorithmic bytes (per ring: node position and inverse
    mass 32 B, position write 24 B and node id 4 B per ring node, 32 B of
    constants per ring tet, 24 B of ring metadata), the compiled reference's
    own ring sweep timed on a bounded host sample, and one frame of the GPU
    against the compiled reference (the GPU's ring and tet orders as vOrder
    and tOrder, the reference's fixed Skin constants, no contacts)."""
    import pyoracle as po

    wl = workloads.beam_yeoh_96x24x24()
    s = dict(wl["settings"], energy=xf.ENERGY_CONTINUOUS_SKIN)
    wl = dict(wl, settings=s)
    S, dt = wl["substeps"], float(wl["dt"])
    T = wl["tets"].size // 4
    sc = workloads.make_scene(ctx, wl, xcd_slabs=xcd_slabs)
    mode, size, _ = sc.autotune(S, dt, reps=2)
    sc.step(S, dt)
    sc.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        sc.step(S, dt)
    sc.sync()
    el = time.perf_counter() - t0
    finite = sc.count_nonfinite() == 0
    info = sc.info()
    R = info["n_rings"]
    ms_ring, _, n_launch = sc.profile(S, dt)
    sc.close()
    node_total = int(xf.mesh_rings(wl["tets"])["nodes"].size)
    bytes_substep = 60 * node_total + 32 * 4 * T + 24 * R
    achieved = bytes_substep * S / (ms_ring * 1e-3) / 1e9
    out = {"workload": f"{wl['name']} with ContinuousSkin: {R} rings ({node_total} ring nodes, {4 * T} ring tets), "
                       f"{info['n_ring_colours']} ring colours, {S} substeps/frame",
           "ring_substeps_per_s": R * S * frames / el, "tet_substeps_per_s": T * S * frames / el,
           "ms_per_frame": 1e3 * el / frames, "finite": finite,
           "launch": {"mode": "per-colour", "size": size, "xcd_slabs": xcd_slabs},
           "roofline": {"bound": "hbm", "kernel": "xf_ring_colour", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "launches": n_launch,
                        "avg_launch_us": 1e3 * ms_ring / max(n_launch, 1),
                        "bytes_per_launch": bytes_substep * S / max(n_launch, 1)}}
    if not po.have_ref():
        return out
    plain = dict(wl, settings=dict(s, collide=0, yeoh=xf.SKIN_YEOH))
    g = workloads.make_scene(ctx, plain, xcd_slabs=xcd_slabs)
    g.step(S, dt)
    xg, _ = g.get_state(0)
    torder, rorder = g.order(0)[0], g.ring_order(0)[0]
    g.close()
    x0 = wl["xyz"].astype(np.float64)
    v0 = wl["velocity"].astype(np.float64)
    rb = po.RefBody(wl["xyz"], wl["tets"], wl["density"], wl["pinned"], order=torder)
    rb.set_ring_order(rorder)
    rb.set_state(X=x0, V=v0)
    st = rb.settings(s["compliance"], nu=s["poissons_ratio"], energy=xf.ENERGY_CONTINUOUS_SKIN)
    rb.substeps(st, np.float32(dt), S)
    out["check"] = {"frames": 1, "rel_pos_err_vs_reference": rel_pos_err(xg, rb.state()[0], wl["xyz"])}
    rb.set_state(X=x0, V=v0)
    rb.substeps(st, np.float32(dt), 1)  # warm-up
    n = 0
    t0 = time.perf_counter()
    while True:
        rb.substeps(st, np.float32(dt), 1)
        n += 1
        cel = time.perf_counter() - t0
        if cel >= budget_s:
            break
    del rb
    out["cpu_baseline"] = {"value": R * n / cel, "unit": "ring-substeps/s", "cores": 1, "kind": "reference",
                           "sample": f"reference Geo3d::Substep with ContinuousSkin: {n} substeps of the "
                                     f"{R}-ring beam, single thread, {cel:.1f} s"}
    out["gpu_over_cpu"] = out["ring_substeps_per_s"] / out["cpu_baseline"]["value"]
    return out


def timed_frames(sc, S, dt, frames):
    """Milliseconds per frame of `frames` steps after one warm-up step."""
    sc.step(S, dt)
    sc.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        sc.step(S, dt)
    sc.sync()
    return 1e3 * (time.perf_counter() - t0) / frames


def contracted_leg(ctx, wl, shape, frames=10):
    """The headline beam with the contracted sweep-block kernels
    (xf_scene_opts.contract: fused multiply-adds, not bit-exact) at the tuned
    launch shape: frame time, and the node positions after one frame against
    the exact kernels' (bit-identical to the reference's serial sweep), per
    node relative to |x| and relative to the largest displacement."""
    S, dt = wl["substeps"], float(wl["dt"])
    states = {}
    for c in (False, True):
        sc = workloads.make_scene(ctx, wl, contract=c, **shape)
        x0, _ = sc.get_state(0)
        sc.step(S, dt)
        states[c] = (x0.reshape(-1, 3), sc.get_state(0)[0].reshape(-1, 3))
        if c:
            ms = timed_frames(sc, S, dt, frames)
            used = sc.info()["contract"]
        sc.close()
    x0, xe = states[False]
    xc = states[True][1]
    d = np.linalg.norm(xc - xe, axis=1)
    return {"ms_per_frame": ms, "tet_substeps_per_s": (wl["tets"].size // 4) * S / (ms * 1e-3),
            "contract": used,
            "one_frame_vs_exact": {"rel_position_error": float(np.linalg.norm(xc - xe) / np.linalg.norm(xe)),
                                   "max_over_max_displacement": float(d.max() / np.linalg.norm(xe - x0, axis=1).max()),
                                   "max_abs_m": float(d.max())},
            "note": "exact kernels = reference rounding (bit-identical); contracted kernels fuse multiply-adds"}


def large_leg(ctx, frames=2, cells=(288, 72, 72), schedule="blocks"):
    """The same Yeoh beam at 288x72x72 cells (8,957,952 tets, 1,540,081
    nodes): colours large enough that the projection is bound by throughput
    (VALU and HBM), not by phase latency. Reports tet-substeps/s and the
    projection kernel's achieved algorithmic bandwidth for the schedule, and
    the per-colour XCD-slab schedule's frame time beside it."""
    wl = workloads.beam_yeoh(*cells)
    S, dt = wl["substeps"], float(wl["dt"])
    sc = workloads.make_scene(ctx, wl, **SCHEDULES[schedule])
    mode, size, _ = sc.autotune(S, dt, reps=1)
    info = sc.info()  # after tuning: the slot packing follows the chosen cap
    T = info["n_tets"]
    ms = timed_frames(sc, S, dt, frames)
    finite = sc.count_nonfinite() == 0
    ms_proj, _, n_launch = sc.profile(S, dt)
    sc.close()
    bytes_proj = projection_bytes(info, mode, S)
    achieved = bytes_proj / (ms_proj * 1e-3) / 1e9 if mode != xf.LAUNCH_PERSISTENT else None
    out = {"workload": f"{wl['name']}: {T} tets, {info['n_nodes']} nodes, {info['n_colours']} colours, "
                       f"{info['n_blocks']} sweep blocks, {S} substeps/frame",
           "tet_substeps_per_s": T * S / (ms * 1e-3), "ms_per_frame": ms, "finite": finite,
           "launch": {"mode": MODE_NAMES[mode], "size": size, "schedule": schedule,
                      "block_lanes": info["block_lanes"], "tet_lanes": info["tet_lanes"],
                      "node_phase": "in the last projection launch" if info["fused_finish"] else "own launch"},
           "roofline": {"bound": "hbm", "kernel": KERNELS[mode], "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                        "launches": n_launch, "avg_launch_us": 1e3 * ms_proj / max(n_launch, 1),
                        "bytes_per_launch": bytes_proj / max(n_launch, 1),
                        "fp32_valu": fp32_valu(T * S, ms_proj)}}
    if schedule != "slabs":
        sc = workloads.make_scene(ctx, wl, **SCHEDULES["slabs"])
        out["per_colour_slabs_ms_per_frame"] = timed_frames(sc, S, dt, frames)
        sc.close()
    return out


def adjoint_leg(ctx, reps=3, use_graph=True):
    """Adjoint of the 16x4x4 fitting beam (1,536 tets, per-tet mu, 30 frames of
    20 substeps, loss at frames 10, 20, 30, checkpoints every 20 substeps):
    wall time of xf_scene_adjoint (forward + checkpointed reverse) against the
    plain forward rollout of the same 600 substeps; the same for the FP64
    element model (xf_scene_adjoint_fp64) with the relative L2 distance
    between the two gradients."""
    truth = workloads.beam_fit_16x4x4()
    guess = dict(truth)
    guess["tet_mu"] = np.full(truth["tets"].size // 4, 5e4, np.float32)
    dt, S = float(truth["dt"]), truth["substeps"]
    n_sub = S * truth["frames"]
    loss_at = np.array([10 * S, 20 * S, 30 * S], np.uint32)
    ref = workloads.make_scene(ctx, truth, use_graph=use_graph)
    targets, done = [], 0
    for la in loss_at:
        ref.step(int(la) - done, dt)
        done = int(la)
        targets.append(ref.get_state(0)[0])
    ref.close()
    targets = np.array(targets)
    fwd = workloads.make_scene(ctx, guess, use_graph=use_graph)
    adj = workloads.make_scene(ctx, guess, use_graph=use_graph)
    x0, v0 = fwd.get_state(0)
    fwd.step(n_sub, dt)
    fwd.sync()
    adj.adjoint(n_sub, dt, loss_at, targets)
    t_f, t_a = [], []
    for _ in range(reps):
        fwd.set_state(0, x0, v0)
        t0 = time.perf_counter()
        fwd.step(n_sub, dt)
        fwd.sync()
        t_f.append(time.perf_counter() - t0)
        adj.set_state(0, x0, v0)
        t0 = time.perf_counter()
        L, g = adj.adjoint(n_sub, dt, loss_at, targets)
        t_a.append(time.perf_counter() - t0)
    adj.set_state(0, x0, v0)
    adj.adjoint(n_sub, dt, loss_at, targets, fp64=True)  # captures its reverse-pass graph
    t_64 = []
    for _ in range(reps):
        adj.set_state(0, x0, v0)
        t0 = time.perf_counter()
        L64, g64 = adj.adjoint(n_sub, dt, loss_at, targets, fp64=True)
        t_64.append(time.perf_counter() - t0)
    fwd.close()
    adj.close()
    tf, ta, t64 = min(t_f), min(t_a), min(t_64)
    return {"workload": f"beam16x4x4_fit: {truth['tets'].size // 4} tets, {n_sub} substeps, loss at frames 10/20/30",
            "forward_ms": 1e3 * tf, "adjoint_ms": 1e3 * ta, "adjoint_over_forward": ta / tf,
            "reverse_over_forward": ta / tf - 1.0, "loss": L, "grad_norm": float(np.linalg.norm(g)),
            "fp64_model": {"adjoint_ms": 1e3 * t64, "over_fp32_forward": t64 / tf, "loss": L64,
                           "grad_rel_l2_fp32_vs_fp64": float(np.linalg.norm(g - g64) / np.linalg.norm(g64))}}


def fit_leg(ctx, iters=200):
    """Batched fitting: 64 Yeoh beams (cells 4-16 per axis, hidden per-tet
    log-mu fields), 30 frames of 40 substeps, node positions of every frame
    in the loss, no prior. Times the forward rollout of all bodies and runs
    `iters` Adam iterations (lr 0.05 in log mu), each one xf_scene_adjoint
    over the whole batch (one gradient per body); the first is untimed (it
    sizes the checkpoint and tape buffers). Reports the relative mu error
    before and after."""
    fit = fitting.BatchFit(ctx, workloads.fit_batch())
    err0 = fit.mu_error()
    fit.forward()
    t0 = time.perf_counter()
    fit.forward()
    tf = time.perf_counter() - t0
    losses = [fit.step()]
    t0 = time.perf_counter()
    for i in range(1, iters):
        losses.append(fit.step())
        if (i + 1) % 50 == 0:
            print(f"fit: {i + 1}/{iters} Adam iterations, loss {losses[-1]:.4e}", file=sys.stderr, flush=True)
    ta = (time.perf_counter() - t0) / max(iters - 1, 1)
    nb = len(fit.tets_per_body)
    err = fit.mu_error()
    out = {"workload": f"fit{nb}: {nb} beams, {sum(fit.tets_per_body)} tets, {fit.n_sub} substeps, loss at frames "
                       f"{fitting.frames_text(fit.loss_frames)}, window {fit.window}",
           "forward_ms": 1e3 * tf, "adjoint_ms": 1e3 * ta, "adjoint_over_forward": ta / tf,
           "reverse_over_forward": ta / tf - 1.0, "body_gradients_per_s": nb / ta, "adam_iterations": iters,
           "loss_first": losses[0], "loss": losses[-1], "body_step_backs": fit.failures,
           "rel_mu_err_initial": {"median": float(np.median(err0)), "mean": float(err0.mean())},
           "rel_mu_err_final": {"iterations": iters, "median": float(np.median(err)), "mean": float(err.mean()),
                                "p90": float(np.quantile(err, 0.9))}}
    fit.close()
    out["cpu_forward"] = fit_cpu(workloads.fit_batch(), tf)
    return out


def fit_cpu(batch, gpu_forward_s, substeps=40):
    """The fit batch's forward substep on the host, all threads: each beam a
    reference GeoLinear3d with the reference's own YeohSkinFast constants and
    one mu per beam (the reference has neither per-tet mu nor other Yeoh
    ratios: the same work per tet), pins at x = 0. `substeps` of the rollout
    timed and scaled to the rollout's length; no reference adjoint exists."""
    import pyoracle as po

    if not po.have_ref():
        return {"error": "compiled reference not present"}
    fast = po.have_ref(fast=True)
    bodies, sts = [], []
    for w in batch["bodies"]:
        rb = po.RefBody(w["xyz"], w["tets"], w["density"], w["pinned"], fast=fast)
        bodies.append(rb)
        sts.append(rb.settings(float(np.float32(1.0) / np.median(w["mu_guess"])), nu=0.5, energy=7,
                               simultaneous=True))
    T = sum(w["tets"].size // 4 for w in batch["bodies"])
    n_sub = batch["frames"] * batch["substeps"]
    dt = np.float32(batch["dt"])
    threads = host_threads()
    po.substeps_many(bodies, sts, dt, 1, threads)
    t0 = time.perf_counter()
    po.substeps_many(bodies, sts, dt, substeps, threads)
    el = time.perf_counter() - t0
    del bodies
    rate = T * substeps / el
    return {"value": rate, "unit": "tet-substeps/s", "cores": threads, "kind": "reference",
            "rollout_s": T * n_sub / rate, "gpu_forward_over_cpu": (T * n_sub / rate) / gpu_forward_s,
            "sample": f"{substeps} substeps of {len(batch['bodies'])} beams over {threads} host threads, {el:.1f} s"}


def copy_probe(nbytes=1 << 28, reps=4):
    """Plain device-to-device copies of a known size (the HIP runtime's copy
    kernel): the reference reading for scaling the FETCH_SIZE / WRITE_SIZE
    counters of a rocprofv3 --pmc pass to bytes."""
    import ctypes as C

    xf.lib()  # binds the HIP runtime the library uses
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(reps):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": reps}))


def batch_profile(frames=1, xcd_slabs=False):
    """One batched frame of the 256-box scene with fixed per-colour launches of
    256 threads (short runs under rocprofv3 counter passes)."""
    ctx = xf.Context(0)
    b = workloads.boxes_batch()
    sc = workloads.make_batch_scene(ctx, b, block_size=256, xcd_slabs=xcd_slabs)
    for _ in range(frames):
        sc.step(b["substeps"], float(b["dt"]))
    sc.sync()
    info = sc.info()
    sc.close()
    ctx.close()
    print(json.dumps({"batch_profile": b["name"], "tets": info["n_tets"], "colours": info["n_colours"],
                      "frames": frames, "xcd_slabs": xcd_slabs}))


def rings_profile(frames=1, xcd_slabs=True):
    """One frame of the 96x24x24 beam with ContinuousSkin, per-colour launches
    of 64 threads (short runs under rocprofv3 counter passes)."""
    ctx = xf.Context(0)
    wl = workloads.beam_yeoh_96x24x24()
    wl = dict(wl, settings=dict(wl["settings"], energy=xf.ENERGY_CONTINUOUS_SKIN))
    sc = workloads.make_scene(ctx, wl, xcd_slabs=xcd_slabs)
    for _ in range(frames):
        sc.step(wl["substeps"], float(wl["dt"]))
    sc.sync()
    info = sc.info()
    sc.close()
    ctx.close()
    print(json.dumps({"rings_profile": wl["name"], "rings": info["n_rings"], "ring_colours": info["n_ring_colours"],
                      "frames": frames, "xcd_slabs": xcd_slabs}))


def beam_profile(schedule="blocks", block=64, tet_lanes=1, frames=1, graph=False, contract=False):
    """Frames of the 96x24x24 beam in a fixed launch shape: block launches or
    per-colour launches of `block` threads, `tet_lanes` lanes per tet in
    block launches; eager launches unless graph (short runs under rocprofv3
    counter and trace passes)."""
    ctx = xf.Context(0)
    wl = workloads.beam_yeoh_96x24x24()
    sc = workloads.make_scene(ctx, wl, block_size=block, tet_lanes=tet_lanes, use_graph=graph, contract=contract,
                              **SCHEDULES[schedule])
    for _ in range(frames):
        sc.step(wl["substeps"], float(wl["dt"]))
    sc.sync()
    info = sc.info()
    sc.close()
    ctx.close()
    print(json.dumps({"beam_profile": wl["name"], "tets": info["n_tets"], "colours": info["n_colours"],
                      "blocks": info["n_blocks"], "block_nodes": info["n_block_nodes"], "schedule": schedule,
                      "block": block, "tet_lanes": tet_lanes, "frames": frames, "graph": graph,
                      "contract": info["contract"]}))


def large_profile(schedule="blocks", block=256):
    """One frame of the 288x72x72 beam in a fixed launch shape: block launches,
    or per-colour launches of `block` threads (short runs under rocprofv3
    counter passes)."""
    ctx = xf.Context(0)
    wl = workloads.beam_yeoh(288, 72, 72)
    sc = workloads.make_scene(ctx, wl, block_size=block, **SCHEDULES[schedule])
    sc.step(wl["substeps"], float(wl["dt"]))
    sc.sync()
    info = sc.info()
    sc.close()
    ctx.close()
    print(json.dumps({"large_profile": wl["name"], "tets": info["n_tets"], "colours": info["n_colours"],
                      "blocks": info["n_blocks"], "block": block, "schedule": schedule}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint legs (single beam and batched fit)")
    ap.add_argument("--no-rings", action="store_true", help="skip the vertex-ring energy leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 288x72x72 bandwidth leg")
    ap.add_argument("--fit-iters", type=int, default=200, help="Adam iterations of the batched fitting leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="per-colour launches of 64 threads instead of the autotuned shape (short profiler passes)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 4 device copies of 256 MiB (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--batch-profile", action="store_true",
                    help="only run one frame of the 256-box batched scene (rocprofv3 counter passes)")
    ap.add_argument("--rings-profile", action="store_true",
                    help="only run one frame of the beam with the ContinuousSkin ring energy (rocprofv3 counter passes)")
    ap.add_argument("--large-profile", action="store_true",
                    help="only run one frame of the 288x72x72 beam (rocprofv3 counter passes)")
    ap.add_argument("--no-xcd-slabs", action="store_true",
                    help="plain blockIdx-linear order in per-colour launches instead of the XCD-slab one")
    ap.add_argument("--schedule", choices=sorted(SCHEDULES), default="blocks",
                    help="sweep schedule of the beam, large and batched legs: sweep blocks (default) or "
                         "per-colour launches in the XCD-slab or plain order")
    ap.add_argument("--profile", action="store_true",
                    help="only run frames of the beam in a fixed launch shape (rocprofv3 counter and trace "
                         "passes; see --block, --tet-lanes, --frames, --graph)")
    ap.add_argument("--block", type=int, default=64, help="--profile: workgroup size")
    ap.add_argument("--tet-lanes", type=int, default=1, help="--profile: lanes per tet of block launches")
    ap.add_argument("--frames", type=int, default=1, help="--profile: frames")
    ap.add_argument("--graph", action="store_true", help="--profile: replay each frame from a hipGraph")
    ap.add_argument("--contract", action="store_true",
                    help="--profile: the sweep-block kernels with multiply-add contraction")
    ap.add_argument("--adjoint-profile", action="store_true",
                    help="only run the adjoint leg's forward and adjoint with eager launches (rocprofv3 kernel "
                         "traces)")
    args = ap.parse_args()
    args.xcd_slabs = not args.no_xcd_slabs
    if args.copy_probe:
        copy_probe()
        return
    if args.batch_profile:
        batch_profile(xcd_slabs=args.xcd_slabs)
        return
    if args.rings_profile:
        rings_profile(xcd_slabs=args.xcd_slabs)
        return
    if args.large_profile:
        large_profile(schedule=args.schedule)
        return
    if args.adjoint_profile:
        ctx = xf.Context(0)
        print(json.dumps({"adjoint_profile": adjoint_leg(ctx, reps=2, use_graph=False)}))
        ctx.close()
        return
    if args.profile:
        beam_profile(schedule=args.schedule, block=args.block, tet_lanes=args.tet_lanes, frames=args.frames,
                     graph=args.graph, contract=args.contract)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is not built",
              file=sys.stderr)
        sys.exit(2)

    wl = workloads.beam_yeoh_96x24x24()
    ctx = xf.Context(0)
    opts = SCHEDULES[args.schedule]
    sc = workloads.make_scene(ctx, wl, use_graph=True, **opts)
    info = sc.info()
    T, N, S = info["n_tets"], info["n_nodes"], wl["substeps"]
    dt = float(wl["dt"])
    if args.no_autotune:
        mode, size = opts.get("launch_mode", xf.LAUNCH_PER_COLOUR), 64
    else:
        mode, size = sc.autotune(S, dt, reps=2)[:2]
    info = sc.info()  # the slot packing may have changed with the tuned cap

    for _ in range(args.warmup):
        sc.step(S, dt)
    sc.sync()
    # Timed region: `steps` frames (graph replays), wall clock around HIP
    # events recorded on the library stream before the first and after the
    # last frame.
    t0 = time.perf_counter()
    ms_dev = sc.time_steps(args.steps, S, dt)
    el = time.perf_counter() - t0
    bad = sc.count_nonfinite()
    if bad:
        print(f"non-finite positions after the run: {bad}", file=sys.stderr)
        sys.exit(1)

    # Dominant kernel: the sweep-block or per-colour projection launches (or
    # the single persistent launch, which also holds the node phases). Its
    # average launch time comes from the timed region: the event-timed frame
    # less the frame's own node launches (the first prediction and the last
    # velocity update, timed per launch in one eager frame), over the frame's
    # projection launches. It includes the gaps between launches of the
    # graph, so the fraction is a lower bound; the eager per-launch event
    # times are reported beside it.
    ms_proj_eager, ms_node, n_launch = sc.profile(S, dt)
    fused_now = bool(sc.info()["fused_finish"])
    bytes_proj = projection_bytes(info, mode, S)
    ms_frame_dev = ms_dev / args.steps
    ms_proj = ms_frame_dev - (ms_node if mode != xf.LAUNCH_PERSISTENT else 0.0)
    achieved = bytes_proj / (ms_proj * 1e-3) / 1e9
    # the same launches against the fixed model of 196 B per tet projection
    # (16 B indices, 64 B constants and material, 64 B gather, 48 B scatter
    # of FP32 positions; the node phases are not counted)
    bytes_model = S * 196 * T

    # HBM traffic per projection launch: calibrated FETCH_SIZE + WRITE_SIZE of the
    # rocprofv3 --pmc passes recorded for this workload and schedule (profiles/),
    # when the launch shape matches.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_beam96x24x24.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f)
        shapes = rec.get("blocks_shapes", [])
        match = [b for b in shapes if b.get("n_blocks") == info["n_blocks"] and b.get("block_size") == size and
                 b.get("block_lanes") == info["block_lanes"]]
        if mode == xf.LAUNCH_BLOCKS and match and info["tet_lanes"] == 1:
            traffic = float(match[0]["traffic_bytes_per_dispatch"])
        elif mode == xf.LAUNCH_PER_COLOUR and args.schedule != "blocks" and \
                rec.get("tets_per_dispatch") == T // max(info["n_colours"], 1):
            traffic = float(rec["traffic_bytes_per_dispatch"] if args.schedule == "slabs" else
                            rec["plain_schedule"]["traffic_bytes_per_dispatch"])

    out = {
        "metric": METRIC,
        "value": T * S * args.steps / el,
        "unit": "tet-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 element math, fp64 node state",
        "data": "synthetic",
        "config": {"workload": f"{wl['name']}: {T} tets, {N} nodes, {S} substeps/frame, {info['n_colours']} "
                               f"colours, {info['n_blocks']} sweep blocks, ground + sphere", "step": "one frame",
                   "launch": {"mode": MODE_NAMES[mode], "size": size, "schedule": args.schedule,
                              "tet_lanes": sc.info()["tet_lanes"] if mode == xf.LAUNCH_BLOCKS else 1,
                              "block_lanes": info["block_lanes"],
                              "node_phase": "in the last projection launch" if fused_now else "own launch"}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": KERNELS[mode],
                     "launches": n_launch, "avg_launch_us": 1e3 * ms_proj / max(n_launch, 1),
                     "bytes_per_launch": bytes_proj / max(n_launch, 1),
                     "frac_196B_model": bytes_model / (ms_proj * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "ms_frame_events": ms_frame_dev, "ms_node_launches": ms_node,
                     "avg_launch_us_eager": 1e3 * ms_proj_eager / max(n_launch, 1),
                     "fp32_valu": fp32_valu(T * S, ms_proj)},
    }
    if args.schedule == "blocks":
        # the per-colour schedule on the same beam (structured 24 colours, XCD slabs)
        pc = workloads.make_scene(ctx, wl, **SCHEDULES["slabs"])
        pc.autotune(S, dt, reps=2)
        ms_pc = timed_frames(pc, S, dt, 10)
        pc.close()
        out["per_colour_slabs"] = {"ms_per_frame": ms_pc, "tet_substeps_per_s": T * S / (ms_pc * 1e-3)}
        if mode == xf.LAUNCH_BLOCKS:
            try:
                out["contracted"] = contracted_leg(
                    ctx, wl, dict(opts, launch_mode=mode, use_graph=True, block_size=size,
                                  tet_lanes=info["tet_lanes"], block_lanes=info["block_lanes"]))
            except xf.XfError as e:
                out["contracted"] = {"error": str(e)}
    legs = []
    if not args.no_batch:
        legs.append(("batched", lambda c: batched_leg(c, schedule=args.schedule)))
    if not args.no_large:
        legs.append(("large", lambda c: large_leg(c, schedule=args.schedule)))
    if not args.no_rings:
        legs.append(("rings", lambda c: rings_leg(c, xcd_slabs=args.xcd_slabs)))
    if not args.no_adjoint:
        legs += [("adjoint", adjoint_leg), ("fit", lambda c: fit_leg(c, iters=args.fit_iters))]
    if not args.no_cpu:
        # the one-frame checks run at the timed launch shape
        shape = dict(opts, launch_mode=mode, use_graph=True, tet_lanes=info["tet_lanes"],
                     block_lanes=info["block_lanes"] if mode == xf.LAUNCH_BLOCKS else opts.get("block_lanes", 0),
                     **({"persistent_blocks": size} if mode == xf.LAUNCH_PERSISTENT else {"block_size": size}))
        legs.append(("cpu_baseline", lambda c: cpu_baseline(c, wl, opts=shape)))
    for name, leg in legs:
        print(f"bench: {name} leg", file=sys.stderr, flush=True)
        try:
            out[name] = leg(ctx)
        except (xf.XfError, FloatingPointError) as e:  # a side leg must not hide the main line
            out[name] = {"error": str(e)}
    sc.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
