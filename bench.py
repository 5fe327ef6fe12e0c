"""Benchmark of the cloth step on one MI355X.

Workload (N=1): the large configuration -- a 512x512 cloth (262,144 vertices,
3,135,492 directed springs of three types) over a 12-capsule body of 501,120
triangles that is deformed every frame with its LBVH refitted, FP32, dt=5e-4.
One step = body deformation + LBVH refit + fused Verlet/spring/collision
kernel.  Prints ONE JSON line.  Beside the headline: "batched" (256 uneven
instances with bodies rebuilt every frame, through one Batch and through the
plain per-instance loop) and "adjoint" (forward + backward of a 200-step
differentiable rollout against the plain rollout).

  python bench.py [--gpus 1] [--steps 500] [--warmup 50] [--no-cpu] [--no-batch] [--no-adjoint]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, GB/s
WORKLOADS = {
    "large": "large: 512x512 cloth, 262,144 vertices, 3,135,492 directed springs (structural/shear/bend), "
             "12-capsule body of 501,120 triangles deformed and LBVH-refitted every frame, dt=5e-4",
    "garment": "garment: 111x111 cloth, 12,321 vertices, 12-capsule body of 50,688 triangles deformed and "
               "LBVH-rebuilt every frame, dt=1e-3",
    "small": "small: 64x64 cloth, 4,096 vertices, static 5,120-triangle icosphere, dt=1e-3",
    "reference": "reference: the reference's own scene (main.cpp): tshirt2.obj, 16,773 vertices, 195,534 "
                 "directed mesh springs, over female.obj, static 13,776 triangles, its constants (dt=1/40)",
}


def algorithmic_bytes(sc):
    """Payload bytes per launch (12 B per FP32 3-vector, 4 B per index/scalar,
    float4 padding and neighbour gathers not counted)."""
    n = sc.n
    E = sum(int(c.shape[0]) for _, c in sc.springs)
    verlet = n * (36 + 36 + 12) + 8 * E          # read/write x, x_prev, n_contact; 3 row ptrs; records
    T = sc.body.tris.shape[0]
    Vb = sc.body.verts.shape[0]
    deform = 24 * Vb                             # read rest, write current
    refit = 12 * Vb + T * (12 + 24) + (T - 1) * (24 + 8 + 4)   # vertices, leaf index+box, internal box+links
    return {"verlet": verlet, "deform": deform, "refit": refit, "springs": E}


def step_kernel_figures(live, tk, ab, rebuild):
    """Kernel times of the report and the step kernel's achieved bandwidth.
    live: event timing over the timed steps (cc_sim_get_timing); tk: the
    same kernel groups one at a time (cc_sim_time_kernels).  The step kernel's
    live time is used when the timed steps recorded it, else its isolated
    time; achieved is None when neither is positive."""
    upd = "build_us" if rebuild else "refit_us"
    kern = {"verlet": live["verlet_us"], "body_update": live[upd]}
    iso = {"verlet": tk["verlet_us"], "build" if rebuild else "refit": tk[upd], "deform": tk["body_us"]}
    if kern["verlet"] > 0:
        t, dom = kern["verlet"], "live"
    elif iso["verlet"] > 0:
        t, dom = iso["verlet"], "isolated"
    else:
        return kern, iso, "none", None
    return kern, iso, dom, ab["verlet"] / (t * 1e-6) / 1e9


def read_traffic(path):
    """HBM bytes per launch of the step kernel from the committed rocprofv3
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by the same counter on a
    plain copy of known size); None when the summary is absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def copy_probe(nbytes):
    """One device-to-device copy of nbytes of freshly allocated (cold) memory,
    through the HIP runtime's own copy kernel."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    big = C.c_void_p()                    # evict a from the caches: touch 1 GiB of other memory
    assert hip.hipMalloc(C.byref(big), C.c_size_t(1 << 30)) == 0
    assert hip.hipMemset(big, 0, C.c_size_t(1 << 30)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpyDtoD(b, a, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    for ptr in (a, b, big):
        hip.hipFree(ptr)


def cpu_baseline(sc, budget_s=15.0, nthreads=1, label="large"):
    """The oracle's FP32 path (the reference algorithm on the host) on a
    bounded sample of the same workload: whole steps (deform + refit or
    rebuild + step; the step alone for a static body) from the initial
    state, one warm-up step, then three runs of the same number of steps
    (sized so the three take about budget_s), the median reported."""
    import oracle as O
    from simulator_amd import _lib as L
    O.set_threads(nthreads)            # boxes, deformation and refit subtrees; the step takes nthreads
    b = sc.body
    v4 = np.concatenate([b.verts, np.ones((b.verts.shape[0], 1), np.float32)], 1)
    rest = O.vertex_extend(v4, b.tris, b.tris, b.normals, sc.params.body_extend)[:, :3].copy()
    caps = b.caps
    d = O.bvh_build(rest, b.tris, sc.params.box_depth)
    prm = O.params_from(sc.params)
    pins = np.zeros(sc.n, np.uint8)
    pins[sc.pins] = 1
    m = b.motion
    moving = b.mode in (L.CC_BODY_REFIT, L.CC_BODY_REBUILD)
    state = {"x": sc.pos.copy(), "xp": sc.pos.copy(), "c": np.zeros_like(sc.pos), "frame": 0, "d": d,
             "h": O.BvhHandle(d)}

    def one_step():
        if moving:
            # frame motion (same formulas as the device deformation, host doubles)
            t = state["frame"] * float(np.float32(sc.params.dt))
            w = np.sin(2 * np.pi * m["freq"] * t)
            th = m["rot_amp"] * w
            R64 = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
            P = np.array(m["pivot"])
            T = (P - R64 @ P + m["amp"] * w * np.array(m["dir"])).astype(np.float32)
            R = R64.astype(np.float32).ravel()
            breath = (1 + m["breath_amp"] * np.sin(2 * np.pi * m["breath_freq"] * t
                                                   + caps["phase"].astype(np.float64))).astype(np.float32)
            verts = O.body_deform(rest, caps["vstart"], caps["p0"], caps["p1"], breath, R, T)
            dd = state["d"]
            if b.mode == L.CC_BODY_REBUILD:
                dd = O.bvh_build(verts, b.tris, sc.params.box_depth)
            else:
                lb = O.prim_boxes(verts, dd["leaf_tri"], sc.params.box_depth)
                dd["box_int"][:] = O.refit(dd["childA"], dd["childB"], dd["parent_int"], lb)
                dd["box_leaf"][:] = lb
                dd["verts"] = np.ascontiguousarray(verts)
            state["d"] = dd
            state["h"] = O.BvhHandle(dd)
        state["x"], state["xp"], state["c"] = O.step_f32(state["x"], state["xp"], state["c"], sc.springs, prm,
                                                         pins, state["h"], nthreads=nthreads)
        state["frame"] += 1

    t0 = time.perf_counter()
    one_step()                          # warm-up
    per = time.perf_counter() - t0
    nsteps = int(max(1, min(200, budget_s / 3.0 / max(per, 1e-6))))
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(nsteps):
            one_step()
        runs.append(time.perf_counter() - t0)
    O.set_threads(1)
    el = float(np.median(runs))
    what = {L.CC_BODY_REFIT: "deform + refit + ", L.CC_BODY_REBUILD: "deform + rebuild + "}.get(b.mode, "")
    return {"value": sc.n * nsteps / el, "unit": "vertex-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"median of 3 runs of {nsteps} whole steps ({what}Verlet/spring/collision) of the same "
                      f"{label} workload after 1 warm-up step from the initial state, FP32 oracle, "
                      f"{el:.2f} s per run"}


def batched_leg(warmup=16, steps=160, plain_steps=48):
    """The batched workload (256 instances, bodies rebuilt every frame) through
    one Batch, beside the plain per-instance loop (one Simulator per instance
    advanced in turn, each call plain_steps frames -- whole graph blocks of
    the simulator -- with no synchronisation between instances) on the same
    GPU; plain_steps=0 times the batch alone."""
    from simulator_amd import batch as B
    from simulator_amd.simulator import Simulator
    scs = B.instances(256, seed=0)
    n = sum(sc.n for sc in scs)
    b = B.Batch(scs)
    b.simulate(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.simulate(steps)
    b.synchronize()
    tb = time.perf_counter() - t0
    nv, tile, order = b.info()
    tuned, fused = b.tuning()
    b.close()
    out = {"workload": "256 instances, cloth side log-uniform [16,192] (seed 0), capsule bodies of "
                       "2,112-36,096 triangles rebuilt every frame, dt=1e-3",
           "instances": len(scs), "vertices": n, "body_triangles": int(sum(sc.body.tris.shape[0] for sc in scs)),
           "value": n * steps / tb, "unit": "vertex-steps/s", "ms_per_step": 1e3 * tb / steps, "steps": steps,
           "tile": tile, "order": "largest-first" if order else "instance",
           "tile_ran_in": "the per-body sort's launch (k_bb_sort_step)" if fused else "k_verlet_batch",
           "tuner_ms_per_step": {f"tile{t}/{o}": round(float(tuned[i, j]), 4)
                                 for i, t in enumerate((64, 128, 256))
                                 for j, o in enumerate(("instance", "largest-first"))}}
    if plain_steps <= 0:
        return out
    sims = [Simulator.from_scene(sc) for sc in scs]
    block = sims[0].block_steps
    for s in sims:
        s.simulate(block)
    for s in sims:
        s.synchronize()
    t0 = time.perf_counter()
    for s in sims:
        s.simulate(plain_steps)
    for s in sims:
        s.synchronize()
    tp = time.perf_counter() - t0
    for s in sims:
        s.close()
    vp = n * plain_steps / tp
    out.update({"plain_loop_value": vp, "plain_loop_ms_per_step": 1e3 * tp / plain_steps,
                "plain_loop_steps": plain_steps,
                "plain_loop_composition": "per simulator: %d graph blocks of %d steps + %d single steps" % (
                    plain_steps // block, block, plain_steps % block),
                "speedup_vs_plain": out["value"] / vp})
    return out


def adjoint_leg(nsteps=200, reps=3):
    """Forward + backward of the differentiable rollout (128x128 cloth over the
    icosphere, 200 steps, checkpoints every 20) beside the plain 200-step
    simulation of the same scene."""
    from simulator_amd import scene as S
    from simulator_amd.adjoint import Adjoint
    from simulator_amd.simulator import Simulator
    sc = S.adjoint_scene(steps=nsteps)
    sim = Simulator.from_scene(sc)
    adj = Adjoint(sim, nsteps, 20)
    x0 = sc.pos.astype(np.float32)
    v0 = np.zeros_like(x0)
    theta = np.array([sc.params.k[0], sc.params.k[1], sc.params.k[2], sc.params.damp, sc.params.rest_scale])
    g = np.ones_like(x0)
    adj.forward(x0, v0, theta)
    adj.backward(g)
    tf = tbw = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        adj.forward(x0, v0, theta)
        t1 = time.perf_counter()
        adj.backward(g)
        t2 = time.perf_counter()
        tf += t1 - t0
        tbw += t2 - t1
    plain = Simulator.from_scene(sc)
    plain.simulate(16)
    plain.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain.simulate(nsteps)
    plain.synchronize()
    tp = (time.perf_counter() - t0) / reps
    adj.close()
    sim.close()
    plain.close()
    tf, tbw = tf / reps, tbw / reps
    return {"workload": "128x128 cloth over a 5,120-triangle icosphere, 200 steps, checkpoints every 20",
            "forward_ms": 1e3 * tf, "backward_ms": 1e3 * tbw, "plain_ms": 1e3 * tp,
            "forward_backward_vs_plain": (tf + tbw) / tp,
            "gradient_rollouts_per_s": 1.0 / (tf + tbw)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--workload", choices=["large", "garment", "small", "reference"], default="large",
                    help="large (default, the headline): 512x512 over 501k refitted triangles; garment: "
                         "111x111 over 50k triangles rebuilt every frame; small: 64x64 over a static sphere; "
                         "reference: the reference's tshirt over its female body")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB of cold memory "
                         "(the reference kernel for calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the cloth step is a "
              "single-GPU kernel; multi-GPU sharding is a design note, not a run)", file=sys.stderr)
        sys.exit(2)

    from simulator_amd import _lib as L
    from simulator_amd import scene as S
    from simulator_amd.simulator import Simulator

    sc = {"large": S.large, "garment": S.garment, "small": S.small,
          "reference": lambda: S.load_mesh_scene(os.path.join(ROOT, "tests", "golden", "ref_tshirt_female.npz"))
          }[args.workload]()
    sim = Simulator.from_scene(sc)
    # HIP events around every step graph, single step kernel and body
    # update, on the streams they run on, read while later work runs
    sim.set_timing(True)
    sim.simulate(args.warmup)
    sim.synchronize()
    sim.set_timing(True)                  # averages over the timed steps only
    t0 = time.perf_counter()
    sim.simulate(args.steps)
    sim.synchronize()
    el = time.perf_counter() - t0
    ms = 1e3 * el / args.steps
    value = sc.n * args.steps / el
    live = sim.timing()

    sim.set_timing(False)

    # the reference's calling pattern, one simulate() call per frame, over
    # the same frames as the timed steps (a fresh simulator, the same warm-up)
    pf = Simulator.from_scene(sc)
    pf.set_timing(True)
    pf.simulate(args.warmup)
    pf.synchronize()
    pf.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pf.simulate(1)
    pf.synchronize()
    el_pf = time.perf_counter() - t0
    live_pf = pf.timing()
    pf.close()

    # the same kernel groups one at a time, nothing overlapped
    tk = sim.time_kernels(args.steps)
    # traversal counters over a separate run of the same steps
    sim.set_counting(True)
    t1 = time.perf_counter()
    sim.simulate(args.steps)
    sim.synchronize()
    el_c = time.perf_counter() - t1
    visits, queries = sim.counters()
    sim.set_counting(False)
    sim.check_finite()
    block = sim.block_steps
    sim.close()

    ab = algorithmic_bytes(sc)
    rebuild = sc.body.mode == L.CC_BODY_REBUILD
    kern, iso, dom, achieved = step_kernel_figures(live, tk, ab, rebuild)
    step_bytes = ab["verlet"] + ab["refit"] + ab["deform"]
    out = {
        "metric": "vertex-steps/s @262k-vert cloth FP32, % HBM peak; batched & adjoint vs plain",
        "value": value,
        "unit": "vertex-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("the reference's own meshes (tshirt2.obj, female.obj) from tests/golden/ref_tshirt_female.npz"
                 if args.workload == "reference" else "synthetic (seeded grid cloth and capsule-union body)"),
        "config": {"workload": WORKLOADS[args.workload],
                   "vertices": sc.n, "springs_directed": ab["springs"],
                   "body_triangles": int(sc.body.tris.shape[0])},
        "roofline": {"bound": "hbm", "kernel": "k_verlet", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
                     "traffic": None, "timing": dom},
        "kernels_us": kern,
        "kernels_timed_steps": live["steps"],
        "kernels_isolated_us": iso,
        "path": "%d graph blocks of %d steps + %d single steps" % (args.steps // block, block, args.steps % block),
        "per_frame_calls": {"calls": args.steps, "ms_per_step": 1e3 * el_pf / args.steps,
                            "value": sc.n * args.steps / el_pf,
                            "verlet_us": live_pf["verlet_us"],
                            "body_update_us": live_pf["build_us" if rebuild else "refit_us"],
                            "vs_headline": (sc.n * args.steps / el_pf) / value},
        "step_algorithmic_bytes": step_bytes,
        "step_hbm_frac": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "bvh_nodes_visited_per_s": visits / el_c,
        "bvh_nodes_per_query": visits / max(queries, 1),
    }
    traffic = read_traffic(os.path.join(ROOT, "profiles", "k_verlet_%s_traffic.txt" % args.workload))
    if traffic:
        out["roofline"]["traffic"] = traffic
    if not args.no_batch:
        out["batched"] = batched_leg()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg()
    if not args.no_cpu:
        cb = cpu_baseline(sc, budget_s=args.cpu_budget, label=args.workload)
        # the same on all the host cores this process may use (OpenMP over
        # vertices, triangles and refit subtrees), at most 16: a GPU's share
        # of its host
        cores = min(16, len(os.sched_getaffinity(0)))
        if cores > 1:
            cb["all_cores"] = cpu_baseline(sc, budget_s=args.cpu_budget, nthreads=cores, label=args.workload)
        out["cpu_baseline"] = cb
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib << 20)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
