#!/usr/bin/env python3
"""Throughput of the VBD time step on one MI355X.

Workload (default): the 128^3-cell noisy cube -- 129^3 = 2,146,689 vertices, 12,582,912
Kuhn tets, mu = 1e5, lambda = 4e5, rho = 1000, bottom face pinned, free vertices jittered by
seeded uniform +-0.3 cell, h = 1/60, 20 VBD iterations per step, Chebyshev rho = 0.95.
A "step" is one whole time step (predict, 20 x (8 color launches + blend), velocity update)
replayed from one captured hipGraph with all state resident in HBM.

Prints one JSON line.  value = vertex-iterations per second (millions): all vertices
(pinned included) x iterations x steps / seconds.
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

METRIC = "M vertex-iters/s + %HBM roofline, 1 MI355X; batched & adjoint vs plain"
HBM_PEAK_GBS = 8000.0
H = 1.0 / 60.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name: str):
    """HBM bytes per color launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    as committed in profiles/<name>.txt on a line 'traffic_bytes_per_launch: <float>'."""
    path = os.path.join(ROOT, "profiles", name + ".txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def host_cpu():
    """The host's CPU as this process sees it: model string, nproc, the affinity set, and the
    threads to use -- the affinity set, capped by OMP_NUM_THREADS where the host sets it (a GPU
    host may allot each GPU a share of its cores that way)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = max(1, min(aff, int(share))) if share.isdigit() and int(share) > 0 else aff
    return {"model": model, "nproc": nproc, "affinity": aff, "omp_num_threads": share or None, "threads": threads}


def cpu_baseline(m, iterations: int, target_seconds: float = 10.0):
    """Reference-cost CPU restatement (oracle mode 2: full 12x12 element Hessian per
    (vertex, tet) call and a per-call std::vector, as VBD.cpp:205-256 / Energy.cpp:90-158)
    timed on a bounded sample of the same mesh: whole sweeps (every free vertex of every
    color, colors in order) with the host's threads inside each color (race-free; the result
    is independent of the thread count), repeated to about `target_seconds`; plus one thread
    on a smaller sample, one sweep in the reference's own OpenMP-over-colors scheme, and the
    diagnostic energy the reference's step evaluates every 10 iterations (VBD.cpp:92-94; serial
    in the reference, VBD.cpp:258-288, so timed on one thread).  Rates count every vertex
    (pinned included) per iteration, like the device value."""
    from oracle import oracle as O
    prep = O.prepare(m)
    y = m.x0.copy()
    y[m.fixed == 0] += np.float32(H * H) * np.asarray(m.gravity, np.float32)
    cpu = host_cpu()
    cores = cpu["threads"]
    colors = []
    for c in range(prep.n_colors):
        vs = prep.color_verts[prep.color_off[c]:prep.color_off[c + 1]]
        colors.append(vs[m.fixed[vs] == 0].astype(np.int32))
    nfree = int((m.fixed == 0).sum())
    nv = m.num_verts

    def run(threads, n_per_color):
        p = O.make_params(dt=H, iterations=1, rho=0.95, mu=m.mu, lam=m.lam, gravity=m.gravity,
                          mode=O.MODE_FAITHFUL_COST, threads=threads)
        secs, cnt = 0.0, 0
        for vs in colors:
            vs = vs[:n_per_color]
            _, s = O.solve_subset(prep, p, m.x0, y, vs)
            secs += s
            cnt += vs.size
        return secs, cnt

    # one thread, ~2 s
    s1, c1 = run(1, 200)
    n1 = int(max(50, min(200 * 2.0 / max(s1, 1e-6), max(len(v) for v in colors))))
    s1, c1 = run(1, n1)
    # all threads: whole sweeps, repeated
    full = max(len(v) for v in colors)
    sw, cw = run(cores, full)
    reps = max(1, int(round(target_seconds / max(sw, 1e-6))) - 1)
    for _ in range(reps):
        s, c = run(cores, full)
        sw += s
        cw += c
    sweeps = reps + 1
    sweep_s = sw / sweeps
    # the reference's own scheme, OpenMP over colors (racy; at most one thread per color)
    pr = O.make_params(dt=H, iterations=1, rho=0.95, mu=m.mu, lam=m.lam, gravity=m.gravity,
                       mode=O.MODE_FAITHFUL_COST, threads=cores)
    _, sr = O.sweep_over_colors(prep, pr, m.x0, y)
    # the diagnostic energy (|det F| for the SVD's singular value product: cheaper than the
    # reference's JacobiSVD per tet), one evaluation on one thread
    t0 = time.perf_counter()
    O.total_energy(prep, pr, m.x0)
    energy_s = time.perf_counter() - t0
    per_free = nv / nfree  # free-vertex solves -> vertex-iterations (pinned included)
    return {"value": cw * per_free / sw / 1e6, "unit": "M vertex-iters/s", "cores": cores, "kind": "port",
            "sample": f"{sweeps} whole sweeps of the same mesh, {m.name} ({cw} vertex solves in {sw:.1f} s), reference "
                      f"cost structure (full 12x12 element Hessian per (vertex, tet) call, per-call std::vector), "
                      f"{cores} threads inside each color; energy diagnostic excluded",
            "host_cpu": cpu,
            "with_energy_every_10_value": nv / (sweep_s + energy_s / 10.0) / 1e6,
            "energy_eval_s": round(energy_s, 3),
            "energy_sample": f"one log-Neo-Hookean diagnostic over the {m.num_tets} tets on 1 thread ({energy_s:.2f} s), "
                             f"added once per 10 iterations as VBD.cpp:92-94 does",
            "single_thread_value": c1 * per_free / s1 / 1e6,
            "single_thread_sample": f"first {n1} free vertices of each color, 1 thread ({c1} solves in {s1:.1f} s)",
            "reference_scheme_value": nv / sr / 1e6,
            "reference_scheme_sample": f"one whole sweep with OpenMP over the {prep.n_colors} colors as VBD.cpp:119-121 "
                                       f"({cores} threads requested, racy; {nfree} solves in {sr:.1f} s)"}


def bench_batch(count: int = 256, frames: int = 60, baseline_frames: int = 3, iterations: int = 20,
                persistent: bool = True, launch=None):
    """Batched workload: `count` independent beams of uneven size (nx ~ U{8..64}, ny = nz ~
    U{2..8}, per-beam mu, nu), per-tet materials, 20 iterations with Chebyshev 0.95, all meshes
    stepped by one hipGraph per frame, against the same meshes stepped one at a time with plain
    (non-graph) launches, and one at a time with a hipGraph per mesh (baseline_frames 0 skips
    them).  `launch` = (lanes, block, pairs) fixes the launch configuration instead of
    autotuning it.  The color kernel's measured HBM traffic is read from the profile committed
    for the configuration that ran (profiles/traffic_color_tiled_batched<count>_<block>x<pairs>.txt)."""
    from vbd_tetmesh_amd import VBDParams, meshgen
    from vbd_tetmesh_amd.solver import DeviceMesh, prepare_host
    ms = meshgen.batch_beams(count, seed=0)
    params = VBDParams(dt=H, numIterations=iterations, gravity=(0.0, -9.8, 0.0), useAcceleration=True,
                       accelerationRho=0.95)
    descs = []
    for m in ms:
        t, off, verts, dminv, vol = prepare_host(m)
        descs.append(dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                          rest_volume=vol, color_off=off, color_verts=verts, tet_mu=t.tetMu, tet_lambda=t.tetLambda))
    nv = sum(m.num_verts for m in ms)
    nt = sum(m.num_tets for m in ms)
    b = DeviceMesh().init_batch(descs)
    b.set_params(params)
    if launch is None:
        tune = b.autotune(8)
    else:
        b.set_launch_config(launch[0], launch[1], launch[2])
        tune = (launch[0], launch[2], launch[1])
    b.set_positions(np.concatenate([d["positions"] for d in descs]))
    b.step(1)  # warm-up (captures the graph)
    b.synchronize()
    b.set_positions(np.concatenate([d["positions"] for d in descs]))
    b.set_velocities(np.zeros((nv, 3), np.float32))
    b.synchronize()
    t0 = time.perf_counter()
    b.step(frames)
    b.synchronize()
    tb = (time.perf_counter() - t0) / frames
    xb = b.positions()
    if not np.all(np.isfinite(xb)):
        raise SystemExit("non-finite positions in the batched run")
    roof = launch_roofline(b, iterations,
                           f"traffic_color_tiled_batched{count}_{tune[2]}x{tune[1]}" if tune[0] == 0 else None)

    def per_mesh(graph: bool):
        ds = []
        for kw in descs:
            d = DeviceMesh().init(**kw)
            d.set_params(params)
            d.set_launch_config(tune[0], tune[2], tune[1])
            d.set_graph(graph)
            d.step(1)
            ds.append(d)
        for d in ds:
            d.synchronize()
        t0 = time.perf_counter()
        for _ in range(baseline_frames):
            for d in ds:
                d.step(1)
        for d in ds:
            d.synchronize()
        return (time.perf_counter() - t0) / baseline_frames

    pab = persistent_ab(b, np.concatenate([d["positions"] for d in descs]), frames) if persistent else None
    tp = per_mesh(False) if baseline_frames > 0 else None
    tg = per_mesh(True) if baseline_frames > 0 else None
    vi = float(nv) * iterations
    return {"workload": f"{count} beams, {nv} verts, {nt} tets, per-tet materials, {iterations} iters/step, "
                        f"Chebyshev 0.95, one hipGraph per frame over all meshes",
            "value": round(vi / tb / 1e6, 2), "unit": "M vertex-iters/s", "ms_per_step": round(tb * 1e3, 4),
            "frames": frames,
            "per_mesh_plain_ms_per_step": round(tp * 1e3, 3) if tp else None,
            "per_mesh_graph_ms_per_step": round(tg * 1e3, 3) if tg else None,
            "speedup_vs_per_mesh_plain": round(tp / tb, 2) if tp else None,
            "speedup_vs_per_mesh_graph": round(tg / tb, 2) if tg else None,
            "baseline_frames": baseline_frames, "roofline": roof, "persistent_ab": pab,
            "launch_config": {"lanes_per_vertex": tune[0], "pairs_per_lane": tune[1], "block_size": tune[2]}}


def persistent_ab(d, x0, frames: int, modes=(1,)):
    """The same frames from x0 with one persistent launch per sweep (vbd_set_persistent: colors
    and blend in one kernel, a device-wide barrier between colors; mode 1 the library's grid, 8k
    at most 8k blocks) against the default per-color launches, all replayed from the step's
    hipGraph: ms per frame each and whether the final positions are bitwise equal.  The handle is
    left in the default mode."""
    zero = np.zeros_like(x0)

    def run(mode):
        try:
            d.set_persistent(mode)
            d.set_positions(x0)
            d.set_velocities(zero)
            d.step(1)  # captures the step's graph
            d.synchronize()
            d.set_positions(x0)
            d.set_velocities(zero)
            d.synchronize()
            t0 = time.perf_counter()
            d.step(frames)
            d.synchronize()
            return (time.perf_counter() - t0) / frames, d.positions(), d.info().persistent_grid
        except Exception as e:  # a barrier that timed out is reported, not fatal to the bench
            return None, None, str(e)

    t0_, x0_, _ = run(0)
    res = []
    for mode in modes:
        t1_, x1_, g1 = run(mode)
        res.append({"mode": mode, "persistent_grid": g1,
                    "persistent_ms_per_step": round(t1_ * 1e3, 4) if t1_ else None,
                    "persistent_over_per_color": round(t1_ / t0_, 3) if (t0_ and t1_) else None,
                    "bitwise_equal": bool(x1_ is not None and x0_ is not None and np.array_equal(x0_, x1_))})
    d.set_persistent(0)
    out = {"frames": frames, "per_color_ms_per_step": round(t0_ * 1e3, 4) if t0_ else None,
           "kept": "per-color launches (default)"}
    if len(res) == 1:
        out.update({k: v for k, v in res[0].items() if k != "mode"})
    else:
        out["modes"] = res
    return out


def bench_persistent(caps=(256, 512, 768, 1)):
    """Persistent-sweep grid sweep on the three workloads at their default launch configuration:
    the cube128 step (20 iterations), the sphere96 step (30 iterations, contact plane z = -1.2)
    and the 256-beam batch (20 iterations), each against its per-color launches."""
    from vbd_tetmesh_amd import VBDParams, meshgen
    from vbd_tetmesh_amd.solver import DeviceMesh, prepare_host, upload
    out = {}
    m = meshgen.noisy_cube(128, seed=0)
    d = upload(m, params=VBDParams(dt=H, numIterations=20, gravity=m.gravity, useAcceleration=True,
                                   accelerationRho=0.95, mu=m.mu, lambda_=m.lam))
    out["cube128"] = persistent_ab(d, m.x0, 5, caps)
    del d
    m = meshgen.voxel_sphere(96, seed=0)
    d = upload(m, params=VBDParams(dt=H, numIterations=30, gravity=m.gravity, useAcceleration=True,
                                   accelerationRho=0.95, mu=m.mu, lambda_=m.lam, contact=(2, -1.2, 1e4)))
    out["sphere96"] = persistent_ab(d, m.x0, 10, caps)
    del d
    descs = []
    for mm in meshgen.batch_beams(256, seed=0):
        t, off, verts, dminv, vol = prepare_host(mm)
        descs.append(dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                          rest_volume=vol, color_off=off, color_verts=verts, tet_mu=t.tetMu, tet_lambda=t.tetLambda))
    b = DeviceMesh().init_batch(descs)
    b.set_params(VBDParams(dt=H, numIterations=20, gravity=(0.0, -9.8, 0.0), useAcceleration=True,
                           accelerationRho=0.95))
    b.set_launch_config(0, 256, 1)
    out["batch256"] = persistent_ab(b, np.concatenate([x["positions"] for x in descs]), 20, caps)
    return out


def launch_roofline(d, sweeps: int, traffic_name=None):
    """The color kernel against the HBM roofline: `sweeps` sweeps of the step's color launches (as
    after its first two sweeps, which alone save their start) replayed from one hipGraph between
    two HIP events (vbd_time_color_launches), the bytes the
    layout must move per launch (vbd_layout_bytes: records as stored + vertex data, gathers and
    tile padding excluded), the measured HBM traffic per launch where a rocprofv3 counter
    summary is committed, and the fixed 60 B-per-record yardstick of the reference's data model
    reported apart (the compact records hold 32 B)."""
    info = d.info()
    ms, n = d.time_color_launches(sweeps)
    lbytes, _ = d.layout_bytes()
    ybytes, _ = d.algorithmic_bytes()
    per_sweep = max(1, n // sweeps)
    sec = ms / 1e3
    achieved = lbytes / per_sweep / sec / 1e9
    kernel = "k_color_tiled" if info.lanes_per_vertex == 0 else "k_color_solve"
    traffic = read_traffic(traffic_name) if traffic_name else None
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "traffic_gbs": round(traffic / sec / 1e9, 1) if traffic else None,
           "traffic_frac": round(traffic / sec / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
           "kernel": kernel, "alg_bytes_per_launch": round(lbytes / per_sweep),
           "alg_bytes": "layout minimum: records as stored (32 B compact, or 52 B full for the per-vertex kernels "
                        "and the corotated material; +8 per-tet mu/lambda) + 56 B per free vertex (row 4, x 16 read "
                        "+ 16 written, inertia 16, mass 4)",
           "avg_launch_ms": round(ms, 4), "launches_timed": n,
           "timing": "graph replay of the step's color launches after its first two sweeps (the blends write each "
                     "sweep's start), HIP events around the replay",
           "yardstick_rate": {"bytes_per_launch": round(ybytes / per_sweep), "gbs": round(ybytes / per_sweep / sec / 1e9, 1),
                              "note": "fixed yardstick of the reference data model (45 B per vertex, 60 B per "
                                      "vertex-tet record); not bytes this kernel moves"}}
    return out


def bench_sphere(n: int = 96, frames: int = 30, iterations: int = 30, slot_order: int = 0, plane: float = -1.2,
                 roofline: bool = True, persistent: bool = True):
    """Irregular workload as configured: the 96^3 voxel sphere (interior vertices jittered, ids
    shuffled), mu = 5e4, lambda = 2e5, gravity -z, 30 iterations/step, Chebyshev 0.95, dropped
    from rest onto the penalty plane z = `plane` (k = 1e4 N/m) for 30 frames.  `value` times the
    frames in one replay; a second pass times each frame with HIP events (the handle on a torch
    stream) and splits the frames at the first one that ends with a vertex below the plane.
    With 30 iterations from x_t (VBD.cpp:72-81) this stiff mesh (k_elastic >> m/h^2 per vertex)
    sinks at about 1.5 mm per frame instead of falling freely: Gauss-Seidel removes the uniform
    translation, its lowest-frequency error, slowly.  Its bottom (z = -1) therefore never
    reaches z = -1.2 within 30 frames; a plane at z = -1.02 is reached after about 12."""
    import torch
    from vbd_tetmesh_amd import VBDParams, meshgen
    from vbd_tetmesh_amd.solver import upload
    m = meshgen.voxel_sphere(n, seed=0)
    params = VBDParams(dt=H, numIterations=iterations, gravity=m.gravity, useAcceleration=True, accelerationRho=0.95,
                       mu=m.mu, lambda_=m.lam, contact=(2, plane, 1e4))
    d = upload(m, params=params, slot_order=slot_order)
    info = d.info()
    zero = np.zeros_like(m.x0)
    d.step(1)  # captures the step's graph
    d.synchronize()
    d.set_positions(m.x0)
    d.set_velocities(zero)
    d.synchronize()
    t0 = time.perf_counter()
    d.step(frames)
    d.synchronize()
    secs = time.perf_counter() - t0
    # per-frame pass on a torch stream (torch's default stream is the null stream, which the
    # handle cannot adopt)
    d.set_positions(m.x0)
    d.set_velocities(zero)
    stream = torch.cuda.Stream()
    d.set_stream(stream)
    frame_ms, first_contact, zmin = [], None, []
    for f in range(frames):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        d.step(1)
        b.record(stream)
        b.synchronize()
        frame_ms.append(a.elapsed_time(b))
        zmin.append(float(d.positions()[:, 2].min()))
        if first_contact is None and zmin[-1] < plane:
            first_contact = f
    d.set_stream(None)
    free_ms = frame_ms[:first_contact] if first_contact is not None else frame_ms
    contact_ms = frame_ms[first_contact:] if first_contact is not None else []
    roof = launch_roofline(d, iterations, "traffic_color_tiled_sphere96") if roofline else None
    pab = persistent_ab(d, m.x0, min(frames, 10)) if (roofline and persistent) else None
    vi = float(info.num_verts) * iterations * frames
    per_frame_vi = float(info.num_verts) * iterations
    return {"workload": f"sphere {n}^3 voxels, {info.num_verts} verts, {info.num_tets} tets, {info.num_colors} colors, "
                        f"max valence {info.max_valence}, shuffled ids, dropped from rest onto the contact plane "
                        f"z={plane}, {iterations} iters/step, {frames} frames",
            "value": round(vi / secs / 1e6, 2), "unit": "M vertex-iters/s", "ms_per_step": round(secs * 1e3 / frames, 3),
            "frames": frames, "min_z_last_frame": round(zmin[-1], 4),
            "first_frame_ending_in_contact": None if first_contact is None else first_contact + 1,
            "frames_without_contact": len(free_ms), "frames_with_contact": len(contact_ms),
            "value_without_contact": round(per_frame_vi / (np.mean(free_ms) / 1e3) / 1e6, 2) if free_ms else None,
            "value_with_contact": round(per_frame_vi / (np.mean(contact_ms) / 1e3) / 1e6, 2) if contact_ms else None,
            "roofline": roof, "persistent_ab": pab,
            "launch_config": {"lanes_per_vertex": info.lanes_per_vertex, "pairs_per_lane": info.pairs_per_lane,
                              "block_size": info.block_size}}


def bench_energy(d, reps: int = 20):
    """GPU energies on the resident mesh: the log-Neo-Hookean diagnostic (VBD.cpp:258-288) and
    the incremental potential G, each a per-tet kernel plus a fixed-order double reduction,
    timed with HIP events over `reps` back-to-back evaluations."""
    res = {}
    for mode, name in ((0, "log_neo_hookean_diagnostic"), (1, "incremental_potential")):
        d.time_energy(mode, 2)
        ms, nb = d.time_energy(mode, reps)
        gbs = nb / ms / 1e6
        res[name] = {"ms_per_eval": round(ms, 4), "alg_bytes": round(nb), "achieved": round(gbs, 1),
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return res


def bench_corotated(d, m, iterations: int, steps: int = 5, cfg=None, tag: str = "cube128"):
    """The reference GPU kernel's material (corotated, VBD_GPU_Wrapper.cu:114-213, with its
    x1.000001 inertia term) on the same resident mesh: autotuned launch configuration, whole steps
    replayed from a hipGraph, the color kernel against the layout's bytes, and its measured traffic
    from the counter profile of the configuration that ran
    (profiles/traffic_color_tiled_corot_<tag>_<block>x<pairs>.txt)."""
    from vbd_tetmesh_amd import _capi as A
    info0 = d.info()
    d.set_material_model(A.VBD_MATERIAL_COROTATED_GPU)
    d.set_positions(m.x0)
    if cfg is None:
        tune = d.autotune(4)
    else:  # an explicit launch configuration (lanes, block, pairs per lane), as for the cube run
        d.set_launch_config(*cfg)
        tune = (cfg[0], cfg[2], cfg[1])
    d.set_positions(m.x0)
    d.set_velocities(np.zeros_like(m.x0))
    d.step(1)
    d.synchronize()
    t0 = time.perf_counter()
    d.step(steps)
    d.synchronize()
    secs = time.perf_counter() - t0
    info = d.info()
    x = d.positions()
    roof = launch_roofline(d, iterations,
                           f"traffic_color_tiled_corot_{tag}_{info.block_size}x{info.pairs_per_lane}"
                           if info.lanes_per_vertex == 0 else None)
    roof["kernel"] = "k_color_tiled (corotated)" if info.lanes_per_vertex == 0 else "k_color_corot"
    roof["note"] = "VALU-bound (5 polar steps per record); HBM fraction on the 52 B full records"
    d.set_material_model(A.VBD_MATERIAL_STABLE_NEO_HOOKEAN)
    d.set_launch_config(info0.lanes_per_vertex, info0.block_size, info0.pairs_per_lane)
    if not np.all(np.isfinite(x)):
        raise SystemExit("non-finite positions after the corotated steps")
    vi = float(info.num_verts) * iterations * steps
    return {"workload": "same cube, corotated material of the reference GPU kernel", "value": round(vi / secs / 1e6, 2),
            "unit": "M vertex-iters/s", "ms_per_step": round(secs * 1e3 / steps, 3), "steps": steps,
            "roofline": roof, "device_bytes": info.device_bytes, "device_bytes_after": d.info().device_bytes,
            "launch_config": {"lanes_per_vertex": tune[0], "pairs_per_lane": tune[1], "block_size": tune[2]}}


def bench_hybrid(m, iterations: int, frames: int = 5):
    """The reference main.cpp:215-232 frame loop on the device solver, host buffers every frame:
    host forwardStep (numpy), updateInertiaFromCPU (12 B per vertex H2D), solveVBD (x <- y,
    Chebyshev 0.95), copyPositionsToCPU (12 B per vertex D2H, synchronising), host
    updateVelocity.  Reports the PCIe-inclusive rate of the device part (upload + solve +
    download) and of the whole loop; neither is the headline value."""
    from vbd_tetmesh_amd.solver import VBD_GPU_Solver, prepare_host
    t, off, verts, dminv, vol = prepare_host(m)
    g = VBD_GPU_Solver()
    t.computeInertia(H)
    g.initFromCPU(t, (off, verts), dminv, vol)
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    dev = host = xfer = 0.0
    for f in range(frames + 2):  # solveVBD's first call runs eagerly, the second captures its graph
        h0 = time.perf_counter()
        t.savePreviousPosition()
        t.applyGravity(m.gravity, H)
        t.computeInertia(H)
        h1 = time.perf_counter()
        g.updateInertiaFromCPU(t)
        g.solveVBD(iterations, sdt, m.mu, m.lam, True, 0.95)
        g.copyPositionsToCPU(t)
        h2 = time.perf_counter()
        t.updateVelocity(H)
        h3 = time.perf_counter()
        if f > 1:
            dev += h2 - h1
            host += (h1 - h0) + (h3 - h2)
    # the transfers alone: one inertia upload and one position download per frame
    for f in range(frames):
        x0 = time.perf_counter()
        g.updateInertiaFromCPU(t)
        g.copyPositionsToCPU(t)
        xfer += time.perf_counter() - x0
    if not np.all(np.isfinite(t.mVertPos)):
        raise SystemExit("non-finite positions in the hybrid loop")
    vi = float(m.num_verts) * iterations * frames
    return {"workload": "same cube, reference main.cpp frame loop (VBD_GPU_Solver, pageable host arrays)",
            "device_value": round(vi / dev / 1e6, 2), "loop_value": round(vi / (dev + host) / 1e6, 2),
            "unit": "M vertex-iters/s", "frames": frames,
            "device_ms_per_frame": round(dev * 1e3 / frames, 3), "host_ms_per_frame": round(host * 1e3 / frames, 3),
            "transfer_ms_per_frame": round(xfer * 1e3 / frames, 3), "pcie_bytes_per_frame": 24 * m.num_verts}


def bench_hybrid_cpp(m, iterations: int, frames: int = 5):
    """The reference main.cpp:215-232 frame loop as a reference user compiles it: the C++ drop-in
    (include/vbd/*.h; vbd_dropin_demo's hybrid mode) with host forwardStep and updateVelocity in
    C++ (serial loops, as the reference's) and updateInertiaFromCPU, solveVBD (x <- y, Chebyshev
    0.95) and copyPositionsToCPU on the device.  The same cube is written as TetGen files (rest
    shape; the loader pins y <= min + 0.1% of the extent, the bottom face; gravity -y, h = 1/60)
    and loaded by the drop-in's loader.  Reports the host and device parts per frame over the
    frames after the first two (solveVBD's first call runs eagerly, the second captures its
    graph) and the drop-in's set-up times."""
    import shutil
    import subprocess
    import tempfile
    from vbd_tetmesh_amd import meshgen
    demo = os.path.join(ROOT, "vbd_tetmesh_amd", "vbd_dropin_demo")
    if not os.path.exists(demo):
        return {"error": "vbd_dropin_demo is not built"}
    tmp = tempfile.mkdtemp(prefix="vbd_cube_")
    try:
        t0 = time.perf_counter()
        meshgen.write_tetgen(os.path.join(tmp, "cube"), m.rest, m.tets)
        write_s = time.perf_counter() - t0
        cmd = [demo, os.path.join(tmp, "cube"), "hybrid", str(frames + 2), str(iterations), "0.95", "1", "0.001",
               repr(float(m.mu)), repr(float(m.lam)), os.devnull]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("timing ")]
    if r.returncode != 0 or not lines:
        return {"error": (r.stderr or r.stdout)[-400:]}
    tok = lines[-1].split()
    kv = {k: float(v) for k, v in zip(tok[1::2], tok[2::2])}
    dev, host = kv["device_ms_per_frame"], kv["host_ms_per_frame"]
    vi = float(m.num_verts) * iterations
    return {"workload": "same cube (rest shape, TetGen files), reference main.cpp frame loop through the C++ drop-in",
            "device_value": round(vi / (dev / 1e3) / 1e6, 2), "loop_value": round(vi / ((dev + host) / 1e3) / 1e6, 2),
            "unit": "M vertex-iters/s", "frames": int(kv["frames"]),
            "device_ms_per_frame": round(dev, 3), "host_ms_per_frame": round(host, 3),
            "setup_s": {"tetgen_load": kv["load_s"], "coloring": kv["coloring_s"], "tet_data": kv["tet_data_s"],
                        "init_from_cpu": kv["init_from_cpu_s"]},
            "tetgen_write_s": round(write_s, 2), "pcie_bytes_per_frame": 24 * m.num_verts}


def bench_adjoint(n: int = 16, steps: int = 30):
    """Differentiable workload: 16^3 cube, per-tet mu field, lambda = 4 mu, 30 converged steps;
    forward (rollout with checkpoints every 5 steps) vs backward (segment recomputation + adjoint
    solves + parameter gradients) wall time on the device.  Precision 64 (the default: every step
    converged to |grad G| < 1e-6, block-Jacobi preconditioned conjugate-gradient adjoint),
    precision 32 (FP32 state, steps end on the stagnation test near 5e-3 |M g|; colored block
    Gauss-Seidel adjoint with Chebyshev), and precision 64 with the steps converged by Newton
    steps with conjugate-gradient solves instead of VBD sweeps (fp64_newton; forward iterations
    are then Newton steps, with the CG iterations inside them beside)."""
    from vbd_tetmesh_amd import meshgen
    from vbd_tetmesh_amd.adjoint import DiffSim
    m, mu_t, lam_t, v0 = meshgen.diff_workload(n, seed=0)
    x_t = m.rest.copy()
    out = {}
    for prec, fwd in ((64, "vbd"), (32, "vbd"), (64, "newton")):
        sim = DiffSim(m, steps=steps, precision=prec, forward_method=fwd)
        sim.loss_and_grads(m.tet_mu, m.tet_lambda, v0, x_t)  # warm-up
        sim.loss_and_grads(m.tet_mu, m.tet_lambda, v0, x_t)
        st = sim.stats
        vi = float(m.num_verts) * sum(st.forward_iterations)
        out[f"fp{prec}" + ("_newton" if fwd == "newton" else "")] = {
            "forward_method": fwd,
            "forward_s": round(st.forward_seconds, 4), "backward_s": round(st.backward_seconds, 4),
            "adjoint_over_forward": round(st.backward_seconds / st.forward_seconds, 2),
            "forward_iterations_mean": round(float(np.mean(st.forward_iterations)), 1),
            "grad_G_max": float(max(st.forward_residuals)),
            "adjoint_method": sim.adj_method,
            "adjoint_iterations_mean": round(float(np.mean(st.adjoint_sweeps)), 1),
            "adjoint_rel_tol": sim.adj_rel_tol,
            "adjoint_chebyshev_rho": sim.adj_rho if sim.adj_method == "gs" else None,
            "adjoint_fallbacks": st.adjoint_fallbacks,
            "forward_inner_iterations_mean": (round(float(np.mean(st.forward_inner_iterations)), 1)
                                              if st.forward_inner_iterations else None),
            "forward_value": round(vi / st.forward_seconds / 1e6, 2) if fwd == "vbd" else None,
            "unit": "M vertex-iters/s"}
    return {"workload": f"cube {n}^3 ({m.num_verts} verts, {m.num_tets} tets), per-tet mu/lambda, {steps} converged "
                        f"steps, checkpoint every 5, gradients w.r.t. mu, lambda, v0", **out}


def copy_probe(mb: int):
    """Device copy of a known size with torch, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = mb * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="cube cells per side")
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the cube's launch configuration first (default: the library's default, which "
                         "is the autotuner's choice for this mesh; timing candidates in the same process leaves "
                         "the cube about 0.5%% slower)")
    ap.add_argument("--no-autotune", action="store_true", help="the default (kept for older command lines)")
    ap.add_argument("--lanes", type=int, default=None, help="color-kernel lanes per vertex (0: tiled kernel)")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--unroll", type=int, default=1)
    ap.add_argument("--slot-order", type=int, default=0, help="vertex order inside a color: 0 auto, 1 vertex id, 2 Morton")
    ap.add_argument("--copy-probe", type=int, default=0, help="MiB; run only the copy probe")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-per-mesh comparison")
    ap.add_argument("--batch-frames", type=int, default=60, help="timed frames of the batched workload")
    ap.add_argument("--batch-baseline-frames", type=int, default=3,
                    help="frames of the per-mesh baselines (0 skips them)")
    ap.add_argument("--batch-launch", default=None,
                    help="lanes,block,pairs: fixed launch configuration of the batch (default: autotuned)")
    ap.add_argument("--no-sphere", action="store_true", help="skip the irregular sphere workload")
    ap.add_argument("--sphere-frames", type=int, default=30, help="frames of the sphere workload (30 as configured)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable workload")
    ap.add_argument("--no-corotated", action="store_true", help="skip the corotated-material run of the cube")
    ap.add_argument("--no-hybrid", action="store_true", help="skip the host-buffer (PCIe-inclusive) frame loops")
    ap.add_argument("--no-persistent-ab", action="store_true",
                    help="skip the persistent-sweep A/B runs (for kernel traces of the default launches)")
    ap.add_argument("--only", choices=["batched", "sphere", "adjoint", "persistent"], default=None,
                    help="run one secondary workload alone and print its JSON object (for profiling)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (multi-GPU sharding "
                                   "is described but not built)"}))
        sys.exit(2)
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if ws > 1:
        print("one GPU only: WORLD_SIZE > 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from vbd_tetmesh_amd import VBDParams, meshgen
    from vbd_tetmesh_amd.solver import upload

    batch_kw = dict(frames=args.batch_frames, baseline_frames=args.batch_baseline_frames,
                    persistent=not args.no_persistent_ab,
                    launch=tuple(int(v) for v in args.batch_launch.split(",")) if args.batch_launch else None)
    if args.only:
        fn = {"batched": lambda: bench_batch(**batch_kw),
              "sphere": lambda: bench_sphere(frames=args.sphere_frames, slot_order=args.slot_order,
                                             persistent=not args.no_persistent_ab),
              "adjoint": bench_adjoint, "persistent": bench_persistent}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return

    t0 = time.time()
    m = meshgen.noisy_cube(args.n, seed=0)
    params = VBDParams(dt=H, numIterations=args.iterations, substeps=1, gravity=m.gravity, useAcceleration=True,
                       accelerationRho=0.95, mu=m.mu, lambda_=m.lam)
    d = upload(m, params=params, slot_order=args.slot_order)
    info = d.info()
    log(f"[bench] cube{args.n}: nV={info.num_verts} nT={info.num_tets} colors={info.num_colors} "
        f"pairs={info.num_pairs} device={info.device_bytes / 2**30:.2f} GiB setup {time.time() - t0:.1f}s")
    tune = None
    if args.lanes is not None:
        d.set_launch_config(args.lanes, args.block, args.unroll)
    elif args.autotune and not args.no_autotune:
        d.set_positions(m.x0)
        tune = d.autotune(8)
        d.set_positions(m.x0)
        log(f"[bench] autotune -> lanes={tune[0]} pairs/lane={tune[1]} block={tune[2]} ({tune[3]:.3f} ms/sweep)")
    info = d.info()

    # warmup
    d.step(args.warmup)
    d.synchronize()
    # timed region
    d.synchronize()
    t1 = time.perf_counter()
    d.step(args.steps)
    d.synchronize()
    t2 = time.perf_counter()
    secs = t2 - t1
    vi = float(info.num_verts) * args.iterations * args.steps
    value = vi / secs / 1e6
    log(f"[bench] {args.steps} steps in {secs * 1e3:.1f} ms -> {value:.1f} M vertex-iters/s")
    x = d.positions()
    if not np.all(np.isfinite(x)):
        raise SystemExit("non-finite positions after the timed steps")

    # dominant kernel (color solve): its launches replayed from a graph, HIP events on the
    # solver's stream
    kernel = "k_color_tiled" if info.lanes_per_vertex == 0 else "k_color_solve"
    roof = launch_roofline(d, args.iterations, f"traffic_{kernel[2:]}_cube{args.n}")
    lbytes, lblend = d.layout_bytes()
    # the step: iterations x the color launches, iterations - 1 blends, and the first two sweeps'
    # color launches also save their start (16 B per free vertex)
    its = args.iterations
    b_vi = (lbytes + lblend * (its - 1) / its + 16.0 * info.num_free * min(2, its) / its) / info.num_verts
    overall = vi / secs * b_vi / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "M vertex-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded noisy cube, generated in-process)",
        "config": {"workload": f"cube{args.n} noisy Kuhn-split cube, {info.num_verts} verts, {info.num_tets} tets, "
                               f"{args.iterations} iters/step, Chebyshev 0.95, h=1/60, bottom face pinned",
                   "iterations_per_step": args.iterations, "colors": info.num_colors,
                   "lanes_per_vertex": info.lanes_per_vertex, "pairs_per_lane": info.pairs_per_lane,
                   "block_size": info.block_size, "device_bytes": info.device_bytes},
        "roofline": roof,
        "roofline_step": {"alg_bytes_per_vertex_iter": round(b_vi, 1), "achieved": round(overall, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(overall / HBM_PEAK_GBS, 4),
                          "note": "whole timed step: layout bytes of the color launches and blends per "
                                  "vertex-iteration x the measured rate"},
    }
    if not args.no_persistent_ab:
        out["persistent_ab"] = persistent_ab(d, m.x0, args.steps)
        pa = out["persistent_ab"]
        log(f"[bench] persistent sweep: {pa['persistent_ms_per_step']} ms/step on {pa['persistent_grid']} blocks vs "
            f"per-color {pa['per_color_ms_per_step']} ms/step, bitwise {pa['bitwise_equal']}")
    out["energy"] = bench_energy(d)
    log(f"[bench] energies: " + ", ".join(f"{k} {v['ms_per_eval']:.3f} ms ({v['frac']})" for k, v in out["energy"].items()))
    if not args.no_corotated:
        cr = bench_corotated(d, m, args.iterations,
                             cfg=None if args.lanes is None else (args.lanes, args.block, args.unroll),
                             tag=f"cube{args.n}")
        out["corotated"] = cr
        log(f"[bench] corotated: {cr['value']:.1f} M vertex-iters/s, color kernel frac {cr['roofline']['frac']}")
    if not args.no_hybrid:
        del d
        hy = bench_hybrid(m, args.iterations)
        out["hybrid_pcie"] = hy
        log(f"[bench] hybrid loop: device part {hy['device_value']:.1f} M vertex-iters/s "
            f"({hy['device_ms_per_frame']:.2f} ms/frame, transfers {hy['transfer_ms_per_frame']:.2f} ms), whole loop "
            f"{hy['loop_value']:.1f}")
        hc = bench_hybrid_cpp(m, args.iterations)
        out["hybrid_cpp"] = hc
        if "error" in hc:
            log(f"[bench] C++ drop-in loop: {hc['error']}")
        else:
            log(f"[bench] C++ drop-in loop: device part {hc['device_ms_per_frame']:.2f} ms/frame "
                f"({hc['device_value']:.1f} M vertex-iters/s), host part {hc['host_ms_per_frame']:.2f} ms, whole loop "
                f"{hc['loop_value']:.1f}; set-up {hc['setup_s']}")
    if not args.no_batch:
        d = None
        bb = bench_batch(**batch_kw)
        out["batched"] = bb
        base = (f"; per-mesh plain {bb['per_mesh_plain_ms_per_step']:.2f} ms/step "
                f"({bb['speedup_vs_per_mesh_plain']:.1f}x), per-mesh graph {bb['per_mesh_graph_ms_per_step']:.2f} "
                f"ms/step ({bb['speedup_vs_per_mesh_graph']:.1f}x)" if bb['per_mesh_plain_ms_per_step'] else
                "; per-mesh baselines skipped")
        log(f"[bench] batched: {bb['value']:.1f} M vertex-iters/s, {bb['ms_per_step']:.3f} ms/step, color kernel frac "
            f"{bb['roofline']['frac']}{base}")
    if not args.no_sphere:
        sp = bench_sphere(frames=args.sphere_frames, slot_order=args.slot_order, persistent=not args.no_persistent_ab)
        # the same drop onto a plane the sinking sphere reaches within the frames
        sc = bench_sphere(frames=args.sphere_frames, slot_order=args.slot_order, plane=-1.02, roofline=False)
        sp["contact_run"] = {k: sc[k] for k in ("workload", "value", "ms_per_step", "first_frame_ending_in_contact",
                                                 "frames_without_contact", "frames_with_contact",
                                                 "value_without_contact", "value_with_contact", "min_z_last_frame")}
        out["sphere"] = sp
        cr_ = sp["contact_run"]
        log(f"[bench] sphere: {sp['value']:.1f} M vertex-iters/s, color kernel frac {sp['roofline']['frac']}; plane "
            f"z=-1.02: {cr_['frames_without_contact']} frames without contact {cr_['value_without_contact']}, "
            f"{cr_['frames_with_contact']} with {cr_['value_with_contact']}")
    if not args.no_adjoint:
        ad = bench_adjoint()
        out["adjoint"] = ad
        log("[bench] adjoint: " + "; ".join(f"{k}: forward {v['forward_s']:.3f}s backward {v['backward_s']:.3f}s "
                                             f"(x{v['adjoint_over_forward']}), |grad G| <= {v['grad_G_max']:.1e}"
                                             for k, v in ad.items() if k.startswith("fp")))
    if not args.no_cpu_baseline:
        cb = cpu_baseline(m, args.iterations)
        out["cpu_baseline"] = cb
        log(f"[bench] cpu baseline {cb['value'] * 1e6:.0f} vertex-iters/s on {cb['cores']} threads")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
