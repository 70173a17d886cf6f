#!/usr/bin/env python3
"""Benchmark: matchScene latency on the 1M-point scene on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (seeded, synthetic): a 100k-point box+cylinder model trained with
samplingDistanceRel 0.04; a 1,000,000-point scene holding 10 rigid copies
(40% half-space occlusion, sigma 0.2 mm noise) plus 20% uniform clutter,
without normals (so normals are estimated from 10-NN); matchScene(0.04, 0.1,
0.2) with numMatches 10.  Each step is one complete drop-in matchScene with
the scene in host memory, as the reference's Detector::matchScene takes it:
the copy to HBM, removeNan and the bounding box, sampling, normals, key
points, voting, clustering, sparse + dense ICP and scoring.  The same match
on a scene already resident in HBM is reported beside it (`resident_ms`).

Prints one JSON line.  `roofline` covers the dominant kernel (scene voting):
algorithmic bytes per launch from the kernel's own counters divided by its
event-timed duration, with the resource that actually binds it beside.
`cpu_baseline` is the CPU restatement of the reference algorithm on the
reference's own kd-tree (oracle/cpu_ref.cpp): the whole matchScene on the
same scene, median of 5 at min(cores, 8) threads as the reference's
Detector caps OpenMP, and median of 3 at all cores.  `icp_batching` is the
batching workload's ICP figure (the 64-scene vote's own hypotheses, batched
vs one at a time).
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

HBM_PEAK_GBS = 8000.0
PROFILE_FILE = os.path.join(ROOT, "profiles", "vote_kernels_pmc_matchScene_1M.md")


def profiled(quantity: str, kernel: str):
    """A per-launch figure of `kernel` on this workload kept with the
    rocprofv3 PMC summaries (e.g. the HBM bytes from the copy-calibrated
    FETCH_SIZE and WRITE_SIZE passes); None when no such measurement is kept."""
    try:
        with open(PROFILE_FILE) as f:
            for line in f:
                if line.startswith(f"{quantity} {kernel}:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


BINDING = {
    # what the PMC counters of each vote kernel show binds it
    "vote_hits_kernel": "L1-to-register data path (TD) and VALU issue, not HBM or LDS atomics",
    "pair_kernel": "VALU issue (per-pair feature arithmetic: correctly rounded divisions, square "
                   "roots, arctangents, MurmurHash3), not HBM",
    "vote_kernel": "memory and LDS latency at one workgroup per CU",
}


def binding(kname: str, lds_rate: float, ceil: dict) -> dict:
    """The resource that binds `kname`, with the figures that show it: the
    counters kept in PROFILE_FILE for that kernel and, for the kernels that
    cast votes, the LDS atomic rate against this GPU's measured ceiling."""
    out = {"resource": BINDING[kname],
           "l1_tag_lookups_per_cu_cycle": profiled("l1_tag_lookups_per_cu_cycle", kname),
           "valu_busy_frac": profiled("valu_busy_frac", kname),
           "alone_ms": profiled("alone_ms", kname),
           "profile": os.path.relpath(PROFILE_FILE, ROOT)}
    if kname != "pair_kernel":
        out.update({"lds_atomics_per_s": float("%.4g" % lds_rate),
                    "lds_atomic_ceiling_per_s": ceil["lds_atomics_per_s"],
                    "lds_frac": round(lds_rate / ceil["lds_atomics_per_s"], 4)})
    return out


def vote_algorithmic_bytes(st: dict) -> float:
    """The whole vote stage, per launch set: 24 B per neighbour pair (position
    + normal), 12 B per voted key point, 8 B per key hit (bucket descriptor),
    8 B per vote (bucket entry: ref + alpha), 20 B per candidate written."""
    return (24.0 * st["nPairs"] + 12.0 * st["nVotedKeypoints"] + 8.0 * st["nKeyHits"] +
            8.0 * st["nVotes"] + 20.0 * st["nPoses"])


def pair_algorithmic_bytes(st: dict) -> float:
    """The pair kernel: 24 B per neighbour pair (position + normal), 16 B per
    pair for its key's table slot, 24 B per key point (position + normal),
    12 B per key hit written (bucket start, length, scene alpha)."""
    return (24.0 * st["nPairs"] + 16.0 * st["nPairs"] + 24.0 * st["nKeypoints"] +
            12.0 * st["nKeyHits"])


def hits_algorithmic_bytes(st: dict) -> float:
    """The hit-voting kernel: 12 B per key hit read, 8 B per vote (bucket
    entry), 24 B of candidates per voted key point."""
    return 12.0 * st["nKeyHits"] + 8.0 * st["nVotes"] + 24.0 * st["nVotedKeypoints"]


def ply_ingest(det, points) -> dict:
    """The 1M-point scene as a binary PLY file: parse + upload into a
    resident scene (ppf_scene_upload_ply), against the upload of the same
    points from memory."""
    import tempfile
    from surface_match_amd import PointCloud, ResidentScene, writePLY, synth
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "scene.ply")
        assert writePLY(path, PointCloud(points))
        mb = os.path.getsize(path) / 1e6
        cloud = PointCloud(points)
        cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        best_f = best_m = 1e9
        for _ in range(3):
            t0 = time.perf_counter()
            ResidentScene(det, ply=path, viewPoint=synth.VIEWPOINT)
            t1 = time.perf_counter()
            ResidentScene(det, cloud)
            t2 = time.perf_counter()
            best_f, best_m = min(best_f, t1 - t0), min(best_m, t2 - t1)
    return {"file_MB": round(mb, 1), "points": int(len(points)),
            "ply_to_resident_ms": round(best_f * 1e3, 2),
            "ply_MB_per_s": round(mb / best_f, 1),
            "arrays_to_resident_ms": round(best_m * 1e3, 2)}


MESH_GRID, MESH_RADIUS, MESH_SEED = 1001, 0.0023, 0x5EED


def mesh_sampling() -> dict:
    """sampleMesh on a 2M-triangle mesh (a 1001 x 1001 height field) at a
    radius that yields about 1M points, host arrays in and out: one
    ppf_sample_mesh call into buffers of the known size (copies in, area
    pass, host CDF, sampling kernel, copies out; best of 3), and the Python
    sampleMesh, which first asks for the size."""
    import ctypes as C
    from surface_match_amd import PointCloud, sampleMesh, synth, _native as N
    P, F = synth.height_field_mesh(MESH_GRID, 5)
    pc = PointCloud(P)
    pc.face = F
    t0 = time.perf_counter()
    out = sampleMesh(pc, MESH_RADIUS, MESH_SEED)
    t_py = time.perf_counter() - t0
    n = len(out.point)
    OP, ON = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
    got = C.c_size_t()
    best = 1e9
    for _ in range(3):
        t0 = time.perf_counter()
        N.check(N.lib().ppf_sample_mesh(N.ptr(P), len(P), N.ptr(F), len(F), MESH_RADIUS,
                                        MESH_SEED, N.ptr(OP), N.ptr(ON), n, C.byref(got)))
        best = min(best, time.perf_counter() - t0)
    assert got.value == n and np.array_equal(OP, out.point)
    return {"faces": int(len(F)), "points": n, "ms": round(best * 1e3, 2),
            "points_per_s": float("%.4g" % (n / best)),
            "python_sampleMesh_ms": round(t_py * 1e3, 2)}


MLS_N, MLS_STRIDE, MLS_RADIUS, MLS_SIGMA = 1_000_000, 4, 0.03, 0.004


def mls_cloud():
    """1M points on a unit sphere with radial noise sigma 0.004 (seeded)"""
    r = np.random.Generator(np.random.PCG64(8))
    d = r.standard_normal((MLS_N, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * (1.0 + MLS_SIGMA * r.standard_normal((MLS_N, 1)))).astype(np.float32)


def mls_normals() -> dict:
    """estimateNormalMLS, order 2, radius 0.03 on the 1M-point noisy sphere
    for every 4th point (250k queries); device time of search + sort + fit
    (best of 3), and the whole call with host arrays in and out."""
    from surface_match_amd import PointCloud, estimateNormalMLS
    P = mls_cloud()
    q = np.arange(0, MLS_N, MLS_STRIDE)
    best, best_call, st = 1e9, 1e9, None
    for _ in range(3):
        pc = PointCloud(P.copy())
        pc.viewPoint = np.zeros(3, np.float32)
        t0 = time.perf_counter()
        st = estimateNormalMLS(pc, q, MLS_RADIUS, 2, invert=True)
        best_call = min(best_call, time.perf_counter() - t0)
        best = min(best, st["device_ms"])
    return {"points": MLS_N, "queries": int(len(q)), "radius": MLS_RADIUS, "order": 2,
            "neighbours": st["neighbours"], "device_ms": round(best, 3),
            "call_ms": round(best_call * 1e3, 2),
            "neighbours_per_s": float("%.4g" % (st["neighbours"] / (best * 1e-3)))}


def cpu_mls_normals(n_queries: int = 5000) -> dict:
    """the oracle's estimateNormalMLS (one thread) on the first queries of
    the same workload"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    P = mls_cloud()
    q = np.arange(0, MLS_N, MLS_STRIDE)[:n_queries]
    t0 = time.perf_counter()
    O.estimate_normals_mls(P, np.full_like(P, np.nan), np.arange(MLS_N), q, MLS_RADIUS, 2,
                           invert=1, vp=np.zeros(3, np.float32))
    ms = (time.perf_counter() - t0) * 1e3
    return {"queries": n_queries, "ms": round(ms, 1),
            "ms_per_250k_queries": round(ms * (MLS_N // MLS_STRIDE) / n_queries, 1)}


def cpu_mesh_sampling() -> float:
    """the same sampleMesh call by the C oracle (one thread), ms"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from surface_match_amd import synth
    P, F = synth.height_field_mesh(MESH_GRID, 5)
    t0 = time.perf_counter()
    O.sample_mesh(P, F, MESH_RADIUS, MESH_SEED)
    return (time.perf_counter() - t0) * 1e3


def measured_ceilings() -> dict:
    """This GPU's streaming-copy bandwidth and LDS atomic rate, measured by
    the library's own microbenchmarks (the roofline's peak stays the
    datasheet's 8 TB/s)."""
    import ctypes as C
    from surface_match_amd import _native as N
    gbps, ats = C.c_float(0), C.c_float(0)
    N.check(N.lib().ppf_measure_ceilings(C.byref(gbps), C.byref(ats)))
    return {"copy_GBps": round(gbps.value, 1), "lds_atomics_per_s": float("%.4g" % ats.value)}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w):
    """The CPU restatement of the reference algorithm on its own data
    structures (oracle/cpu_ref.cpp: the reference's nanoflann kd-tree,
    parallel greedy sampling, kNN normals, a hash map, the OpenMP vote loop;
    built -O3 -mavx -fopenmp like the reference) on the same 1M-point scene,
    the whole matchScene: one warm-up, then the median of 5 runs at
    min(nproc, 8) threads (the reference's Detector caps OpenMP at 8) and of 3
    at nproc threads (about 30 s of CPU work)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from surface_match_amd import synth
    m = O.train(*w["model"], 0.04)
    cm = O.CpuRefModel(m)
    param = O.default_match_param(numMatches=10)
    pts = w["scene"].points

    def one(threads):
        _, _, st = cm.match(pts, 0.04, 0.1, 0.2, param, vp=synth.VIEWPOINT, threads=threads)
        return st

    nproc = os.cpu_count() or 1
    t8 = min(nproc, 8)
    one(t8)  # warm-up
    out = {}
    for threads, reps in ((t8, 5), (nproc, 3)):
        if threads in out:
            continue
        runs = [one(threads) for _ in range(reps)]
        tot = [r["stage_ms"]["total"] for r in runs]
        med = runs[int(np.argsort(tot)[len(tot) // 2])]
        out[threads] = {"ms": float(np.median(tot)), "runs_ms": [round(t, 1) for t in tot],
                        "stage_ms": med["stage_ms"],
                        "work": {k: med[k] for k in ("nSampled", "nKeypoints", "nVotedKeypoints",
                                                      "nPairs", "nVotes", "nPoses", "nResults")}}
    return out


def icp_batching(det, H_per_instance=64, n_inst=8, seed=0x5EED0003):
    """Batched vs plain GPU point-to-plane ICP on the same hypotheses."""
    import math
    from surface_match_amd import synth
    from surface_match_amd.icp import IcpTarget
    r = np.random.Generator(np.random.PCG64(seed))
    _, _, src, srcn = det.sampled_model()          # the trained resampled model
    tp, tn, gts = [], [], []
    for k in range(n_inst):
        p, n = synth.make_model(40000, seed + 10 + k)
        R = synth.random_rotation(r)
        t = r.uniform(-250, 250, 3)
        tp.append(p @ R.T + t)
        tn.append(n @ R.T)
        T = np.eye(4)
        T[:3, :3], T[:3, 3] = R, t
        gts.append(T)
    dst = (np.concatenate(tp) + r.normal(scale=0.1, size=(40000 * n_inst, 3))).astype(np.float32)
    dstn = np.concatenate(tn).astype(np.float32)
    inits, owner = [], []
    for k in range(n_inst):
        for _ in range(H_per_instance):
            w = np.radians(r.normal(scale=2.0, size=3))
            th = np.linalg.norm(w)
            a = w / th
            K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
            P = np.eye(4)
            P[:3, :3] = np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K
            P[:3, 3] = r.normal(scale=2.0, size=3)
            inits.append(P @ gts[k])
            owner.append(k)
    inits = np.array(inits)
    tgt = IcpTarget(dst, dstn, 2.0)
    D = det.model_info().diameter
    kw = dict(iterations=10, reject_dist=0.1 * D, mse_min=0.0, mse_max=1e30, tolerance=0.0,
              mode=1, scoring_dist=0.01 * D)
    for _ in range(2):                                         # warm-up; the second captures the graph
        tgt.refine(src, srcn, inits, batched=True, **kw)
    tgt.refine(src, srcn, inits, batched="persistent", **kw)
    tgt.refine(src, srcn, inits[:8], batched=False, **kw)
    t0 = time.perf_counter()
    pb, _, msb = tgt.refine(src, srcn, inits, batched=True, **kw)
    wb = time.perf_counter() - t0
    t0 = time.perf_counter()
    pq, _, msq = tgt.refine(src, srcn, inits, batched="persistent", **kw)
    wq = time.perf_counter() - t0
    nplain = 64
    t0 = time.perf_counter()
    pp, _, msp = tgt.refine(src, srcn, inits[:nplain], batched=False, **kw)
    wp = time.perf_counter() - t0

    def ok(P, k):
        ang = synth.rotation_angle_deg(gts[k][:3, :3].T @ P[:3, :3])
        return ang <= 0.1 and np.linalg.norm(gts[k][:3, 3] - P[:3, 3]) <= 0.001 * D
    fb = float(np.mean([ok(pb[h], owner[h]) for h in range(len(inits))]))
    fp = float(np.mean([ok(pp[h], owner[h]) for h in range(nplain)]))
    diff = max(synth.rotation_angle_deg(pb[h][:3, :3].T @ pp[h][:3, :3]) for h in range(nplain))
    return {"hypotheses": len(inits), "source_points": int(len(src)), "iterations": 10,
            "mode": "point-to-plane", "batched_poses_per_s": round(len(inits) / wb, 1),
            "plain_poses_per_s": round(nplain / wp, 1),
            "persistent_poses_per_s": round(len(inits) / wq, 1),
            "persistent_device_ms": round(msq, 3),
            "max_persistent_vs_spread_deg": round(float(max(
                synth.rotation_angle_deg(pb[h][:3, :3].T @ pq[h][:3, :3]) for h in range(len(inits)))), 6),
            "speedup": round((len(inits) / wb) / (nplain / wp), 2),
            "batched_device_ms": round(msb, 3), "plain_device_ms_per_pose": round(msp / nplain, 4),
            "batched_within_0.1deg_0.1pctD": fb, "plain_within_0.1deg_0.1pctD": fp,
            "max_batched_vs_plain_deg": round(float(diff), 6)}


def scene_batching_in_child():
    """scene_batching() run by `bench.py --only batch` in a child process."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--only", "batch"],
                       capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"scene batching child failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def scene_batching(seed=0x5EED0003, n_scenes=64, one_by_one=True):
    """Many scenes of uneven size matched together (one vote launch for all
    key points, one batched refinement per stage) vs one matchScene each,
    with the box+cylinder model of 20k points.  one_by_one=False times the
    batched call alone (no one-by-one leg, no refinement side measurement),
    e.g. to profile it."""
    from surface_match_amd import Detector, PointCloud, MatchParam, ResidentScene, synth
    mp, mn = synth.make_model(20000, 0x5EED0001)
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    D = det.model_info().diameter
    scenes = synth.batch_scenes(seed, n_scenes)
    rs = []
    for sc in scenes:
        c = PointCloud(sc.points)
        c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        rs.append(ResidentScene(det, c))
    param = MatchParam(numMatches=8)
    ResidentScene.match_many(rs, 0.04, 0.1, 0.2, param)       # warm-up
    if not one_by_one:
        tb = []
        for _ in range(3):
            t0 = time.perf_counter()
            _, st = ResidentScene.match_many(rs, 0.04, 0.1, 0.2, param)
            tb.append(time.perf_counter() - t0)
        return {"scenes": n_scenes, "batched_ms": round(float(np.median(tb)) * 1e3, 2),
                "batched_stage_ms": {k: round(float(st[k]), 3) for k in
                                     ("msSample", "msVote", "msCluster", "msIcp", "msTotal")}}
    single = [r.match(0.04, 0.1, 0.2, param) for r in rs]
    tb, ts = [], []
    for _ in range(3):                                         # median of 3 each
        t0 = time.perf_counter()
        many, st = ResidentScene.match_many(rs, 0.04, 0.1, 0.2, param)
        tb.append(time.perf_counter() - t0)
    for _ in range(3):
        t0 = time.perf_counter()
        single = [r.match(0.04, 0.1, 0.2, param) for r in rs]
        ts.append(time.perf_counter() - t0)
    wb, ws = float(np.median(tb)), float(np.median(ts))
    same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
               for a, b in zip(many, single))
    icp = vote_hypotheses_icp(det, rs, scenes)
    found = tot = 0
    for sc, (P, S) in zip(scenes, many):
        tot += len(sc.gt_poses)
        hit = set()
        for T in P:
            k, ang, dt = synth.pose_error(T, sc.gt_poses, D)
            if ang <= 1.0 and dt <= 0.01 * D:
                hit.add(k)
        found += len(hit)
    return {"scenes": n_scenes, "points": int(sum(len(s.points) for s in scenes)),
            "instances": tot, "keypoints": int(st["nKeypoints"]),
            "hypotheses_refined": int(st["nSparseIcp"]), "dense_refined": int(st["nDenseIcp"]),
            "batched_ms": round(wb * 1e3, 2), "one_by_one_ms": round(ws * 1e3, 2),
            "speedup": round(ws / wb, 2), "scenes_per_s": round(n_scenes / wb, 1),
            "results_identical": bool(same),
            "instances_found_within_1deg_1pctD": found,
            "batched_stage_ms": {k: round(float(st[k]), 3) for k in
                                 ("msSample", "msVote", "msVoteKernel", "msCluster", "msIcp",
                                  "msTotal")},
            "icp": icp}


def vote_hypotheses_icp(det, rs, scenes):
    """The batching workload's ICP comparison: every pose hypothesis the
    64-scene vote and clustering produce (refinement off, key point
    fraction 0.2), refined by point-to-plane ICP (10 iterations) against its
    own scene's sampled cloud with the normals the match estimated -- all of
    them in one batched refinement across the scenes, and one at a time with
    the conventional per-hypothesis GPU refiner.  Both give the same poses;
    the pose error of each against the nearest ground-truth instance."""
    from surface_match_amd import MatchParam, MatchResult, PointCloud, ResidentScene, synth
    from surface_match_amd.icp import IcpTarget, refine_many
    D = float(det.model_info().diameter)
    sp, sn, _, _ = det.sampled_model()
    param = MatchParam(numMatches=100000, sparsePoseRefinement=0, densePoseRefinement=0)
    hyps, _ = ResidentScene.match_many(rs, 0.04, 0.2, 0.2, param)
    targets, owner, inits = [], [], []
    for s_, (sc, (P, _)) in enumerate(zip(scenes, hyps)):
        cloud = PointCloud(sc.points)
        cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        mr = MatchResult()
        det.matchScene(cloud, 0.04, 0.2, 0.2, param, matchResult=mr)
        targets.append(IcpTarget(mr.sampledScene.point, mr.sampledScene.normal, 0.08 * D))
        owner += [s_] * len(P)
        inits += list(P)
    owner = np.array(owner, np.int32)
    inits = np.array(inits)
    kw = dict(iterations=10, reject_dist=0.1 * D, mse_min=0.0, mse_max=1e30, tolerance=0.0,
              mode=1, scoring_dist=0.01 * D)
    for _ in range(2):                                 # warm-up; the second captures the graph
        refine_many(targets, owner, sp, sn, inits, batched=True, **kw)
    t0 = time.perf_counter()
    pb, _, msb = refine_many(targets, owner, sp, sn, inits, batched=True, **kw)
    wb = time.perf_counter() - t0
    nplain = min(len(inits), 256)
    refine_many(targets, owner[:4], sp, sn, inits[:4], batched=False, **kw)
    t0 = time.perf_counter()
    pp, _, msp = refine_many(targets, owner[:nplain], sp, sn, inits[:nplain], batched=False, **kw)
    wp = time.perf_counter() - t0
    errs = [synth.pose_error(P, scenes[o].gt_poses, D) for P, o in zip(pb, owner)]
    within = np.array([ang <= 0.1 and dt <= 0.001 * D for _, ang, dt in errs])
    start = [synth.pose_error(P, scenes[o].gt_poses, D) for P, o in zip(inits, owner)]
    near = np.array([ang <= 5.0 and dt <= 0.05 * D for _, ang, dt in start])
    diff = max(synth.rotation_angle_deg(pb[h][:3, :3].T @ pp[h][:3, :3]) for h in range(nplain))
    dt_diff = max(float(np.abs(pb[h][:3, 3] - pp[h][:3, 3]).max()) for h in range(nplain))
    return {"hypotheses": len(inits), "scenes": len(scenes), "source_points": int(len(sp)),
            "iterations": 10, "mode": "point-to-plane",
            "batched_poses_per_s": round(len(inits) / wb, 1),
            "plain_poses_per_s": round(nplain / wp, 1), "plain_timed_on": nplain,
            "speedup": round((len(inits) / wb) / (nplain / wp), 2),
            "batched_device_ms": round(msb, 3),
            "within_0.1deg_0.1pctD_of_an_instance": round(float(np.mean(within)), 4),
            "started_within_5deg_5pctD": int(near.sum()),
            "of_those_within_0.1deg_0.1pctD": round(float(np.mean(within[near])), 4)
            if near.any() else None,
            "max_batched_vs_plain_deg": round(diff, 6), "max_batched_vs_plain_mm": dt_diff}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-icp-batching", action="store_true")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-scene-batching", action="store_true")
    ap.add_argument("--lean", action="store_true",
                    help="the matchScene measurement only (no side measurements), e.g. to profile")
    ap.add_argument("--only", choices=["mls", "mesh", "batch", "batch-call"],
                    help="run one side measurement (MLS normals, mesh sampling, the 64-scene "
                         "batch, or the batched call of the 64 scenes alone) and print its "
                         "JSON, e.g. to profile it")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the PPF path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)",
              file=sys.stderr)
        sys.exit(2)
    if args.only:
        side = {"mls": mls_normals, "mesh": mesh_sampling, "batch": scene_batching,
                "batch-call": lambda: scene_batching(one_by_one=False)}
        print(json.dumps(side[args.only]()))
        return

    from surface_match_amd import Detector, PointCloud, MatchParam, ResidentScene, synth

    w = synth.workload(2)
    mp, mn = w["model"]
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    cloud = PointCloud(w["scene"].points)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    scene = ResidentScene(det, cloud)
    param = MatchParam(numMatches=10)
    info = det.model_info()

    # launch shapes: the library's defaults, or (--autotune) those the
    # in-project autotuner picks on this scene; results do not depend on them
    if args.autotune:
        scene.autotune(0.04, 0.1, 0.2, param)
    for _ in range(args.warmup):
        det.matchScene(cloud, 0.04, 0.1, 0.2, param)
        scene.match(0.04, 0.1, 0.2, param)
    times, stats = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        poses, scores = det.matchScene(cloud, 0.04, 0.1, 0.2, param)
        times.append(time.perf_counter() - a)
        stats.append(det.last_stats)
    total = time.perf_counter() - t0
    ms_per_step = total * 1e3 / args.steps
    # the same match on the scene resident in HBM (no upload)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        rposes, rscores, _ = scene.match(0.04, 0.1, 0.2, param)
    resident_ms = (time.perf_counter() - t1) * 1e3 / args.steps
    assert np.array_equal(rposes, poses) and np.array_equal(rscores, scores)

    st = stats[-1]
    vote_ms = float(np.mean([s["msVoteKernel"] for s in stats]))
    pair_ms = float(np.mean([s["msPairKernel"] for s in stats]))
    hits_ms = float(np.mean([s["msHitsKernel"] for s in stats]))
    # the dominant kernel of the path (the vote stage's longer kernel on the
    # split path, else the fused vote kernel)
    if pair_ms > 0 and pair_ms >= hits_ms:
        kname, kms, alg = "pair_kernel", pair_ms, pair_algorithmic_bytes(st)
    elif pair_ms > 0:
        kname, kms, alg = "vote_hits_kernel", hits_ms, hits_algorithmic_bytes(st)
    else:
        kname, kms, alg = "vote_kernel", vote_ms, vote_algorithmic_bytes(st)
    achieved = alg / (kms * 1e-3) / 1e9
    ceil = measured_ceilings()
    lds_rate = st["nVotes"] / ((hits_ms or vote_ms) * 1e-3)  # one LDS atomic add per vote
    D = info.diameter
    found = []
    for P in poses:
        k, ang, dt = synth.pose_error(P, w["scene"].gt_poses, D)
        found.append(dict(instance=k, rot_deg=round(ang, 3), trans_mm=round(dt, 3)))

    out = {
        "metric": "matchScene ms @1M-pt scene",
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded box+cylinder model, 1M-pt scene: 10 instances, 40% occlusion, "
                "sigma 0.2 mm, 20% clutter)",
        "config": {
            "workload": "matchScene, 100k-pt model (trainModel 0.04), 1M-pt scene without normals "
                        "in host memory, matchScene(0.04, 0.1, 0.2), numMatches 10",
            "scene_points": int(st["nScene"]), "sampled": int(st["nSampled"]),
            "keypoints": int(st["nKeypoints"]), "pairs": int(st["nPairs"]),
            "votes": int(st["nVotes"]), "poses": int(st["nPoses"]),
            "hypotheses": int(st["nHypotheses"]), "model_refs": int(info.nSampled),
        },
        "resident_ms": round(resident_ms, 3),
        "best_step_ms": round(min(times) * 1e3, 3),
        "stage_ms": {k: round(float(np.mean([s[k] for s in stats])), 3) for k in
                     ("msIngest", "msSample", "msNormals", "msKeypoints", "msVote", "msVoteKernel",
                      "msPairKernel", "msHitsKernel", "msCluster", "msIcpPrepare", "msIcp",
                      "msIcpSparse", "msIcpDense", "msTotal")},
        "vote_fallback_keypoints": int(st["nVoteFallback"]),
        "vote_throughput": {"keypoints_per_s": round(st["nKeypoints"] / (vote_ms * 1e-3)),
                            "votes_per_s": round(st["nVotes"] / (vote_ms * 1e-3)),
                            # one LDS atomic add per vote (the count slot is derived)
                            "lds_atomics_per_s": round(st["nVotes"] / ((hits_ms or vote_ms) * 1e-3))},
        # the dominant kernel against HBM (its algorithmic bytes; MFMA is
        # unused), with what the PMC counters show binds it
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled("traffic_bytes_per_launch", kname),
                     "kernel": kname, "algorithmic_bytes": alg, "kernel_ms": round(kms, 4),
                     "binding": binding(kname, lds_rate, ceil),
                     "vote_stage": {"algorithmic_bytes": vote_algorithmic_bytes(st),
                                    "ms": round(vote_ms, 4),
                                    "GBps": round(vote_algorithmic_bytes(st) / (vote_ms * 1e-3) / 1e9, 1)},
                     # both kernels of the split vote path (they take about the
                     # same time, so either can be the dominant one in a run)
                     "vote_kernels": {k: {"ms": round(ms_, 4), "algorithmic_bytes": b,
                                          "GBps": round(b / (ms_ * 1e-3) / 1e9, 1),
                                          "frac": round(b / (ms_ * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                          "binding": BINDING[k]}
                                      for k, ms_, b in (("pair_kernel", pair_ms, pair_algorithmic_bytes(st)),
                                                        ("vote_hits_kernel", hits_ms,
                                                         hits_algorithmic_bytes(st)))
                                      if ms_ > 0}},
        "results": found,
        "tuning": det.get_tuning(),
    }
    out["measured_ceilings"] = ceil
    if not args.lean:
        out["ply_ingest"] = ply_ingest(det, w["scene"].points)
        out["mesh_sampling"] = mesh_sampling()
        out["mls_normals"] = mls_normals()
    if not args.no_scene_batching and not args.lean:
        # in a process of its own, with one detector, as a process serving
        # batches runs it (after this process's other measurements it
        # measured 49 instead of 46 ms: its refinement preparation overlapped
        # its vote less)
        out["scene_batching"] = scene_batching_in_child()
        # the batching configuration's ICP figure: the 64-scene vote's own hypotheses
        out["icp_batching"] = out["scene_batching"].pop("icp")
    if not args.no_icp_batching and not args.lean:
        out["icp_batching_perturbed"] = icp_batching(det)
    if not args.no_cpu_baseline and not args.lean:
        cb = cpu_baseline(w)
        t8 = min(os.cpu_count() or 1, 8)
        out["cpu_baseline"] = {
            "value": round(cb[t8]["ms"], 1), "unit": "ms", "cores": t8, "kind": "port",
            "label": "CPU restatement of the reference algorithm on the reference's own kd-tree "
                     "(nanoflann), greedy sampling and a hash map; -O3 -mavx -fopenmp",
            "sample": "the full 1M-point matchScene (numMatches 10); one warm-up, median of 5",
            "runs_ms": cb[t8]["runs_ms"], "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "stage_ms": cb[t8]["stage_ms"], "work": cb[t8]["work"],
            "all_cores": {"cores": max(cb), "value": round(cb[max(cb)]["ms"], 1),
                          "runs_ms": cb[max(cb)]["runs_ms"], "stage_ms": cb[max(cb)]["stage_ms"]},
            "mesh_sampling_ms": round(cpu_mesh_sampling(), 1),
            "mls_normals": cpu_mls_normals()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
