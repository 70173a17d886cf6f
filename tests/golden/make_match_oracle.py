"""Writes the CPU oracle's matchScene results for the large synthetic
workloads, with SHA-256 digests of the generated clouds so a test can tell
whether its own generator run reproduced the same inputs.  The oracle takes
seconds to minutes per workload on 8 threads, too long to repeat inside every
GPU test run.

    python tests/golden/make_match_oracle.py 2       # workload2_oracle.json
    python tests/golden/make_match_oracle.py 5       # workload5_oracle.json
    python tests/golden/make_match_oracle.py batch   # batch64_oracle.json

Workload 2: a 100k-point box+cylinder model, a 1M-point scene of 10
instances (40% occlusion, sigma 0.2 mm, 20% clutter), numMatches 10.
Workload 5: a 200k-point box+cylinder+torus model, a 4M-point scene of 20
instances (50% occlusion, sigma 0.3 mm, 30% clutter), numMatches 20.
batch: the 20k-point box+cylinder model against 64 scenes of log-uniform
50k-500k points with 1-8 instances each (`synth.batch_scenes(0x5EED0003)`),
numMatches 8, one oracle match per scene.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

MATCH = {"samplingDistanceRel": 0.04, "keyPointFraction": 0.1, "minScore": 0.2}
NUM_MATCHES = {2: 10, 5: 20}
DESCR = {2: "synth.workload(2): 100k-pt model (box+cylinder), 1M-pt scene",
         5: "synth.workload(5): 200k-pt model (box+cylinder+torus), 4M-pt scene"}
BATCH_SEED, BATCH_SCENES, BATCH_MODEL = 0x5EED0003, 64, (20000, 0x5EED0001)
BATCH_NUM_MATCHES = 8


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def threads():
    return min(os.cpu_count() or 1, 16)


def oracle_match(O, m, points, num_matches):
    from surface_match_amd import synth
    t = time.time()
    P, S, st = O.match(m, points, None, MATCH["samplingDistanceRel"], MATCH["keyPointFraction"],
                       MATCH["minScore"], param=O.default_match_param(numMatches=num_matches),
                       vp=np.array(synth.VIEWPOINT, np.float32), threads=threads())
    return P, S, st, time.time() - t


def result_record(P, S, st):
    return {
        "stats": {"nSampled": int(st.nSampled), "nKeypoints": int(st.nKeypoints),
                  "nPoses": int(st.nPoses), "nPairs": int(st.nPairs), "nVotes": int(st.nVotes)},
        "scores_f32_bits": [int(v) for v in np.asarray(S, np.float32).view(np.uint32)],
        "scores": [float(v) for v in S],
        "poses_3x4": [[float(v) for v in p.ravel()] for p in P],
    }


def workload_fixture(idx: int):
    import oracle as O
    from surface_match_amd import synth
    w = synth.workload(idx)
    mp, mn = w["model"]
    m = O.train(mp, mn, 0.04)
    P, S, st, sec = oracle_match(O, m, w["scene"].points, NUM_MATCHES[idx])
    out = {"workload": DESCR[idx], "match": dict(MATCH, numMatches=NUM_MATCHES[idx]),
           "model_sha256": digest(mp), "scene_sha256": digest(w["scene"].points),
           "scene_points": int(len(w["scene"].points)), "oracle_seconds": round(sec, 1),
           "oracle_threads": threads()}
    out.update(result_record(P, S, st))
    with open(os.path.join(HERE, f"workload{idx}_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("scene_points", "oracle_seconds", "scores")}))


def batch_fixture():
    import oracle as O
    from surface_match_amd import synth
    mp, mn = synth.make_model(*BATCH_MODEL)
    m = O.train(mp, mn, 0.04)
    scenes = synth.batch_scenes(BATCH_SEED, BATCH_SCENES)
    recs, total = [], 0.0
    for i, sc in enumerate(scenes):
        P, S, st, sec = oracle_match(O, m, sc.points, BATCH_NUM_MATCHES)
        total += sec
        r = {"scene_sha256": digest(sc.points), "scene_points": int(len(sc.points)),
             "instances": len(sc.gt_poses)}
        r.update(result_record(P, S, st))
        recs.append(r)
        print(f"scene {i}: {len(sc.points)} points, {len(S)} results, {sec:.1f} s", flush=True)
    out = {"workload": "20k-pt box+cylinder model; synth.batch_scenes(0x5EED0003, 64)",
           "match": dict(MATCH, numMatches=BATCH_NUM_MATCHES), "model_sha256": digest(mp),
           "oracle_seconds": round(total, 1), "oracle_threads": threads(), "scenes": recs}
    with open(os.path.join(HERE, "batch64_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    arg = sys.argv[1] if len(sys.argv) > 1 else "2"
    if arg == "batch":
        batch_fixture()
    else:
        workload_fixture(int(arg))
