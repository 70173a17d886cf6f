"""Writes workload5_oracle.json: the CPU oracle's matchScene result on the
4M-point scene (20 instances of the 200k-point box+cylinder+torus model, 50%
occlusion, sigma 0.3 mm, 30% clutter), with SHA-256 digests of the generated
model and scene so a test can tell whether its own generator run reproduced
the same clouds.  The oracle takes about two minutes on 8 threads, too long
to repeat inside every GPU test run.

    python tests/golden/make_workload5_oracle.py
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

NUM_MATCHES = 20


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def oracle_result(w, threads):
    import oracle as O
    from surface_match_amd import synth
    mp, mn = w["model"]
    m = O.train(mp, mn, 0.04)
    t = time.time()
    P, S, st = O.match(m, w["scene"].points, None, 0.04, 0.1, 0.2,
                       param=O.default_match_param(numMatches=NUM_MATCHES),
                       vp=np.array(synth.VIEWPOINT, np.float32), threads=threads)
    return P, S, st, time.time() - t


def main():
    from surface_match_amd import synth
    w = synth.workload(5)
    threads = min(os.cpu_count() or 1, 16)
    P, S, st, sec = oracle_result(w, threads)
    out = {
        "workload": "synth.workload(5): 200k-pt model (box+cylinder+torus), 4M-pt scene",
        "match": {"samplingDistanceRel": 0.04, "keyPointFraction": 0.1, "minScore": 0.2,
                  "numMatches": NUM_MATCHES},
        "model_sha256": digest(w["model"][0]),
        "scene_sha256": digest(w["scene"].points),
        "scene_points": int(len(w["scene"].points)),
        "oracle_seconds": round(sec, 1), "oracle_threads": threads,
        "stats": {"nSampled": int(st.nSampled), "nKeypoints": int(st.nKeypoints),
                  "nPoses": int(st.nPoses), "nPairs": int(st.nPairs), "nVotes": int(st.nVotes)},
        "scores_f32_bits": [int(v) for v in np.asarray(S, np.float32).view(np.uint32)],
        "scores": [float(v) for v in S],
        "poses_3x4": [[float(v) for v in p.ravel()] for p in P],
    }
    with open(os.path.join(HERE, "workload5_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("scene_points", "oracle_seconds", "scores")}))


if __name__ == "__main__":
    main()
