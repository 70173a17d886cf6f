"""Write murmur_kat.json: MurmurHash3_x86_32 known answers for 16-byte keys
of four little-endian int32 words (the PPF hash key, util.cpp:690-695) with
seed 42, computed by scikit-learn's murmurhash3_32 (an independent
implementation of the published algorithm).

    python tests/golden/make_murmur_kat.py
"""
import json
import os

import numpy as np
from sklearn.utils import murmurhash3_32


def main():
    r = np.random.Generator(np.random.PCG64(42))
    keys = [[8, 8, 0, 5], [0, 0, 0, 0], [15, 15, 15, 25], [-2147483648, 0, 0, 0],
            [1, 2, 3, 4], [-1, -1, -1, -1]]
    keys += r.integers(0, 16, (40, 3)).tolist()
    keys = [k if len(k) == 4 else k + [int(r.integers(0, 26))] for k in keys]
    keys += r.integers(-2 ** 31, 2 ** 31, (20, 4)).tolist()
    out = []
    for k in keys:
        b = np.array(k, dtype="<i4").tobytes()
        out.append({"key": [int(x) for x in k], "hash": int(murmurhash3_32(b, seed=42, positive=True))})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "murmur_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
