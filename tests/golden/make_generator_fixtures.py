"""Write tests/golden/generator.json: checksums that pin the synthetic input generator.

The generator is this project's own input spec (the reference renders scenes), so its
fixtures are produced by the oracle's C generator and committed; the device generator must
reproduce them bit for bit.  Run:  python tests/golden/make_generator_fixtures.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

SEED0 = 0x43505653
CASES = [
    dict(N=64, seed=SEED0, boxes=4, noise=False, x0=0, y0=0, tile=64),
    dict(N=256, seed=SEED0 + 1, boxes=8, noise=True, x0=0, y0=0, tile=256),
    dict(N=1024, seed=SEED0, boxes=32, noise=False, x0=0, y0=0, tile=1024),
    dict(N=2048, seed=SEED0 + 1, boxes=128, noise=True, x0=512, y0=1024, tile=512),
]
UNIFORM = [dict(seed=SEED0, count=4096, stride=4), dict(seed=7, count=1000, stride=3)]


def main():
    out = {"depth": [], "uniform": []}
    for c in CASES:
        D = O.gen_depth(c["N"], c["seed"], c["boxes"], c["noise"], c["x0"], c["y0"], c["tile"])
        out["depth"].append(dict(c, sha256=hashlib.sha256(D.tobytes()).hexdigest(),
                                 first=[float(v) for v in D.reshape(-1)[:4]]))
    for c in UNIFORM:
        q = O.gen_queries_uniform(c["seed"], c["count"], c["stride"])
        out["uniform"].append(dict(c, sha256=hashlib.sha256(q.tobytes()).hexdigest()))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generator.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
