"""Write the oracle's results for a few fixed cases to tests/golden/*.json.

Each fixture holds the case's parameters, the SHA-256 of the moments
p, u, v, w (float64, [z][y][x], with -0.0 folded to 0.0) after the given
number of steps, their L2 norms, and a fixed sample of values, so that the GPU
tests can check bit-exactness at sizes the oracle takes tens of seconds on.

    python tests/golden/make_golden.py           # all fixtures
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import CONFIG, FAITHFUL, LDC, SPHERE, TGV, Oracle  # noqa: E402

PI_R = 3.1415926


def tgv_tau(n, re, u0):
    return u0 * n / re / (1.0 / 3.0) / 2 / PI_R  # tgv/src/System.cpp:21


CASES = {
    # the TGV parity configuration: 64^3, U0 0.05, Re 100, 1000 steps
    "tgv64_re100_1000": dict(case=TGV, mode=FAITHFUL, n=(64, 64, 64), tau0=tgv_tau(64, 100.0, 0.05), u0=0.05,
                             steps=1000),
    "cavity_faithful_24_300": dict(case=LDC, mode=FAITHFUL, n=(24, 24, 24), tau0=0.1, u0=0.1, steps=300),
    "cavity_config_32_500": dict(case=LDC, mode=CONFIG, n=(32, 32, 32), tau0=3 * 0.1 * 32 / 400.0, u0=0.0,
                                 lid_u=0.1, steps=500),
    "sphere_faithful_72x20x20_200": dict(case=SPHERE, mode=FAITHFUL, n=(72, 20, 20), tau0=0.08, u0=0.1,
                                         steps=200),
    "sphere_config_48x24x24_300": dict(case=SPHERE, mode=CONFIG, n=(48, 24, 24), tau0=0.03, u0=0.05,
                                       sphere=(12.0, 12.0, 12.0, 4.0), steps=300),
}


def digest(fields):
    h = hashlib.sha256()
    for a in fields:
        h.update((np.ascontiguousarray(a, dtype=np.float64) + 0.0).tobytes())
    return h.hexdigest()


def sample_index(shape, k=64, seed=7):
    rng = np.random.default_rng(seed)
    return [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(k)]


def make(name, c):
    o = Oracle(c["case"], *c["n"], tau0=c["tau0"], u0=c["u0"], mode=c["mode"], lid_u=c.get("lid_u", 0.1),
               sphere=c.get("sphere", (0.0, 0.0, 0.0, 0.0)))
    o.Initialize()
    o.run(c["steps"])
    m = o.moments()
    idx = sample_index(m[0].shape)
    out = {k: (list(v) if isinstance(v, tuple) else v) for k, v in c.items()}
    out.update({
        "sha256_moments": digest(m),
        "l2": [float(np.linalg.norm(a)) for a in m],
        "sample_index": [list(i) for i in idx],
        "sample": [[float(a[i]) for i in idx] for a in m],
    })
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(out, f, indent=1)
    print(name, out["sha256_moments"][:16])


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for nm in names:
        make(nm, CASES[nm])
