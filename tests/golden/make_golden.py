"""Writes the committed golden fixtures for the synthetic workloads.

For the 50K-vertex / 120-level DAG (K=10,000) it stores the oracle's
top-K costs; for the 1.5M-vertex / 300-level DAG (K=1,000,000) a SHA-256 of
the top-K cost bytes and a few ranks.  Both record a SHA-256 of the
generated fanout CSR so that a generator change is caught.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-cpg_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cpgamd  # noqa: E402
import pyoracle  # noqa: E402


def csr_sha256(og):
    a = og.arrays()
    h = hashlib.sha256()
    for key in ("oadjp", "oadj", "owgt"):
        h.update(np.ascontiguousarray(a[key]).tobytes())
    return h.hexdigest()


def main():
    meta = {}
    for which, k, name in ((1, 10_000, "c1"), (2, 1_000_000, "c2")):
        n, u, v, w = cpgamd.gen_preset(which)
        og = pyoracle.OracleGraph.from_edges(n, u, v, w)
        dist, succ, level, L = og.sfxt()
        costs, st = og.topk(dist, succ, k)
        entry = dict(n=int(n), m=int(og.m), levels=int(L), k=int(k), csr_sha256=csr_sha256(og),
                     costs_sha256=hashlib.sha256(costs.astype(np.float32).tobytes()).hexdigest(),
                     count=int(len(costs)),
                     ranks={str(r): float(costs[r]) for r in (0, 999, 9999, len(costs) - 1)},
                     nonmonotone=int(st["nonmonotone"]), roots=int(st["roots"]),
                     roots_dropped=int(st["roots_dropped"]), H_K=int(st["hops"]),
                     E_K=int(st["edges"]), C_K=int(st["children"]))
        meta[name] = entry
        if which == 1:
            np.save(os.path.join(HERE, "c1_k10000_costs.npy"), costs.astype(np.float32))
        print(name, entry)
    with open(os.path.join(HERE, "synthetic_golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
