"""The seeded synthetic workloads the benchmarks and parity tests run on.

Each is regenerated bit-identically from its parameters on any machine with
this image (integer RNG streams; glibc log/pow in the Chung-Lu skip sampler).
"""
from __future__ import annotations

import math

from . import graph as G

# Chung-Lu n=4096, exponent 2.5, ~88k CSR entries (walk config 1, tgtp config 2)
CL4096 = dict(n=4096, beta=2.5, mean_deg=22.0, w_cap=0.0, seed=1)
# Chung-Lu n=2^20, exponent 2.2, ~10M undirected edges, hub weight capped at
# sqrt(sum w) so per-vertex counters stay far below 2^32 (walk config 3)
CL1M = dict(n=1 << 20, beta=2.2, mean_deg=21.8, w_cap=math.sqrt(21.8 * (1 << 20)), seed=1)
# 512x512 grid with 5% of edges rewired (tgtp precision config 4)
GRID512 = dict(rows=512, cols=512, p_rewire=0.05, seed=1)


def chung_lu_4096():
    return G.chung_lu(**CL4096)


def chung_lu_1m():
    return G.chung_lu(**CL1M)


def grid_512():
    return G.grid_rewired(**GRID512)


def describe(params: dict) -> str:
    return ",".join(f"{k}={v:.6g}" if isinstance(v, float) else f"{k}={v}"
                    for k, v in params.items())


# Batched tgtp workload: 256 Chung-Lu graphs, n = floor(2^U(10, 16))
# (log-uniform in [1024, 65536]), exponent U(2.1, 3.0), mean degree U(8, 40),
# graph k seeded 1000 + k; the parameters come from one seeded PCG64 stream.
BATCH256 = dict(count=256, log2n=(10.0, 16.0), beta=(2.1, 3.0), mean_deg=(8.0, 40.0),
                seed=5, graph_seed0=1000)


def batch_params(count: int | None = None, spec: dict = BATCH256):
    """[(n, beta, mean_deg, seed)] of the batched workload (first `count`)."""
    import numpy as np

    rng = np.random.default_rng(spec["seed"])
    k = spec["count"]
    e = rng.uniform(*spec["log2n"], size=k)
    b = rng.uniform(*spec["beta"], size=k)
    d = rng.uniform(*spec["mean_deg"], size=k)
    out = [(int(np.floor(2.0 ** e[i])), float(b[i]), float(d[i]), spec["graph_seed0"] + i)
           for i in range(k)]
    return out if count is None else out[:count]


def batch_graphs(count: int | None = None, spec: dict = BATCH256):
    return [G.chung_lu(n, beta, md, seed=s) for n, beta, md, s in batch_params(count, spec)]
