"""Writes the golden fixtures under tests/golden/ (run where the reference
tree is present; the tests themselves only read the .npz files).

Sources of truth:
  * saba_kats.json (not written here): known answers of the reference's own
    saba_randomwalk (src/rw_experiment_saba.cpp:65-145), recorded from a
    build of its own source with a lane-wise % in place of Intel SVML's
    _mm256_rem_epu32 (which only icpx provides): C9 (40 hops, 64 walks) and
    K5 (20 hops, 64 walks), both forcing the whole-bouquet SplitMix32
    fallback.
  * ref_rwmt_<case>.npz: per-vertex visit counts of the reference's own
    scalar_randomwalk (src/rw_experiment.cpp:34-47, per-source std::mt19937
    reseeded 42 + source) through oracle/_ref/ref_mt_walk, for several
    (depth, n_walks) per graph.
  * ref_mt_<case>.npz: outputs of oracle/_ref/ref_tgtp_mt, i.e. the
    reference's own graph.h loader, EigenVecs::load_txt, calculate_taus and
    AESC::tgtp() of src/aesc_baseline.h (scalar MT19937 walks), compiled from
    /root/reference by `make -C oracle ref`, run on graphs from this
    package's seeded generators with eigenpairs from tgtp_amd.eigenprep.

usage: python tests/golden/make_golden.py [--only-mt-walks]
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from tgtp_amd import eigenprep as E  # noqa: E402
from tgtp_amd import graph as G  # noqa: E402

HERE = Path(__file__).resolve().parent
REF = ROOT / "oracle" / "_ref" / "ref_tgtp_mt"
REF_WALK = ROOT / "oracle" / "_ref" / "ref_mt_walk"

# text of a graph with duplicates, a self-loop and an isolated vertex (4)
ODD_GRAPH = "0 1\n0 1\n1 2\n2 2\n2 3\n5 3\n5 0\n6 5\n"

# (name, graph factory, [(depth, n_walks), ...])
MT_WALK_CASES = [
    ("cycle15", lambda: G.cycle(15), [(5, 3), (5, 64)]),
    ("petersen", lambda: G.petersen(), [(15, 8), (1, 40)]),
    ("odd7", None, [(6, 16), (3, 5)]),
    ("grid24", lambda: G.grid_rewired(24, 24, 0.05, seed=2), [(100, 20), (2, 200)]),
    ("chunglu2000", lambda: G.chung_lu(2000, 2.5, 21.3, seed=5), [(5, 64), (10, 24)]),
    ("chunglu4096", lambda: G.chung_lu(4096, 2.5, 22.0, seed=1), [(5, 2048)]),
]


def mt_walk_fixtures(tmp: str) -> None:
    for name, make, runs in MT_WALK_CASES:
        gtxt = f"{tmp}/{name}.graph.txt"
        if make is None:
            Path(gtxt).write_text(ODD_GRAPH)
        else:
            make().write_txt(gtxt)
        g = G.load_txt(gtxt)
        out = {"idx": g.idx, "adj": g.adj, "runs": np.array(runs, dtype=np.int64)}
        for depth, n_walks in runs:
            f = f"{tmp}/{name}.d{depth}.w{n_walks}.f32"
            subprocess.run([str(REF_WALK), gtxt, str(depth), str(n_walks), f], check=True)
            c = np.fromfile(f, np.float32)
            assert c.max() < 2 ** 24 and np.array_equal(c, np.floor(c))
            out[f"counts_d{depth}_w{n_walks}"] = c.astype(np.uint32)
        np.savez_compressed(HERE / f"ref_rwmt_{name}.npz", **out)
        print("mt walks", name, "n", g.n, "m", g.m, runs)

# (name, graph factory, omega, epsilons, gamma)
CASES = [
    ("petersen_w4", lambda: G.petersen(), 4, (0.05, 0.01), 10),
    ("cycle15_w5", lambda: G.cycle(15), 5, (0.05,), 10),
    ("grid24_w32", lambda: G.grid_rewired(24, 24, 0.05, seed=2), 32, (0.05, 0.01), 10),
    ("grid32_w128", lambda: G.grid_rewired(32, 32, 0.05, seed=1), 128, (0.05,), 10),
    ("chunglu2000_w16", lambda: G.chung_lu(2000, 2.5, 21.3, seed=5), 16, (0.05, 0.01), 10),
    ("chunglu300_w64_g1", lambda: G.chung_lu(300, 2.5, 12.0, seed=3), 64, (0.2,), 1),
]

SABA_KATS = {
    "source": "reference saba_randomwalk (src/rw_experiment_saba.cpp:65-145), recorded run",
    "cases": [
        {"graph": "cycle", "n": 9, "depth": 40, "n_walks": 64,
         "visit_count": 23040, "checksum": 130632},
        {"graph": "complete", "n": 5, "depth": 20, "n_walks": 64,
         "visit_count": 6400, "checksum": 19771},
    ],
}


def main() -> None:
    if not REF.exists() or not REF_WALK.exists():
        raise SystemExit(f"{REF} / {REF_WALK} missing: run `make -C oracle ref` where "
                         "/root/reference exists")
    with tempfile.TemporaryDirectory() as tmp:
        mt_walk_fixtures(tmp)
    if "--only-mt-walks" in sys.argv:
        return
    (HERE / "saba_kats.json").write_text(json.dumps(SABA_KATS, indent=2) + "\n")
    with tempfile.TemporaryDirectory() as tmp:
        for name, make, omega, epsilons, gamma in CASES:
            g = make()
            w, X = E.compute_eigens(g, omega)
            lam, X32 = E.eigens_f32(w, X)
            gtxt, etxt = f"{tmp}/{name}.graph.txt", f"{tmp}/{name}.eig.txt"
            g.write_txt(gtxt)
            E.write_eigens_txt(etxt, lam, X32)
            out = {"omega": np.int32(omega), "gamma": np.int32(gamma),
                   "epsilons": np.array(epsilons, dtype=np.float64)}
            for k, eps in enumerate(epsilons):
                pre = f"{tmp}/{name}.{k}"
                subprocess.run([str(REF), gtxt, etxt, str(omega), repr(eps), str(gamma), pre],
                               check=True)
                if k == 0:
                    out["idx"] = np.fromfile(pre + ".idx.u32", np.uint32)
                    out["adj"] = np.fromfile(pre + ".adj.u32", np.uint32)
                    out["lam"] = np.fromfile(pre + ".lam.f64", np.float64)
                    out["X"] = np.fromfile(pre + ".vec.f32", np.float32).reshape(-1, omega)
                out[f"taus_{k}"] = np.fromfile(pre + ".taus.i32", np.int32)
                out[f"scores_mt_{k}"] = np.fromfile(pre + ".scores.f32", np.float32)
            # the text inputs exactly as the reference read them, for loader tests
            out["graph_txt"] = np.frombuffer(Path(gtxt).read_bytes(), dtype=np.uint8)
            np.savez_compressed(HERE / f"ref_mt_{name}.npz", **out)
            print(name, "n", g.n, "m", g.m, "omega", omega)


if __name__ == "__main__":
    main()
