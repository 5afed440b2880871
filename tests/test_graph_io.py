"""Host graph I/O (graph.h / aesc.h loader semantics) and the seeded
generators.  CPU only."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from tgtp_amd import eigenprep as E
from tgtp_amd import graph as G
from tgtp_amd import workloads as W


def _both(path):
    g = G.load_txt(path)
    oi, oa = oracle.load_txt(path)
    assert np.array_equal(g.idx, oi) and np.array_equal(g.adj, oa)
    return g


def test_loader_comments_blanks_duplicates_selfloops(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# header\n% mm comment\n\n   \n0 1\n1 2 extra fields ok\n  2 0\n"
                 "0 1\n3 3\n\t# indented comment\n4\t0\r\n")
    g = _both(p)
    assert g.n == 5
    # duplicates kept, self-loop stored twice (graph.h:63-66)
    assert g.adj[g.idx[0]:g.idx[1]].tolist() == [1, 1, 2, 4]
    assert g.adj[g.idx[3]:g.idx[4]].tolist() == [3, 3]


def test_loader_max_id_defines_n_and_isolated_gap(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 1\n5 1\n")
    g = _both(p)
    assert g.n == 6 and g.degree(3) == 0
    with pytest.raises(ValueError, match="isolated"):
        G.check_aesc_graph(g, 2)
    with pytest.raises(ValueError, match="omega"):
        G.check_aesc_graph(g, 7)


@pytest.mark.parametrize("text", ["0 -1\n", "0\n", "a b\n", "1.5 2\n", "0 4294967295\n",
                                  "99999999999999999999 1\n"])
def test_loader_rejects_invalid_edges(tmp_path, text):
    p = tmp_path / "g.txt"
    p.write_text("0 1\n" + text)
    with pytest.raises(ValueError, match="invalid edge"):
        G.load_txt(p)
    with pytest.raises(ValueError, match="invalid edge"):
        oracle.load_txt(p)


def test_loader_empty_and_missing(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# nothing\n")
    with pytest.raises(ValueError, match="empty graph"):
        G.load_txt(p)
    with pytest.raises(OSError):
        G.load_txt(tmp_path / "missing.txt")


def test_eigen_loader_matches_oracle_and_rejects_nonfinite(tmp_path):
    g = G.chung_lu(200, 2.5, 8.0, seed=4)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 12))
    p = tmp_path / "e.txt"
    E.write_eigens_txt(p, lam, X)
    l1, X1 = G.load_eigens_txt(p, g.n, 12)
    l2, X2 = oracle.load_eigens(p, g.n, 12)
    assert np.array_equal(l1, lam) and np.array_equal(X1, X)
    assert np.array_equal(l1, l2) and np.array_equal(X1, X2)
    bad = tmp_path / "bad.txt"
    bad.write_text(p.read_text().replace(p.read_text().split()[5], "nan", 1))
    with pytest.raises(ValueError, match="invalid eigen"):
        G.load_eigens_txt(bad, g.n, 12)


@pytest.mark.parametrize("text, expect", [
    ("1.5 0x10", 0.0),         # `in >> x` gathers "0" and stops at 'x' (strtod: 16)
    ("1.5 0x1p3", 0.0),
    ("1.5 -.25", -0.25),
    ("1.5 2.5E+2", 250.0),
    ("1.5 1e-400", 0.0),       # an underflow keeps 0
    ("1.5 1e-40", float(np.float32(1e-40))),
    ("1.5 1e+", None),         # the gathered "1e+" is no number (strtod: 1)
    ("1.5 1e", None),
    ("1.5 3.4028236e38", None),  # overflows float
    ("1.5 inf", None),
    ("+.e1 2", None),
])
def test_eigen_loader_reads_fields_as_istream_does(tmp_path, text, expect):
    # aesc.h:51-64 reads every field with `in >> x`; the oracle's loader does
    # the same through libstdc++'s num_get
    p = tmp_path / "e.txt"
    p.write_text(text)
    if expect is None:
        with pytest.raises(ValueError):
            oracle.load_eigens(p, 1, 1)
        with pytest.raises(ValueError, match="invalid eigen"):
            G.load_eigens_txt(p, 1, 1)
        return
    l2, X2 = oracle.load_eigens(p, 1, 1)
    l1, X1 = G.load_eigens_txt(p, 1, 1)
    assert l1[0] == l2[0] == 1.5
    assert X1.tobytes() == X2.tobytes() == np.array([[expect]], np.float32).tobytes()


def test_loaders_under_sanitizers_match_oracle(tmp_path):
    """graph_io.cpp built with AddressSanitizer and UBSan, driven over
    seeded-random and malformed graph texts, eigenpair files and damaged
    binary caches (tests/cpp/graph_io_sanitize.cpp)."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    root = G.__file__.rsplit("/tgtp_amd/", 1)[0]
    exe = tmp_path / "graph_io_sanitize"
    build = subprocess.run(
        [cxx, "-std=c++20", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", f"-I{root}/include",
         f"{root}/tests/cpp/graph_io_sanitize.cpp", f"{root}/tgtp_amd/csrc/graph_io.cpp",
         f"{root}/oracle/orc_tgtp.cpp", f"{root}/oracle/orc_walk.cpp", "-o", str(exe)],
        capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("sanitizer runtime not available")
    assert build.returncode == 0, build.stderr[-2000:]
    out = subprocess.run([str(exe), str(tmp_path), "4000"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, (out.stdout + out.stderr)[-4000:]
    f = out.stdout.split()
    assert f[0] == "ok" and f[1] == "4000"
    stats = dict(zip(f[2::2], map(int, f[3::2])))
    # both outcomes were exercised
    assert 0 < stats["graphs_parsed"] < 4000
    assert 0 < stats["eigen_files_parsed"] < 4000
    assert stats["damaged_caches_refused"] == 4000
    assert stats["damaged_eigen_caches_refused"] == 4000


def _simple_connected(g):
    u = np.repeat(np.arange(g.n), g.degree())
    assert not np.any(u == g.adj), "self-loop"
    key = u.astype(np.int64) * g.n + g.adj
    assert np.unique(key).shape[0] == g.m, "duplicate"
    from oracle.exact import _components
    nc, _ = _components(g.idx, g.adj)
    assert nc == 1
    assert g.degree().min() >= 1
    for v in range(0, g.n, max(1, g.n // 50)):
        row = g.adj[g.idx[v]:g.idx[v + 1]]
        assert np.all(row[:-1] <= row[1:])


def test_chung_lu_deterministic_simple_connected():
    a = G.chung_lu(4096, 2.5, 22.0, seed=1)
    b = G.chung_lu(4096, 2.5, 22.0, seed=1)
    assert np.array_equal(a.idx, b.idx) and np.array_equal(a.adj, b.adj)
    _simple_connected(a)
    assert 85000 < a.m < 91000  # ~88k CSR entries
    c = G.chung_lu(4096, 2.5, 22.0, seed=2)
    assert not np.array_equal(a.adj[:1000], c.adj[:1000])


def test_workload_cl4096_shape():
    g = W.chung_lu_4096()
    assert g.n == 4096 and 85000 < g.m < 91000 and g.degree().max() > 1000


def test_grid_rewired_preserves_edge_count():
    g = G.grid_rewired(40, 30, 0.05, seed=3)
    E_ = 40 * 29 + 39 * 30
    _simple_connected(g)
    assert g.m >= 2 * E_ and g.m <= 2 * E_ + 2 * g.n  # repairs add at most one edge each
    plain = G.grid_rewired(40, 30, 0.0, seed=3)
    assert plain.m == 2 * E_ and plain.degree().max() == 4


def test_write_txt_roundtrip(tmp_path):
    g = G.chung_lu(500, 2.5, 9.0, seed=11)
    p = tmp_path / "g.txt"
    g.write_txt(p)
    h = G.load_txt(p)
    assert np.array_equal(g.idx, h.idx) and np.array_equal(g.adj, h.adj)


def test_reverse_index_is_first_occurrence(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 1\n0 1\n1 2\n2 2\n")
    g = G.load_txt(p)
    rev = g.reverse_index()
    u = np.repeat(np.arange(g.n), g.degree())
    for pos in range(g.m):
        v = g.adj[pos]
        row = g.adj[g.idx[v]:g.idx[v + 1]]
        assert rev[pos] == g.idx[v] + int(np.searchsorted(row, u[pos], side="left"))


def test_loader_parallel_chunks_match_oracle(tmp_path):
    """A text graph of several MiB is parsed by all threads in line-aligned
    chunks; the CSR equals the oracle's sequential loader, comments, blank
    lines, extra fields, duplicates and self-loops included."""
    rng = np.random.default_rng(7)
    k = 400_000
    a = rng.integers(0, 50_000, k)
    b = rng.integers(0, 50_000, k)
    lines = []
    for i in range(k):
        if i % 9973 == 0:
            lines.append("# comment line")
        if i % 7919 == 0:
            lines.append("   ")
        lines.append(f"{a[i]} {b[i]}" + (" 1.0 extra" if i % 5 == 0 else ""))
    lines.append("123 123")
    p = tmp_path / "big.txt"
    p.write_text("\n".join(lines) + "\n")
    assert p.stat().st_size > 4 << 20
    g = _both(p)
    assert g.m == 2 * (k + 1)


def test_loader_reports_first_invalid_line_across_chunks(tmp_path):
    """Invalid lines in several chunks: the message names the first one in
    file order, as a sequential reader would."""
    body = "".join(f"{i % 1000} {(i * 7) % 1000}\n" for i in range(600_000))
    p = tmp_path / "bad.txt"
    p.write_text(body + "7 x-first\n" + body + "9 y-second\n" + body)
    with pytest.raises(ValueError, match="x-first"):
        G.load_txt(p)


def test_csr_binary_cache_round_trip_and_damage(tmp_path):
    g = W.G.chung_lu(3000, 2.5, 12.0, seed=5)
    p = tmp_path / "g.csr"
    G.save_bin(g, p)
    h = G.load_bin(p)
    assert np.array_equal(g.idx, h.idx) and np.array_equal(g.adj, h.adj)
    raw = bytearray(p.read_bytes())
    for pos in (30, len(raw) // 2, len(raw) - 3):  # idx, adj, digest
        bad = bytearray(raw)
        bad[pos] ^= 0x10
        q = tmp_path / f"bad{pos}.csr"
        q.write_bytes(bytes(bad))
        with pytest.raises(ValueError, match="not a valid CSR cache"):
            G.load_bin(q)
    q = tmp_path / "short.csr"
    q.write_bytes(bytes(raw[:-1]))
    with pytest.raises(ValueError):
        G.load_bin(q)
    q = tmp_path / "magic.csr"
    q.write_bytes(b"XGTPCSR1" + bytes(raw[8:]))
    with pytest.raises(ValueError):
        G.load_bin(q)
    with pytest.raises(OSError):
        G.load_bin(tmp_path / "missing.csr")


def test_csr_binary_cache_checks_invariants(tmp_path):
    """A file whose digest matches but whose rows are unsorted (written
    through the same writer) is refused."""
    g = G.Graph.from_edges(4, [(0, 1), (0, 2), (2, 3)])
    adj = g.adj.copy()
    adj[0], adj[1] = adj[1], adj[0]  # row 0 = [2, 1]
    G.save_bin(G.Graph(g.idx, adj), tmp_path / "u.csr")
    with pytest.raises(ValueError, match="invariants"):
        G.load_bin(tmp_path / "u.csr")


def test_eigen_binary_cache_round_trip_and_damage(tmp_path):
    g = G.chung_lu(300, 2.5, 8.0, seed=4)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 16))
    p = tmp_path / "e.bin"
    G.save_eigens_bin(p, lam, X)
    l2, X2 = G.load_eigens_bin(p, g.n, 16)
    assert l2.tobytes() == lam.tobytes() and X2.tobytes() == X.tobytes()
    # the caller's sizes must match the file's
    for n, om in ((g.n - 1, 16), (g.n, 15), (g.n + 1, 16)):
        with pytest.raises(ValueError, match="expected n"):
            G.load_eigens_bin(p, n, om)
    raw = bytearray(p.read_bytes())
    for pos in (3, 12, 30, len(raw) // 2, len(raw) - 3):  # magic, header, values, vectors, digest
        bad = bytearray(raw)
        bad[pos] ^= 0x10
        q = tmp_path / f"bad{pos}.bin"
        q.write_bytes(bytes(bad))
        with pytest.raises(ValueError, match="not a valid eigenpair cache"):
            G.load_eigens_bin(q, g.n, 16)
    q = tmp_path / "short.bin"
    q.write_bytes(bytes(raw[:-1]))
    with pytest.raises(ValueError):
        G.load_eigens_bin(q, g.n, 16)
    with pytest.raises(OSError):
        G.load_eigens_bin(tmp_path / "missing.bin", g.n, 16)
    Xn = X.copy()
    Xn[5, 3] = np.nan
    with pytest.raises(ValueError, match="finite"):
        G.save_eigens_bin(tmp_path / "nan.bin", lam, Xn)
    assert not (tmp_path / "nan.bin").exists()


def test_load_eigens_txt_cached(tmp_path):
    import os
    g = G.chung_lu(200, 2.5, 8.0, seed=4)
    lam, X = E.eigens_f32(*E.compute_eigens(g, 12))
    p = tmp_path / "sorted_eigens_12.txt"
    E.write_eigens_txt(p, lam, X)
    l1, X1 = G.load_eigens_txt_cached(p, g.n, 12)
    cache = tmp_path / "sorted_eigens_12.txt.bin"
    assert cache.exists()
    assert l1.tobytes() == lam.tobytes() and X1.tobytes() == X.tobytes()
    # a cache holding other values is used while it is newer than the text
    G.save_eigens_bin(cache, lam * 2, X)
    l2, _ = G.load_eigens_txt_cached(p, g.n, 12)
    assert np.array_equal(l2, lam * 2)
    # ... and rebuilt once the text is newer
    st = cache.stat()
    os.utime(p, (st.st_atime + 10, st.st_mtime + 10))
    l3, X3 = G.load_eigens_txt_cached(p, g.n, 12)
    assert l3.tobytes() == lam.tobytes() and X3.tobytes() == X.tobytes()
    # a cache of other sizes is rebuilt, not trusted
    os.utime(cache, (st.st_atime + 20, st.st_mtime + 20))
    G.save_eigens_bin(cache, lam[:11], X[:, :11])
    l4, _ = G.load_eigens_txt_cached(p, g.n, 12)
    assert l4.tobytes() == lam.tobytes()


def test_load_txt_cached(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 1\n1 2\n2 0\n")
    g1 = G.load_txt_cached(p)
    cache = tmp_path / "g.txt.csr"
    assert cache.exists()
    g2 = G.load_txt_cached(p)
    assert np.array_equal(g1.adj, g2.adj) and np.array_equal(g1.idx, g2.idx)
    cache.write_bytes(b"junk")
    import os
    os.utime(cache, (p.stat().st_mtime + 10, p.stat().st_mtime + 10))
    g3 = G.load_txt_cached(p)  # damaged cache: parsed again and rewritten
    assert np.array_equal(g1.adj, g3.adj)
    assert G.load_bin(cache).m == g1.m


_THREADS_PROBE = r"""
import hashlib, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from tgtp_amd import graph as G
h = hashlib.sha256()
g = G.chung_lu(200000, 2.3, 12.0, seed=9)   # above the generator's parallel thresholds
h.update(g.idx.tobytes()); h.update(g.adj.tobytes())
g2 = G.grid_rewired(300, 300, 0.05, seed=4)
h.update(g2.idx.tobytes()); h.update(g2.adj.tobytes())
t = G.load_txt(sys.argv[2])
h.update(t.idx.tobytes()); h.update(t.adj.tobytes())
lam, X = G.load_eigens_txt(sys.argv[3], int(sys.argv[4]), int(sys.argv[5]))
h.update(lam.tobytes()); h.update(X.tobytes())
print(h.hexdigest())
"""


def test_host_results_independent_of_thread_count(tmp_path):
    """The generators, the graph loader and the eigenpair loader give the
    same bits on 1 and on 8 host threads (inputs large enough for every
    parallel path)."""
    import os
    import subprocess
    import sys
    root = str(Path(__file__).resolve().parents[1])
    g = G.chung_lu(150000, 2.5, 10.0, seed=3)
    gp = tmp_path / "g.txt"
    g.write_txt(gp)  # several MiB: parsed in chunks
    n, om = 40000, 8
    rng = np.random.default_rng(2)
    ep = tmp_path / "e.txt"
    E.write_eigens_txt(ep, rng.standard_normal(om), rng.standard_normal((n, om)).astype(np.float32))
    digests = []
    for threads in ("1", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=threads)
        r = subprocess.run([sys.executable, "-c", _THREADS_PROBE, root, str(gp), str(ep), str(n),
                            str(om)], capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        digests.append(r.stdout.strip())
    assert digests[0] == digests[1]
