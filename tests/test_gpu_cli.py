"""GPU tests of the C++ tools that mirror the reference executables and file formats
(lstc_preprocess ~ without_part_preprocess.py / Halo_creation_preprocess.py,
lstc_tc ~ TC_128, lstc_halo ~ round_clean_sg, lstc_part_tc ~ wo_sort_tc_256)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "trianglecount_amd", "_bin")


def _run(args, cwd):
    print("run:", " ".join(args), flush=True)
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def _ints(path):
    return [int(x) for x in open(path).read().split()]


def _write_tsv(path, s, d):
    with open(path, "w") as f:
        for a, b in zip(s.tolist(), d.tolist()):
            f.write(f"{a}\t{b}\t1\n")


@pytest.fixture(scope="module")
def rmat_files(tmp_path_factory, oracle):
    wd = tmp_path_factory.mktemp("cli")
    n, s, d = oracle.gen_rmat(11, 8, seed=3)
    _write_tsv(str(wd / "g.tsv"), s, d)
    out = _run([os.path.join(BIN, "lstc_preprocess"), "g.tsv", "--parts", "4", "--seed", "9"], str(wd))
    assert "wrote g_output.csr" in out
    sym, g = oracle.clean_orient(n, s, d)
    return wd, sym, g


def test_preprocess_writes_reference_formats(lstc, oracle, rmat_files):
    wd, sym, g = rmat_files
    v = _ints(str(wd / "g_output.csr"))
    # V E |row_ptr| |col| / row_ptr / col (Med_dataset/without_part_preprocess.py:298-311)
    assert v[:4] == [g.n, 2 * g.m, g.n + 1, g.m]
    assert v[4:4 + g.n + 1] == g.off.tolist()
    assert v[4 + g.n + 1:] == g.col.tolist()
    p = _ints(str(wd / "g_part.csr"))
    # V E / V V+1 |col| 0 / row_ptr / col / in_deg (Halo_creation_preprocess.py:259-278)
    assert p[:6] == [sym.n, sym.m, sym.n, sym.n + 1, sym.m, 0]
    body = p[6:]
    assert body[: sym.n + 1] == sym.off.tolist()
    assert body[sym.n + 1: sym.n + 1 + sym.m] == sym.col.tolist()
    assert body[sym.n + 1 + sym.m:] == np.diff(sym.off).tolist()
    parts = _ints(str(wd / "g_part.csr_node_parts_random_4"))
    assert parts == oracle.partition_random(sym.n, 4, seed=9).tolist()
    contig = _ints(str(wd / "g_part.csr_node_parts_contig_4"))
    l = sym.n // 4
    assert contig == [(i // l if i < 4 * l else 0) for i in range(sym.n)]


def test_tc_matches_oracle(lstc, oracle, rmat_files):
    wd, sym, g = rmat_files
    for extra in ([], ["--plain"]):
        out = _run([os.path.join(BIN, "lstc_tc"), "g_output.csr"] + extra, str(wd))
        tri = [int(x) for x in re.findall(r"Triangle : (\d+)", out)]
        assert tri == [oracle.count(g)] * 3
        assert "Total avg of 3 RUNS" in out
        # the split count's one-time input is reported apart from the runs
        assert ("In-lists build :" in out) == (extra == [])


@pytest.mark.parametrize("prune", [False, True])
def test_halo_and_partitioned_count(lstc, oracle, rmat_files, prune):
    wd, sym, g = rmat_files
    args = [os.path.join(BIN, "lstc_halo"), "4", "g_part.csr", "g_part.csr_node_parts_random_4"]
    _run(args + (["--prune"] if prune else []), str(wd))
    vals = _ints(str(wd / "g_part.csr_node_parts_random_4_output.csr"))
    assert vals[:2] == [sym.n, sym.m]
    part = oracle.partition_random(sym.n, 4, seed=9)
    pos = 2
    for p in range(4):
        h = oracle.halo_extract(sym, part, p, 1 if prune else 0)
        nl, rps, ml, no = vals[pos: pos + 4]
        assert (nl, rps, ml, no) == (h.csr.n, h.csr.n + 1, h.csr.m, h.n_owned)
        pos += 4
        assert vals[pos: pos + rps] == h.csr.off.tolist()
        pos += rps
        assert vals[pos: pos + ml] == h.csr.col.tolist()
        pos += ml
    want = oracle.count(g)
    for extra in ([], ["--per-part"], ["--plain"], ["--per-part", "--plain"]):
        out = _run([os.path.join(BIN, "lstc_part_tc"), "g_part.csr_node_parts_random_4_output.csr", "4"] + extra,
                   str(wd))
        assert int(re.search(r"Total Triangle : (\d+)", out).group(1)) == want


@pytest.mark.parametrize("ext", [".txt", ".mtx"])
def test_preprocess_snap_and_matrix_market_inputs(lstc, oracle, tmp_path, ext):
    """lstc_preprocess reads SNAP .txt (4 header lines) and 1-based .mtx inputs to the same
    <name>_output.csr as the oracle's cleaning of the edges they hold."""
    n, s, d = oracle.gen_rmat(10, 8, seed=11)
    e = np.stack([s, d], 1).astype(np.int64)
    if ext == ".txt":
        body = "# Undirected graph: g.txt\n# synthetic\n# Nodes: x Edges: y\n# FromNodeId\tToNodeId\n"
        body += "".join(f"{a}\t{b}\n" for a, b in e)
    else:
        body = f"%%MatrixMarket matrix coordinate pattern general\n% synthetic\n{n} {n} {len(e)}\n"
        body += "".join(f"{a + 1} {b + 1}\n" for a, b in e)
    (tmp_path / f"g{ext}").write_text(body)
    _run([os.path.join(BIN, "lstc_preprocess"), f"g{ext}"], str(tmp_path))
    sym, g = oracle.clean_orient(int(e.max()) + 1, s, d)
    v = _ints(str(tmp_path / "g_output.csr"))
    assert v[:4] == [g.n, 2 * g.m, g.n + 1, g.m]
    assert v[4:4 + g.n + 1] == g.off.tolist() and v[4 + g.n + 1:] == g.col.tolist()
