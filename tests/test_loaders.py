"""CPU tests of the edge-list loaders the preprocessing tool uses (trianglecount_amd/cli/lstc_text.h),
against the reference's loading rules (Med_dataset/without_part_preprocess.py:126-177):
.tsv ids as written, .txt with its first 4 lines skipped, .mtx/.mmio 1-based with a size line,
n = largest id + 1, other extensions refused."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def reader(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("rd") / "read_edges")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "trianglecount_amd", "cli"),
                    os.path.join(ROOT, "tests", "native", "read_edges.cpp"), "-o", exe], check=True)

    def run(path):
        r = subprocess.run([exe, path], capture_output=True, text=True)
        lines = r.stdout.split("\n")
        if r.returncode != 0:
            return lines[0]
        n, m = (int(x) for x in lines[0].split())
        e = np.array([[int(x) for x in ln.split()] for ln in lines[1:1 + m]], dtype=np.int64).reshape(m, 2)
        return n, e
    return run


def _edges(seed=1, n=50, m=300):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, size=(m, 2))


def test_tsv_as_written(reader, tmp_path):
    e = _edges()
    p = tmp_path / "g.tsv"
    p.write_text("".join(f"{a}\t{b}\t1\n" for a, b in e))
    n, got = reader(str(p))
    assert n == e.max() + 1 and np.array_equal(got, e)


def test_txt_skips_exactly_four_header_lines(reader, tmp_path):
    e = _edges(2)
    hdr = "# Directed graph (each unordered pair of nodes is saved once): x.txt\n# comment\n" \
          "# Nodes: 50 Edges: 300\n# FromNodeId\tToNodeId\n"
    p = tmp_path / "g.txt"
    p.write_text(hdr + "".join(f"{a}\t{b}\n" for a, b in e))
    n, got = reader(str(p))
    assert n == e.max() + 1 and np.array_equal(got, e)
    # the first four lines are skipped whatever they hold (pandas skiprows=4)
    q = tmp_path / "h.txt"
    q.write_text("".join(f"{a}\t{b}\n" for a, b in e))
    n2, got2 = reader(str(q))
    assert np.array_equal(got2, e[4:])


@pytest.mark.parametrize("ext", [".mtx", ".mmio"])
def test_matrix_market_one_based_with_size_line(reader, tmp_path, ext):
    e = _edges(3)
    body = "%%MatrixMarket matrix coordinate pattern general\n% a comment\n"
    body += f"{e.max() + 1} {e.max() + 1} {len(e)}\n" + "".join(f"{a + 1} {b + 1}\n" for a, b in e)
    p = tmp_path / f"g{ext}"
    p.write_text(body)
    n, got = reader(str(p))
    assert n == e.max() + 1 and np.array_equal(got, e)
    # value columns are ignored
    q = tmp_path / f"v{ext}"
    q.write_text(body.replace(" pattern ", " real ").replace("\n", " 0.5\n").replace("% a comment 0.5", "% c")
                 .replace("general 0.5", "general"))
    n2, got2 = reader(str(q))
    assert n2 == n and np.array_equal(got2, got)
    z = tmp_path / f"z{ext}"
    z.write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n0 2\n")
    assert reader(str(z)).startswith("error: 0 index")


def test_formats_clean_to_the_same_graph(reader, tmp_path, oracle):
    """The same edges written as .tsv, .txt and .mtx load and clean to one oriented CSR."""
    n0, s, d = oracle.gen_rmat(9, 8, seed=5)
    e = np.stack([s, d], 1).astype(np.int64)
    (tmp_path / "g.tsv").write_text("".join(f"{a}\t{b}\t1\n" for a, b in e))
    (tmp_path / "g.txt").write_text("#\n#\n#\n#\n" + "".join(f"{a}\t{b}\n" for a, b in e))
    (tmp_path / "g.mtx").write_text(f"%%MatrixMarket matrix coordinate pattern general\n{n0} {n0} {len(e)}\n" +
                                    "".join(f"{a + 1} {b + 1}\n" for a, b in e))
    outs = []
    for ext in (".tsv", ".txt", ".mtx"):
        n, got = reader(str(tmp_path / f"g{ext}"))
        _, g = oracle.clean_orient(n, got[:, 0].astype(np.uint32), got[:, 1].astype(np.uint32))
        outs.append((g.off.tolist(), g.col.tolist()))
    assert outs[0] == outs[1] == outs[2]


def test_unsupported_extension_refused(reader, tmp_path):
    p = tmp_path / "g.csv"
    p.write_text("0,1\n")
    assert reader(str(p)).startswith("error: unsupported file type .csv")
