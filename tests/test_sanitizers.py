"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): the oracle's
whole path on small seeded graphs (its independent pieces must agree), and the text-format
readers and writers of the command-line tools on valid, truncated and corrupted files (each
malformed file must be refused with a message, never read out of bounds)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc and g++")


def _run(args, **kw):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=300, **kw)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    return r


def test_oracle_path_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_check")
    subprocess.run(["gcc", "-std=c11", "-fopenmp", *SAN, "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "native", "oracle_check.c"),
                    os.path.join(ROOT, "oracle", "lstc_oracle.c"), "-lm", "-o", exe], check=True)
    r = _run([exe])
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert r.stdout.count(" ok") == 5, r.stdout


@pytest.fixture(scope="module")
def text_tool(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tf") / "text_formats")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", *SAN, "-I", os.path.join(ROOT, "trianglecount_amd", "cli"),
                    os.path.join(ROOT, "tests", "native", "text_formats.cpp"), "-o", exe], check=True)
    return exe


def test_text_formats_round_trip(text_tool, tmp_path):
    r = _run([text_tool, "roundtrip", str(tmp_path)])
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])


GOOD_ORIENTED = "5 12 6 6\n0 2 4 5 6 6\n1 2 2 3 4 4\n"
BAD_ORIENTED = {
    "truncated_col": "5 12 6 6\n0 2 4 5 6 6\n1 2 2\n",
    "truncated_rowptr": "5 12 6 6\n0 2 4\n",
    "huge_sizes": "5 12 99999999999999 99999999999999\n0 1\n",
    "overflowing_number": "5 12 6 6\n0 2 4 5 6 6\n1 2 2 3 4 " + "9" * 40 + "\n",
    "decreasing_rowptr": "5 12 6 6\n0 4 2 5 6 6\n1 2 2 3 4 4\n",
    "rowptr_end_mismatch": "5 12 6 6\n0 2 4 5 6 7\n1 2 2 3 4 4\n",
    "column_beyond_rows": "5 12 6 6\n0 2 4 5 6 6\n1 2 2 3 4 9\n",
    "empty": "",
}
BAD_PARTITIONED = {
    "missing_part": "5 6\n5 6 6 3\n0 2 4 5 6 6\n1 2 2 3 4 4\n",
    "owned_beyond_rows": "5 6\n5 6 6 9\n0 2 4 5 6 6\n1 2 2 3 4 4\n5 6 6 3\n0 2 4 5 6 6\n1 2 2 3 4 4\n",
    "huge_part": "5 6\n5 99999999999 6 3\n0 2\n",
}


def test_oriented_reader_accepts_good_and_refuses_bad(text_tool, tmp_path):
    p = tmp_path / "good.csr"
    p.write_text(GOOD_ORIENTED)
    assert _run([text_tool, "read", "oriented", str(p)]).returncode == 0
    for name, body in BAD_ORIENTED.items():
        q = tmp_path / f"{name}.csr"
        q.write_text(body)
        r = _run([text_tool, "read", "oriented", str(q)])
        assert r.returncode == 1 and r.stdout.startswith("error:"), (name, r.returncode, r.stdout, r.stderr[-500:])


def test_part_and_partitioned_readers_refuse_bad_files(text_tool, tmp_path):
    for name, body in BAD_PARTITIONED.items():
        q = tmp_path / f"{name}.csr"
        q.write_text(body)
        r = _run([text_tool, "read", "partitioned", str(q)])
        assert r.returncode == 1 and r.stdout.startswith("error:"), (name, r.returncode, r.stdout)
    bad_part = {"huge": "5 6\n5 99999999999 6 0\n0 1\n", "short_deg": "5 6\n5 6 6 0\n0 2 4 5 6 6\n1 2 2 3 4 4\n1 1\n"}
    for name, body in bad_part.items():
        q = tmp_path / f"{name}_part.csr"
        q.write_text(body)
        r = _run([text_tool, "read", "part", str(q)])
        assert r.returncode == 1 and r.stdout.startswith("error:"), (name, r.returncode, r.stdout)
    q = tmp_path / "np"
    q.write_text("0 1 0\n")
    assert _run([text_tool, "read", "parts", str(q)]).returncode == 1  # 3 of the 5 vertices


@pytest.mark.parametrize("ext,body,ok", [
    (".tsv", "0\t1\t1\n1\t2\t1\n", True),
    (".tsv", "", True),                                       # no edges: n = 0
    (".tsv", "0\t" + "9" * 30 + "\t1\n", False),             # id beyond 32 bits
    (".tsv", "-1\t2\t1\n", False),                           # negative id
    (".mtx", "%%MatrixMarket\n3 3 2\n0 1\n", False),          # 0 in a 1-based file
    (".mtx", "%%MatrixMarket\n3 3 2\n1 2\n3 1\n", True),
    (".txt", "a\nb\nc\nd\n" + "x" * 10000 + "\n1 2\n", True),  # a line longer than the read buffer
    (".csv", "0,1\n", False),                                 # refused extension
])
def test_edge_list_reader_under_sanitizers(text_tool, tmp_path, ext, body, ok):
    p = tmp_path / f"g{ext}"
    p.write_text(body)
    r = _run([text_tool, "read", "edges", str(p)])
    assert (r.returncode == 0) == ok, (r.returncode, r.stdout, r.stderr[-500:])
