"""The C++ drop-in classes (GraphMatchingGPU*, Graph readers) driven the way
the reference's bin/match.cpp drives them, through tests/cpp/dropin_check."""
import os
import subprocess

import numpy as np
import pytest

import mmamd
import oracle
from mmamd import CSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_check")


def run_driver(*args):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/dropin_check is not built (run __graft_entry__.build())")
    return subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=300)


def parse_read(out):
    lines = out.strip().splitlines()
    n, m, nnz = map(int, lines[0].split())
    rows = []
    for ln in lines[1:]:
        _, rest = ln.split(":", 1)
        rows.append([(int(a), float(b)) for a, b in (t.split("/") for t in rest.split())])
    return n, m, nnz, rows


def test_peek_rand_does_not_consume_the_stream():
    r = run_driver("peek")
    assert r.returncode == 0 and "peek ok" in r.stdout


def test_metis_reader_conventions(tmp_path):
    # fmt 1 = edge weights; vertex 2 lists itself and vertex 1 twice
    p = tmp_path / "g.graph"
    p.write_text("% comment\n4 4 1\n2 0.5 3 2.0\n1 0.5 2 9.0 1 7.0 4 1.5\n1 2.0\n2 1.5\n")
    r = run_driver("read", "metis", p)
    assert r.returncode == 0, r.stderr
    n, m, nnz, rows = parse_read(r.stdout)
    assert (n, m) == (4, 3)
    assert rows == [[(1, 0.5), (2, 2.0)], [(0, 0.5), (3, 1.5)], [(0, 2.0)], [(1, 1.5)]]
    assert "duplicate" in r.stderr and "v--v" in r.stderr


def test_matrix_market_reader_conventions(tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n4 4 5\n"
                 "1 1 5.0\n2 1 -3.0\n3 2 4.0\n4 1 0.5\n4 3 2.0\n")
    r = run_driver("read", "mtx", p)
    assert r.returncode == 0, r.stderr
    n, m, nnz, rows = parse_read(r.stdout)
    assert (n, m, nnz) == (4, 4, 8)
    # edges (1,0,3) (2,1,4) (3,0,0.5) (3,2,2), each appended to both rows in order
    assert rows == [[(1, 3.0), (3, 0.5)], [(0, 3.0), (2, 4.0)], [(1, 4.0), (3, 2.0)], [(0, 0.5), (2, 2.0)]]
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n2 1 1.0\n")
    assert run_driver("read", "mtx", bad).returncode != 0


def write_metis(csr, path):
    with open(path, "w") as f:
        f.write(f"{csr.n} {csr.nnz // 2} 1\n")
        for a, b in csr.ranges:
            f.write(" ".join(f"{int(u) + 1} {float(w):.9g}" for u, w in
                             zip(csr.neighbours[a:b], csr.weights[a:b])) + "\n")


@pytest.mark.gpu
@pytest.mark.parametrize("type_code", [6, 7, 8, 9])
def test_dropin_classes_match_oracle_and_consume_one_seed_per_round(tmp_path, type_code):
    print("dropin type", type_code, flush=True)
    csr = mmamd.gen_rgg(3000, seed=21, weighted=True)
    path = tmp_path / "g.graph"
    write_metis(csr, path)
    out = tmp_path / "out.txt"
    r = run_driver("match", "metis", path, type_code, 3, out)
    assert r.returncode == 0, r.stderr + r.stdout
    lines = out.read_text().splitlines()
    kind, maximal = mmamd.TYPE_TO_KIND[type_code]
    stream = mmamd.reference_seeds(3 * 255 + 1)
    pos = 0
    for k in range(3):
        rounds, ok, size, weight, _ = lines[2 * k].split()
        got = np.array(lines[2 * k + 1].split(), np.int32)
        exp, R, _, _ = oracle.match(csr, kind == mmamd.MM_WEIGHTED, bool(maximal), stream[pos:pos + 255])
        assert int(rounds) == R and int(ok) == 1
        assert got.tolist() == exp.tolist()
        assert (float(weight), int(size)) == pytest.approx(oracle.get_weight(csr, exp))
        pos += R
    assert lines[-1] == f"next_rand {stream[pos]}"


def test_matrix_market_array_storage_and_banner_words(tmp_path):
    # Array format: entry k goes to row k mod (rows - 1), column k div (rows - 1)
    # (the reference's reader wraps its row counter at rows - 1); only entries
    # below the diagonal become edges, so this 3x3 file yields the edge (1, 0)
    # from its second value and nothing else.
    p = tmp_path / "a.mtx"
    p.write_text("%%MatrixMarket matrix array real symmetric\n3 3\n" + "".join(f"{k + 1}.5\n" for k in range(9)))
    r = run_driver("read", "mtx", p)
    assert r.returncode == 0, r.stderr
    n, m, nnz, rows = parse_read(r.stdout)
    assert (n, m, nnz) == (3, 1, 2)
    assert rows == [[(1, 2.5)], [(0, 2.5)], []]
    # banner words are case-insensitive; complex entries weigh |z|; pattern entries weigh 1
    p = tmp_path / "c.mtx"
    p.write_text("%%MatrixMarket MATRIX Coordinate COMPLEX Hermitian\n3 3 3\n2 1 3.0 4.0\n3 3 1 1\n3 1 0 -2\n")
    n, m, nnz, rows = parse_read(run_driver("read", "mtx", p).stdout)
    assert rows == [[(1, 5.0), (2, 2.0)], [(0, 5.0)], [(0, 2.0)]]
    p = tmp_path / "p.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern skew-symmetric\n%\n\n3 3 2\n3 2\n2 1\n")
    n, m, nnz, rows = parse_read(run_driver("read", "mtx", p).stdout)
    assert rows == [[(1, 1.0)], [(2, 1.0), (0, 1.0)], [(1, 1.0)]]
    for bad in ("%%MatrixMarket matrix coordinate real unsymmetric\n2 2 1\n2 1 1\n",   # unknown symmetry
                "%%MatrixMarket vector coordinate real symmetric\n2 2 1\n2 1 1\n",     # not a matrix
                "%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n2 1 1\n",     # not square
                "%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n3 1 1\n",     # outside
                "2 2 1\n2 1 1\n"):                                                     # no banner
        p.write_text(bad)
        assert run_driver("read", "mtx", p).returncode != 0, bad


def test_metis_vertex_sizes_weights_comments_and_short_files(tmp_path):
    # fmt 111: vertex size, one vertex weight, edge weights; '%' lines inside
    # the body are comments; the empty line of the last vertex may be absent
    p = tmp_path / "g.graph"
    p.write_text("%c\n\n4 2 111\n7 3 2 0.25\n% a comment\n8 4 1 0.25 3 2\n9 5 2 2\n")
    r = run_driver("read", "metis", p)
    assert r.returncode == 0, r.stderr
    n, m, nnz, rows = parse_read(r.stdout)
    assert (n, m) == (4, 2)
    assert rows == [[(1, 0.25)], [(0, 0.25), (2, 2.0)], [(1, 2.0)], []]
    for bad in ("3 1\n2\n",                   # two vertex lines missing
                "2 1\n3\n1\n",                # neighbour out of range
                "2 1 1\n2\n1 1.0\n",          # edge weight missing
                "2 1 0 2\n2\n1\n",            # ncon without vertex weights
                "x y\n"):                     # no header
        p.write_text(bad)
        assert run_driver("read", "metis", p).returncode != 0, bad
