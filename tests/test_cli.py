"""The match command-line driver (bin/match.cpp's flags and output) on the
CPU: the host matcher types 0-5, 10, 11 against the oracle, the output
layout, the gnuplot file and argument errors.  No GPU calls."""
import os

import numpy as np
import pytest

import oracle
from cli_util import distinct_weight_graph, parse, run, write_metis
from mmamd import reference_seeds

SEEDS = reference_seeds(255)


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    csr = distinct_weight_graph(3000, 8, seed=4)
    path = str(tmp_path_factory.mktemp("cli") / "g3000.graph")
    write_metis(path, csr)
    return csr, path


def test_host_types_match_the_oracle(graph_file):
    csr, path = graph_file
    rc, out, err = run(["-m", "0 1 2 3 4 5 10 11", "-t", "-a", "2", "--threads", "1 3", path])
    assert rc == 0, err
    head, rows = parse(out)
    assert head == ["g3000.graph", "3000", str(csr.nnz // 2), str(0x88B81733)]
    def two_reps(weighted):
        # srand(12345) once per type, then one rand() per executed round: the
        # second repetition starts where the first one's rounds left the stream
        m1, r1, _, _ = oracle.match(csr, weighted, True, SEEDS)
        m2 = oracle.match(csr, weighted, True, reference_seeds(r1 + 255)[r1:])[0]
        return [m1, m2]

    exp = {0: [oracle.cpu_random(csr)] * 2, 1: [oracle.cpu_stat_min_deg(csr)] * 2,
           2: [oracle.cpu_karp_sipser(csr)] * 2, 3: [oracle.cpu_min_deg(csr)] * 2, 4: [oracle.cpu_weighted(csr)] * 2,
           5: [oracle.cpu_weighted_edge(csr)] * 2, 10: two_reps(False), 11: two_reps(True)}
    for t, ms in exp.items():
        ws = [oracle.get_weight(csr, m) for m in ms]
        size = np.mean([x[1] for x in ws])
        w = np.mean([x[0] for x in ws])
        threads = [1, 3] if t >= 10 else [256]
        for th in threads:
            got = rows[(t, th)]
            assert got[0] == size, (t, th)
            assert abs(got[1] - w) <= 2e-5 * w, (t, th)   # printed with 6 significant digits
    assert out.count("0 can still be matched.") == 20   # -t: every result verified and maximal


def test_gnuplot_layout(graph_file, tmp_path):
    csr, path = graph_file
    gp = str(tmp_path / "out.dat")
    rc, out, err = run(["-m", "4 11", "--gnuplot", gp, path])
    assert rc == 0, err
    fields = open(gp).read().rstrip("\n").split("\t")
    deg = np.diff(csr.ranges, axis=1).ravel()
    assert fields[:5] == ["g3000.graph", "3000", str(csr.nnz // 2), str(deg.min()), str(deg.max())]
    # per type: threads, 4 x (avg, dev), then barrier / UINT_MAX
    assert len(fields) == 5 + 2 * (1 + 8 + 1)


@pytest.mark.parametrize("args,msg", [([], "You have to specify an input file!"),
                                      (["--bogus", "x.graph"], "Invalid command line arguments!"),
                                      (["nofile.txt"], "An exception occured when reading")])
def test_argument_errors(args, msg):
    rc, out, err = run(args)
    assert rc == 255 and msg in err


def test_unknown_types_are_refused(graph_file):
    _, path = graph_file
    rc, out, err = run(["-m", "12", path])
    assert rc == 255 and "Unknown matching type!" in err


def test_random_shuffle_changes_order_not_validity(graph_file):
    _, path = graph_file
    rc, out, err = run(["-m", "4 11", "-r", "-t", "-a", "3", path])
    assert rc == 0, err
    assert "Invalid matching" not in err


@pytest.mark.parametrize("ext", [".graph", ".mtx"])
def test_bzip2_inputs_read_like_plain_files(graph_file, tmp_path, ext):
    """*.graph.bz2 / *.mtx.bz2 (bin/match.cpp:85-110): decompressed while read,
    also when the file holds several concatenated bzip2 streams; the reported
    name drops the .bz2 suffix."""
    import bz2

    csr, path = graph_file
    if ext == ".mtx":
        r = csr.ranges
        lines = ["%%MatrixMarket matrix coordinate real symmetric", f"{csr.n} {csr.n} {csr.nnz // 2}"]
        for i in range(csr.n):
            for j in range(r[i, 0], r[i, 1]):
                if csr.neighbours[j] < i:
                    lines.append(f"{i + 1} {csr.neighbours[j] + 1} {float(csr.weights[j]):.9g}")
        text = ("\n".join(lines) + "\n").encode()
    else:
        text = open(path, "rb").read()
    plain = str(tmp_path / ("g" + ext))
    open(plain, "wb").write(text)
    one = str(tmp_path / ("one" + ext + ".bz2"))
    open(one, "wb").write(bz2.compress(text))
    multi = str(tmp_path / ("multi" + ext + ".bz2"))
    cut = len(text) // 3
    open(multi, "wb").write(bz2.compress(text[:cut]) + bz2.compress(text[cut:2 * cut]) + bz2.compress(text[2 * cut:]))
    outs = []
    for f in (plain, one, multi):
        rc, out, err = run(["-m", "0 4 5", "-t", f])
        assert rc == 0, err
        head, rows = parse(out)
        outs.append((head[1:], {k: v[:2] for k, v in rows.items()}))  # size, weight (not the timings)
        assert head[0] == os.path.basename(f).replace(".bz2", "")
    assert outs[0] == outs[1] == outs[2]


def test_corrupt_bzip2_input_is_an_error(tmp_path):
    bad = str(tmp_path / "bad.graph.bz2")
    open(bad, "wb").write(b"BZh91AY&SY" + bytes(range(200)))
    rc, out, err = run(["-m", "4", bad])
    assert rc == 255
