"""The match driver's GPU types (6-9) against the oracle and against the host
round matchers (10, 11) that share their semantics; the per-round trace."""
import numpy as np
import pytest

import oracle
from cli_util import distinct_weight_graph, parse, run, write_metis
from mmamd import reference_seeds

pytestmark = pytest.mark.gpu

SEEDS = reference_seeds(600)


def test_gpu_types_through_the_driver(tmp_path):
    csr = distinct_weight_graph(20000, 10, seed=6)
    path = str(tmp_path / "g20000.graph")
    write_metis(path, csr)
    rc, out, err = run(["-m", "6 7 8 9 10 11", "-t", "-a", "2", "--threads", "4", path])
    assert rc == 0, err
    head, rows = parse(out)
    # fixed types: 20 seeds per repetition; maximal types: one per executed round
    for t, weighted in ((6, False), (8, True)):
        ms = [oracle.match(csr, weighted, False, SEEDS[20 * k:20 * (k + 1)])[0] for k in range(2)]
        ws = [oracle.get_weight(csr, m) for m in ms]
        assert rows[(t, 256)][0] == np.mean([x[1] for x in ws])
        assert abs(rows[(t, 256)][1] - np.mean([x[0] for x in ws])) <= 2e-5 * np.mean([x[0] for x in ws])
    for t, host, weighted in ((7, 10, False), (9, 11, True)):
        m1, r1, _, _ = oracle.match(csr, weighted, True, SEEDS)
        m2 = oracle.match(csr, weighted, True, SEEDS[r1:])[0]
        ws = [oracle.get_weight(csr, m) for m in (m1, m2)]
        assert rows[(t, 256)][0] == np.mean([x[1] for x in ws])
        assert rows[(host, 4)][:2] == rows[(t, 256)][:2]
    assert out.count("0 can still be matched.") >= 8


def test_intermediate_trace(tmp_path):
    csr = distinct_weight_graph(5000, 8, seed=7)
    path = str(tmp_path / "g5000.graph")
    write_metis(path, csr)
    rc, out, err = run(["-m", "9", "--intermediate", path])
    assert rc == 0, err
    lines = [ln.split("\t") for ln in out.strip().splitlines()]
    assert lines[0] == ["0", "0", "0"]
    m, R, _, mat = oracle.match(csr, True, True, SEEDS)
    assert len(lines) == R + 1
    assert [int(ln[0]) for ln in lines[1:]] == list(range(1, R + 1))
    assert [int(ln[2]) for ln in lines[1:]] == [int(x) for x in mat]
