"""The reference's file-based wrappers (parallelSOSPUpdate,
parallelCombinedGraph; C ABI mosp_parallel_*, called in this process) against
the oracle: every objective's tree written by the per-objective update, and
the combined graph's distances and tree built from those files, equal the
oracle's MOSP update on the same inputs."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import gen
import oracle_lib as O

pytestmark = pytest.mark.gpu


def M():
    import mosp_amd
    return mosp_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", ["uniform-unsafe", "disconnecting", "reweight"])
def test_file_wrappers_equal_oracle(tmp_path, kind):
    g = gen.grid_graph(20, 20, 3, 50, 0.2, 11)
    b = gen.make_batch(kind, g, 60, 12)
    K, pref = 3, [1, 2, 3]
    d0, p0 = gen.initial_trees(g, K)
    want = O.mosp_update(g, b, d0, p0, K, 0, pref)
    assert want["ok"]
    say(f"case legacy {kind} n={g.n} m={g.m} ins={b.n_ins} del={b.n_del}")
    pre = str(tmp_path / "g" / "g")
    O.write_csr(pre, g)
    ins, dele = tmp_path / "c" / "insert.txt", tmp_path / "c" / "delete.txt"
    O.write_batch(str(ins), str(dele), b)
    trees = []
    for k in range(K):
        din, tin = tmp_path / f"in{k}" / "d.txt", tmp_path / f"in{k}" / "t.txt"
        O.write_tree(str(din), str(tin), d0[k], p0[k])
        dout, tout = tmp_path / f"out{k}" / "d.txt", tmp_path / f"out{k}" / "t.txt"
        assert M().parallel_sosp_update(pre, din, tin, ins, dele, k, 0, dout, tout)
        d, p = O.read_tree(str(dout), str(tout), g.n)
        assert np.array_equal(d, want["dist"][k]), f"obj{k} distances"
        assert np.array_equal(p, want["parents"][k]), f"obj{k} parents"
        trees.append(tout)
    cd, ct = tmp_path / "comb" / "d.txt", tmp_path / "comb" / "t.txt"
    assert M().parallel_combined_graph(pre, trees, 0, cd, ct, pref)
    d, p = O.read_tree(str(cd), str(ct), g.n)
    assert np.array_equal(d, want["comb_dist"]), "combined distances"
    assert np.array_equal(p, want["comb_parent"]), "MOSP tree"
