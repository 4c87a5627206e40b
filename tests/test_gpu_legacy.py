"""The reference's file-based wrappers (parallelSOSPUpdate,
parallelCombinedGraph; bin/mosp_legacy runs them) against the oracle: every
objective's tree written by the per-objective update, and the combined
graph's distances and tree built from those files, equal the oracle's MOSP
update on the same inputs."""
from __future__ import annotations

import pathlib
import subprocess
import sys

import numpy as np
import pytest

import gen
import oracle_lib as O

ROOT = pathlib.Path(__file__).resolve().parents[1]
LEGACY = ROOT / "mosp-cuda_amd" / "bin" / "mosp_legacy"
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(*args):
    r = subprocess.run([str(LEGACY), *map(str, args)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


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
        run("sosp", pre, din, tin, ins, dele, k, 0, dout, tout)
        d, p = O.read_tree(str(dout), str(tout), g.n)
        assert np.array_equal(d, want["dist"][k]), f"obj{k} distances"
        assert np.array_equal(p, want["parents"][k]), f"obj{k} parents"
        trees.append(tout)
    cd, ct = tmp_path / "comb" / "d.txt", tmp_path / "comb" / "t.txt"
    out = run("combined", pre, K, 0, cd, ct, ",".join(map(str, pref)), *trees)
    assert "parallelCombinedGraph:" in out
    d, p = O.read_tree(str(cd), str(ct), g.n)
    assert np.array_equal(d, want["comb_dist"]), "combined distances"
    assert np.array_equal(p, want["comb_parent"]), "MOSP tree"
