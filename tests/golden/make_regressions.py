"""Write the count-to-infinity regression fixtures of the reference's oracle
suite (src/mospTest.cu:550-595) from the reference's OWN generators.

The three instances are defined by seeds fed to the reference's
generateGraphCSR (src/generateGraphCSR.cu) and generateChangedEdges
(src/generateChangedEdges.cu), whose byte stream comes from libstdc++'s
mt19937 + uniform_int_distribution; the instances can only be reproduced by
running that code.  This script calls it through oracle/_ref/libmospref.so
(built by `make -C oracle` from the reference checkout) and stores, per case:
the graph, the batch (as the reference read it back), and the expected
updated trees = the reference's dijkstraCsrGraph on the reference's
applyChangeBatch result.  Run:  python tests/golden/make_regressions.py
"""
from __future__ import annotations

import json
import pathlib
import sys
import tempfile

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

import oracle_lib as O  # noqa: E402

# (n, m, K, objective, changes, insertPct, maxWeight, graphSeed, changeSeed)
# -- the case table of src/mospTest.cu:555-558.
CASES = [
    (6, 10, 2, 0, 6, 55, 50, 621705, 250813),
    (13, 22, 3, 0, 6, 26, 50, 770968, 694580),
    (9, 17, 2, 1, 6, 81, 50, 115080, 943676),
]


def main() -> None:
    r = O.ref()
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        for idx, (n, m, K, obj, ch, ins, wmax, gseed, cseed) in enumerate(CASES):
            base = f"{tmp}/regression_{idx}"
            prefix = f"{base}/random/graphCsr"
            assert r.ref_generate_graph_csr(n, m, 1, prefix.encode(), K, 1, wmax, gseed)
            ip, dp = f"{base}/c/insert.txt", f"{base}/c/delete.txt"
            assert r.ref_generate_changed_edges(1, wmax, K, n, ch, float(ins),
                                                float(100 - ins), 1, 1, 1, 0,
                                                prefix.encode(), ip.encode(),
                                                dp.encode(), cseed)
            g = O.read_csr(prefix)
            b = O.read_batch(ip, dp, K)
            upd, mask = O.ref_apply_batch(g, b)
            init = [O.ref_dijkstra(g, k, 0) for k in range(K)]
            expect = [O.ref_dijkstra(upd, k, 0) for k in range(K)]
            out.append({
                "case": list(CASES[idx]),
                "n": n, "K": K,
                "row": g.row.tolist(), "col": g.col.tolist(), "w": g.w.tolist(),
                "ins_from": b.ins_from.tolist(), "ins_to": b.ins_to.tolist(),
                "ins_w": b.ins_w.tolist(), "del_from": b.del_from.tolist(),
                "del_to": b.del_to.tolist(),
                "updated_row": upd.row.tolist(), "updated_col": upd.col.tolist(),
                "updated_w": upd.w.tolist(), "increase_mask": mask.tolist(),
                "init_dist": [[int(x) for x in d] for d, _ in init],
                "init_parent": [p.tolist() for _, p in init],
                "expected_dist": [[int(x) for x in d] for d, _ in expect],
                "expected_parent": [p.tolist() for _, p in expect],
            })
    path = HERE / "regressions.json"
    path.write_text(json.dumps({
        "_source": "src/mospTest.cu:550-595 instances, generated by the reference's "
                   "generateGraphCSR/generateChangedEdges/applyChangeBatch/"
                   "dijkstraCsrGraph via tests/golden/make_regressions.py",
        "cases": out}, indent=None))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
