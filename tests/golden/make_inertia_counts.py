"""Writes FP64 Sylvester-inertia fixtures: the number of eigenvalues of the
pencil (H(k), S(k)) strictly inside an absolute interval, at points of a
Y-G-X-M path, for one of the seeded synthetic bundles.

H(k), S(k) come from the oracle's brute-force dense Bloch sum and the counts
from the oracle's dense FP64 Bunch-Kaufman factorisation of H - sigma S at
both interval ends (``oracle.dense.negatives_below``).  With ``--margin d``
the counts at lower -/+ d and upper -/+ d are recorded too: a count is
unambiguous at that margin when no eigenvalue lies within d of either end.

The output file is updated after every k point and finished points are
skipped on a rerun, so a long run can be resumed.  Run from the repository
root, e.g.

    python tests/golden/make_inertia_counts.py bulk3d --interval -1.02 1.02 \
        --path-points 40 --k 0:121 --out tests/golden/bulk3d_counts.json
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from aospectrum_amd import synth  # noqa: E402
from aospectrum_amd.band import ygxm_path  # noqa: E402
from oracle import bloch, dense  # noqa: E402


def _indices(spec: str, total: int) -> list[int]:
    out: list[int] = []
    for part in spec.split(","):
        if ":" in part:
            a, b = part.split(":")
            out.extend(range(int(a or 0), min(total, int(b or total))))
        elif part:
            out.append(int(part))
    return sorted(set(i for i in out if 0 <= i < total))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("bundle", choices=sorted(synth.SPECS))
    ap.add_argument("--interval", nargs=2, type=float, required=True, metavar=("LO", "HI"))
    ap.add_argument("--path-points", type=int, default=40, help="Y-G-X-M points per segment")
    ap.add_argument("--k", default="0:3", help="path indices, e.g. 0:121 or 0,29,30")
    ap.add_argument("--margin", type=float, default=0.0)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    bundle = synth.generate(synth.SPECS[args.bundle])
    n = bundle.n_orbitals
    path = ygxm_path(args.path_points)
    lo, hi = args.interval
    out_path = Path(args.out)
    doc = {"bundle": args.bundle, "interval_ev": [lo, hi], "path_points_per_segment": args.path_points,
           "k_index": [], "kpoints": [], "counts": [], "margin_ev": args.margin, "margin": {}}
    if out_path.exists():
        old = json.loads(out_path.read_text())
        if old.get("bundle") == args.bundle and old.get("interval_ev") == [lo, hi]:
            doc.update({k: old[k] for k in ("k_index", "kpoints", "counts") if k in old})
            doc["margin"] = old.get("margin", {})
            if old.get("margin_ev"):
                doc["margin_ev"] = old["margin_ev"]
            if not doc["k_index"]:  # files written before indices were recorded
                doc["k_index"] = list(range(len(doc["counts"])))
    # n above 20,000: build the dense matrices in complex64 (the bundle's values
    # are float32) to fit host memory; the factorisation is FP64 either way
    dtype = np.complex64 if n > 20000 else np.complex128
    for i in _indices(args.k, path.n_points):
        need_count = i not in doc["k_index"]
        need_margin = args.margin > 0 and str(i) not in doc["margin"]
        if not (need_count or need_margin):
            continue
        k = path.fractional_points[i]
        t = time.time()
        H, S = bloch.dense_bloch(bundle.operators, bundle.basis.atom_offsets, n, k, dtype=dtype)
        if need_count:
            count = dense.negatives_below(H, S, hi) - dense.negatives_below(H, S, lo)
            doc["k_index"].append(int(i))
            doc["kpoints"].append([float(v) for v in k])
            doc["counts"].append(int(count))
        if need_margin:
            d = args.margin
            neg = {name: dense.negatives_below(H, S, s) for name, s in
                   (("lo-", lo - d), ("lo+", lo + d), ("hi-", hi - d), ("hi+", hi + d))}
            doc["margin"][str(i)] = {"inner": neg["hi-"] - neg["lo+"], "outer": neg["hi+"] - neg["lo-"]}
        del H, S
        order = np.argsort(doc["k_index"])
        for key in ("k_index", "kpoints", "counts"):
            doc[key] = [doc[key][j] for j in order]
        doc["margin"] = {str(j): doc["margin"][str(j)] for j in sorted(int(x) for x in doc["margin"])}
        out_path.write_text(json.dumps(doc, indent=1) + "\n")
        c = doc["counts"][doc["k_index"].index(i)]
        print(f"k {i} count {c} margin {doc['margin'].get(str(i))} ({time.time() - t:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
