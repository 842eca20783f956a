#!/usr/bin/env python3
"""Writes tests/golden/online_water8.json: the oracle's qrcp_online anchors for
(H2O)8 / cc-pVDZ on the library's level-3 Becke grid, rank 8, 16k-point
batches (run from the repository root; several CPU minutes).  The oracle is
the C++ restatement of cpp/src/anchor_select_online_qr.cpp; a grid checksum
guards the fixture against changes of the grid generator."""
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from mp_isdf_amd import grid as gridmod, molecule  # noqa: E402


def main():
    mol = molecule.water_cluster(8, "cc-pvdz")
    xyz, w = gridmod.make_grid(mol, 3)
    t = time.time()
    anchors = oracle.select(mol.atm, mol.bas, mol.env, xyz, w, "ao", max_rank=8, algorithm="qrcp_online",
                            batch_size=16384)
    out = {"workload": "(H2O)8 cc-pvdz level-3 grid, qrcp_online, rank 8, batch 16384",
           "grid_points": int(len(w)), "grid_sum_xyz": float(np.sum(xyz)), "grid_sum_w": float(np.sum(w)),
           "anchors": [int(a) for a in anchors], "oracle_seconds": round(time.time() - t, 1)}
    (Path(__file__).resolve().parent / "online_water8.json").write_text(json.dumps(out, indent=1) + "\n")
    print(out)


if __name__ == "__main__":
    main()
