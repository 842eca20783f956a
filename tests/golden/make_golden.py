#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ (run from the repository root).

Every expected value comes from a closed form or from scipy/NumPy; nothing is
taken from the reference implementation.

* ao_three_s.npz   -- the reference tests' hand-built 3-s-shell table
  (python/test/test_online_anchor_select.py:19-30: atoms (0,0,0), (0,0,1),
  (0,1,0), alpha = 0.8 / 1.1 / 1.3, c = 1): points and closed-form values
  phi = exp(-alpha r^2) / sqrt(4 pi) and gradients -2 alpha (r - R) phi.
* qrcp_three_s.npz -- the weighted AO-pair matrix of that table on seeded
  points (A[:, g] = sqrt|w_g| vec(phi_g phi_g^T), python/test/test_anchor_select.py:41-48)
  and scipy's dgeqp3 pivots (scipy.linalg.qr(pivoting=True)), plus the
  selection that stops at rank 4 / tolerance 1e-6 as the reference rules give.
* khatri_rao.npz   -- random phi, u, v, sqrt-weights and the sketch
  Y[g, s] = ell^-1/2 sw_g (u_s . phi_g)(v_s . phi_g) by direct NumPy sums.
* pair_error_kat.json -- the known answer of python/test/test_isdf.py:157-164.
"""
import json
from pathlib import Path

import numpy as np
from scipy.linalg import qr

HERE = Path(__file__).resolve().parent
CENTRES = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])
ALPHAS = np.array([0.8, 1.1, 1.3])


def three_s_values(points):
    d = points[:, None, :] - CENTRES[None, :, :]
    r2 = (d ** 2).sum(-1)
    phi = np.exp(-ALPHAS[None, :] * r2) / np.sqrt(4 * np.pi)
    grad = np.moveaxis(-2 * ALPHAS[None, :, None] * d * phi[:, :, None], 2, 0)  # [3][n][3]
    return phi, grad


def main():
    rng = np.random.default_rng(20240611)
    pts = rng.normal(scale=1.0, size=(40, 3))
    phi, grad = three_s_values(pts)
    np.savez(HERE / "ao_three_s.npz", points=pts, values=phi, gradient=grad)

    qp = rng.normal(scale=1.2, size=(60, 3))
    qw = rng.uniform(0.1, 1.0, size=60)
    qw[[7, 31]] = 0.0  # zero-weight points are never selected
    qphi, _ = three_s_values(qp)
    keep = np.flatnonzero(qw != 0)
    pairs = (qphi[keep, :, None] * qphi[keep, None, :]).reshape(len(keep), -1).T * np.sqrt(np.abs(qw[keep]))
    _, r, piv = qr(pairs, pivoting=True, mode="economic")
    diag = np.abs(np.diag(r))
    # Symmetric pairs of 3 AOs span d(d+1)/2 = 6 directions; later pivots only
    # order rounding noise, so the fixture keeps the structural prefix.
    structural = 6
    assert diag[structural - 1] > 1e-8 * diag[0] and diag[structural] < 1e-12 * diag[0]
    rank4 = keep[piv[:4]]
    tol = 1e-6
    ntol = int(np.argmax(diag <= tol * diag[0]))
    np.savez(HERE / "qrcp_three_s.npz", points=qp, weights=qw, pair_matrix=pairs,
             scipy_pivots=piv[:structural], rdiag=diag[:structural], kept=keep, select_rank4=rank4,
             select_tol_1e6=keep[piv[:ntol]])

    n, d, ell = 9, 5, 7
    kphi = rng.normal(size=(n, d))
    u = rng.normal(size=(d, ell))
    v = rng.normal(size=(d, ell))
    sw = rng.uniform(0.2, 1.5, size=n)
    y = np.empty((n, ell))
    for g in range(n):
        for s in range(ell):
            y[g, s] = sw[g] / np.sqrt(ell) * sum(u[a, s] * kphi[g, a] for a in range(d)) * \
                sum(v[a, s] * kphi[g, a] for a in range(d))
    np.savez(HERE / "khatri_rao.npz", phi=kphi, u=u, v=v, sqrt_weights=sw, y=y)

    kat = {"exact_pairs": [[1.0, 2.0], [3.0, 4.0]], "anchor_pairs": [[1.0, 0.0]], "theta": [[1.0, 3.0]],
           "anchor_weights": [1.0], "grid_weights": [2.0, -1.0],
           "l2": float(np.sqrt(24.0)), "relative": float(np.sqrt(24.0 / 35.0))}
    (HERE / "pair_error_kat.json").write_text(json.dumps(kat, indent=1) + "\n")


if __name__ == "__main__":
    main()
