"""ISDF driver on the GPU: the hot-path counterpart of the reference's
``build_isdf`` (python/isdf.py:287-672).

Stages and their timing keys follow the reference (test_isdf.py:62-65):
``grid, context, anchor_selection, ao_evaluation, isdf_fit, pair_space_check,
overlap_check, coulomb_metric``.  The refit is the same least-squares problem as the
reference's SVD refit (isdf.py:238-263) solved through the anchor-pair Gram
on the GPU; the overlap check compares with the analytic AO overlap, as the
reference does with PySCF's int1e_ovlp (isdf.py:665-671), and also with the
grid-quadrature overlap.  rand_sketch takes the reference's three Khatri-Rao
sketch techniques (quadrature, density, density_col_norm); the density
importance is an input here.  The Coulomb metric is the reference's
spherical-cell grid metric (isdf.py:48-67) on the GPU, and
``evaluate_isdf_jk`` contracts it into J
and K on the GPU (isdf.py:130-151).  The integral-direct Cholesky metrics and
the PySCF RKS density behind the default density importance are outside this
build.
"""
from __future__ import annotations

import time

import numpy as np

from . import grid as gridmod
from .basis import Molecule

ALGORITHMS = {
    "qrcp_dense": "QRColumnPivoting",
    "qrcp_online": "OnlineQRColumnPivoting",
    "rand_sketch": "RandomizedQuadratureKhatriRao",
    "cvt": "CentroidalVoronoi",
    "rbrp": None,      # mp_isdf.rbrp_ao_device (python/isdf.py:564-590)
    "rbrp_id": None,   # the same, with its interpolative decomposition as the fit
}
METRIC_ALGORITHMS = ("grid", "unblocked_greedy", "unblocked_uniform")
# Khatri-Rao sketch techniques of rand_sketch (python/isdf.py:30-36)
KHATRIRAO_BACKENDS = {
    "quadrature": "RandomizedQuadratureKhatriRao",
    "density": "RandomizedAODensityKhatriRao",
    "density_col_norm": "RandomizedAODensityNormalizedKhatriRao",
}
DENSITY_SKETCHES = ("density", "density_col_norm")
# interp_vecs=None returns theta[k, n] only up to this size; larger fits keep
# theta on the device (the error and theta . w come from the grid pass).
INTERP_VECS_AUTO_BYTES = 8 << 30
AXES = {
    "ao": "AtomicOrbital",
    "dx": "AtomicOrbitalGradX",
    "dy": "AtomicOrbitalGradY",
    "dz": "AtomicOrbitalGradZ",
    "occ": "OccupiedOrbital",
}


def build_isdf(mol: Molecule, *, level: int = 3, grid=None, anchor_rank: int = 0, anchor_tol=None,
               axes=("ao",), axis_weights=None, anchor_algorithm: str = "qrcp_dense", precision: str = "fp64",
               anchor_batch_size: int = 128, fit_batch_size: int = 4096, anchor_sketch_size: int = 128,
               anchor_oversampling: int = 16, anchor_seed: int = 7, anchor_audit: bool = False,
               mo_coeff=None, nocc: int = 0,
               selection_only: bool = False, pair_space_error: bool = False, interp_vecs=None,
               compute_metric: bool = False, metric_algorithm: str = "unblocked_greedy", metric_tol: float = 1e-12,
               metric_rank: int = 0, metric_batch_size: int = 64, metric_seed: int = 0, sketch: str = "quadrature",
               point_importance=None, density_weight: str = "electron_density",
               ao_screening_threshold: float = 1e-24, cvt_weight: str = "electron_density",
               cvt_max_iter: int = 100, cvt_tol: float = 1e-5, rbrp_block_size: int = 16, rbrp_tau_block=None,
               rbrp_rank_rtol: float = 1e-12, rbrp_refresh_every: int = 8, rbrp_svd_rtol: float = 1e-12,
               space: str = "auto", report_timings: bool = False, debug: bool = False):
    """Select anchors on the molecular grid and fit the AO-pair ISDF on the GPU.

    ``anchor_audit=True`` (rand_sketch only; python/isdf.py:358-359, :461-462)
    prints the projection residuals of the selected and of equally many dense
    QRCP pivots on the actual pair matrix; small problems only.

    ``interp_vecs``: True returns theta[k, n] (the reference always does);
    None (default) returns it when it takes at most 8 GiB -- at (H2O)64 with
    12,000 anchors it would be 221 GB -- and False never.  The pair-space
    error and the overlap check need no theta on the host.

    ``compute_metric=True`` adds the anchor Coulomb metric
    ``anchor_weights[:, None] * raw * anchor_weights[None, :]`` as
    ``result["coulomb_metric"]`` (python/isdf.py:674-689), with ``raw`` from
    ``metric_algorithm``: ``"unblocked_greedy"`` (default) or
    ``"unblocked_uniform"`` -- the integral-direct ERI Cholesky metric of the
    fit's transform pinv(anchor pairs^T) / w_anchor (RBRP's transform for
    ``rbrp_id``), ``mp_isdf.coulomb_metric`` with ``metric_tol``,
    ``metric_rank``, ``metric_batch_size`` and ``metric_seed`` -- or
    ``"grid"``, the grid Coulomb metric of theta.

    ``anchor_algorithm="rbrp"`` selects by robust blockwise random pivoting on
    the kept grid points; ``"rbrp_id"`` also takes RBRP's interpolative
    decomposition as the interpolation vectors instead of the refit.
    ``anchor_algorithm="cvt"`` selects centroidal Voronoi sites weighted by
    ``point_importance`` (or |w| with ``cvt_weight="quadrature_weights"``).
    ``sketch="density" | "density_col_norm"`` (rand_sketch only) selects with
    the AO-density Khatri-Rao sketches weighted by ``point_importance`` (ngrid,
    finite, nonnegative; python/isdf.py:231-235).  Without it only
    ``density_weight="quadrature_weights"`` (|w|) is available: the reference
    computes the electron density with a PySCF RKS calculation.
    """
    from . import mp_isdf

    started = time.perf_counter()
    last = started
    timings = {}

    def record(stage):
        nonlocal last
        now = time.perf_counter()
        timings[stage] = now - last
        last = now
        if report_timings and not debug:
            print(f"  {stage:18s} {timings[stage]:10.6f} s  (elapsed {last - started:.6f} s)", flush=True)

    if space not in ("auto", "device", "host"):
        raise ValueError("space must be 'auto', 'host', or 'device'")
    if space == "host":
        raise ValueError("this build runs on the GPU only; space='host' is not available (use 'auto' or 'device')")
    if isinstance(axes, str):
        axes = (axes,)
    try:
        axis_types = [getattr(mp_isdf.AxisType, AXES[a]) if isinstance(a, str) else a for a in axes]
    except KeyError as exc:
        raise ValueError(f"unknown axis: {exc.args[0]!r}; choose from {list(AXES)}") from exc
    if anchor_algorithm not in ALGORITHMS:
        raise ValueError(f"unknown anchor algorithm: {anchor_algorithm!r}")
    if sketch not in KHATRIRAO_BACKENDS:
        raise ValueError(f"unknown Khatri-Rao sketch technique: {sketch!r}")
    density_sketch = anchor_algorithm == "rand_sketch" and sketch in DENSITY_SKETCHES
    if cvt_weight not in ("electron_density", "quadrature_weights"):
        raise ValueError("cvt_weight must be 'electron_density' or 'quadrature_weights'")
    if anchor_audit and anchor_algorithm != "rand_sketch":
        raise ValueError("anchor_audit requires a rand sketch algorithm")
    if anchor_algorithm == "cvt" and anchor_rank < 1:
        raise ValueError("CVT requires a positive anchor_rank")
    if anchor_rank and anchor_tol is not None:
        raise ValueError("anchor_rank and anchor_tol are mutually exclusive")
    if not isinstance(fit_batch_size, (int, np.integer)) or fit_batch_size < 1:
        raise ValueError("fit_batch_size must be a positive integer")
    if pair_space_error and selection_only:
        raise ValueError("pair_space_error requires fitting (selection_only=False)")
    if metric_algorithm not in METRIC_ALGORITHMS:
        raise ValueError(f"unknown metric algorithm: {metric_algorithm!r}")
    if selection_only and compute_metric:
        raise ValueError("selection_only requires compute_metric=False")
    if not isinstance(metric_seed, (int, np.integer)) or not 0 <= metric_seed < 2**64:
        raise ValueError("metric_seed must be an unsigned 64-bit integer")
    if compute_metric and metric_algorithm == "grid" and interp_vecs is False:
        raise ValueError("compute_metric needs the interpolation vectors (interp_vecs=True)")
    tolerance = 1e-10 if anchor_tol is None else float(anchor_tol)

    if grid is None:
        all_coords, all_weights = gridmod.make_grid(mol, level)
    else:
        all_coords, all_weights = (np.ascontiguousarray(a, dtype=np.float64) for a in grid)
    keep = all_weights != 0
    if not np.any(keep):
        raise ValueError("grid has no nonzero weights")
    record("grid")

    importance = None
    if density_sketch or anchor_algorithm == "cvt":
        convention = cvt_weight if anchor_algorithm == "cvt" else density_weight
        if point_importance is None:
            if convention != "quadrature_weights":
                raise NotImplementedError("the electron-density importance needs a PySCF RKS density; pass "
                                          "point_importance or use the 'quadrature_weights' convention")
            point_importance = np.abs(all_weights)
        importance = np.asarray(point_importance, dtype=np.float64)
        if importance.shape != (len(all_coords),) or not np.isfinite(importance).all() or np.any(importance < 0):
            raise ValueError("point_importance must be a finite nonnegative array of shape (ngrid,)")
        importance = np.ascontiguousarray(importance)

    context = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=mol.cart)
    selector = mp_isdf.AnchorSelection(context)
    record("context")
    rbrp_result = None
    if anchor_algorithm in ("rbrp", "rbrp_id"):
        rbrp_result = mp_isdf.rbrp_ao_device(
            selector, all_coords[keep], all_weights[keep], batch_size=anchor_batch_size, block_size=rbrp_block_size,
            tolerance=tolerance, max_rank=anchor_rank, seed=anchor_seed, tau_block=rbrp_tau_block,
            rank_rtol=rbrp_rank_rtol, refresh_every=rbrp_refresh_every, svd_rtol=rbrp_svd_rtol,
            compute_id=anchor_algorithm == "rbrp_id" and not selection_only)
        anchor_indices = np.flatnonzero(keep)[np.asarray(rbrp_result["indices"], dtype=np.intp)]
    else:
        algorithm = getattr(mp_isdf.LocalAnchorSelectionAlgorithm,
                            KHATRIRAO_BACKENDS[sketch] if anchor_algorithm == "rand_sketch"
                            else ALGORITHMS[anchor_algorithm])
        anchor_indices = np.asarray(selector.select(
            all_coords, all_weights, axis_types, tolerance=tolerance, max_rank=anchor_rank, algorithm=algorithm,
            mo_coeff=mo_coeff, nocc=nocc, batch_size=anchor_batch_size, debug=debug,
            sketch_size=anchor_sketch_size, oversampling=anchor_oversampling, seed=anchor_seed, audit=anchor_audit,
            axis_weights=list(axis_weights) if axis_weights is not None else [], precision=precision,
            point_importance=importance if density_sketch else None, ao_screening_threshold=ao_screening_threshold,
            cvt_weights=importance if anchor_algorithm == "cvt" else None, cvt_max_iter=cvt_max_iter,
            cvt_tol=cvt_tol), dtype=np.intp)
    if len(anchor_indices) == 0:
        raise ValueError("anchor selection found no anchors; adjust settings or grid")
    record("anchor_selection")
    anchor_coords = all_coords[anchor_indices]
    anchor_weights = all_weights[anchor_indices]
    coords = all_coords[keep]
    weights = all_weights[keep]
    result = {
        "anchor_indices": anchor_indices,
        "point_importance": importance,
        "rbrp_diagnostics": ({key: rbrp_result[key] for key in ("relative_residual", "status", "history")}
                             if rbrp_result is not None else None),
        "anchor_coords": anchor_coords,
        "anchor_weights": anchor_weights,
        "coordinates": coords,
        "weights": weights,
        "interp_vecs": None,
        "fit_diagnostics": None,
        "pair_space_errors": None,
        "overlap_reconstruction": None,
        "overlap_relative_error": None,
        "overlap_quadrature_relative_error": None,
        "X1": None, "X2": None, "X3": None, "X4": None,
        "coulomb_metric": None,
        "derivative_integrals": {},
        "target_factors": {},
        "timings": timings,
    }
    if selection_only:
        result["elapsed"] = last - started
        return result

    anchor_ao = context.eval_ao(anchor_coords)
    result["X1"] = result["X2"] = result["X3"] = result["X4"] = anchor_ao
    record("ao_evaluation")
    want_theta = (bool(interp_vecs) if interp_vecs is not None
                  else len(anchor_indices) * len(coords) * 8 <= INTERP_VECS_AUTO_BYTES)
    if compute_metric and metric_algorithm == "grid" and not want_theta:
        raise ValueError("compute_metric needs the interpolation vectors, which exceed 8 GiB here; pass "
                         "interp_vecs=True to form them on the host")
    if anchor_algorithm == "rbrp_id":
        # RBRP's ID is the fit (python/isdf.py:606-620): the reference reports
        # no fit diagnostics and keeps the ID's under rbrp_diagnostics["id"].
        theta = rbrp_result["interp_vecs"]
        fit = {"interp_vecs": theta, "theta_weight_sums": theta @ weights, "fit_diagnostics": None}
        result["rbrp_diagnostics"]["id"] = {key: rbrp_result[key] for key in
                                            ("retained_rank", "condition", "relative_residual", "status")}
    else:
        fit = mp_isdf.isdf_fit(context, anchor_coords, anchor_weights, coords, weights, batch_size=fit_batch_size,
                               interp_vecs=want_theta, weight_sums=True)
        if result["rbrp_diagnostics"] is not None:
            result["rbrp_diagnostics"]["id"] = None
    result["interp_vecs"] = fit["interp_vecs"]
    result["fit_diagnostics"] = fit["fit_diagnostics"]
    record("isdf_fit")
    if pair_space_error:
        if anchor_algorithm == "rbrp_id":
            # the error of the returned ID, evaluated as the reference does
            # (python/isdf.py:176-191), on the GPU
            err = mp_isdf.pair_space_error(context, anchor_coords, anchor_weights, coords, weights,
                                           np.ascontiguousarray(fit["interp_vecs"]), batch_size=fit_batch_size)
            result["pair_space_errors"] = {"scalar": err}
            record("pair_space_check")
        else:
            result["pair_space_errors"] = {"scalar": fit["pair_space_error"]}
            timings["pair_space_check"] = 0.0  # evaluated inside the fit's grid pass
    # Overlap check (isdf.py:665-671): S~ = sum_k phi_k phi_k^T w_k (theta_k . w_grid)
    # against the analytic overlap (the reference's mol.intor("int1e_ovlp")).
    # The error against the grid-quadrature overlap isolates the fit's part.
    tsum = fit["theta_weight_sums"]
    overlap = np.einsum("ku,kv,k->uv", anchor_ao, anchor_ao, anchor_weights * tsum, optimize=True)
    exact = context.overlap()
    quadrature = context.quadrature_overlap(coords, weights)  # on the GPU, in grid batches
    result["overlap_reconstruction"] = overlap
    result["overlap_relative_error"] = float(np.linalg.norm(overlap - exact) /
                                             max(np.linalg.norm(exact), 1e-30))
    result["overlap_quadrature_relative_error"] = float(np.linalg.norm(overlap - quadrature) /
                                                        max(np.linalg.norm(quadrature), 1e-30))
    record("overlap_check")
    if compute_metric:
        if metric_algorithm == "grid":
            raw = grid_coulomb_metric(result["interp_vecs"], coords, weights)
        else:
            transform = (rbrp_result["transform"] if anchor_algorithm == "rbrp_id"
                         else metric_transform(anchor_ao, anchor_weights))
            raw = mp_isdf.coulomb_metric(context, np.ascontiguousarray(transform).reshape(len(anchor_ao), -1),
                                         algorithm=metric_algorithm, tolerance=metric_tol, max_rank=metric_rank,
                                         batch_size=metric_batch_size, debug=debug, seed=metric_seed)
        result["coulomb_metric"] = anchor_weights[:, None] * raw * anchor_weights[None, :]
        record("coulomb_metric")
    else:
        timings["coulomb_metric"] = 0.0
    result["elapsed"] = last - started
    return result


def metric_transform(anchor_ao, anchor_weights):
    """The refit's metric transform pinv(P_a^T) / w_anchor with P_a the anchor
    AO pairs [k, nao * nao] (python/isdf.py:238-248: the truncated SVD, singular
    values above 1e-12 of the largest), on the host as the reference forms it."""
    a = np.asarray(anchor_ao, dtype=np.float64)
    pairs = (a[:, :, None] * a[:, None, :]).reshape(len(a), -1)
    u, s, vh = np.linalg.svd(pairs.T, full_matrices=False)
    keep = s > (1e-12 * s[0] if len(s) else 0)
    physical = (vh[keep].T / s[keep]) @ u[:, keep].T
    return physical / np.asarray(anchor_weights, dtype=np.float64)[:, None]


def grid_coulomb_metric(interp_vecs, coords, weights):
    """Spherical-cell grid Coulomb metric of ``interp_vecs[nvec, ngrid]`` (not
    weight scaled), symmetrised: the GPU form of python/isdf.py:48-67.  Weights
    must be nonzero (the self-cell term is c |w|^(-1/3))."""
    from . import mp_isdf
    return mp_isdf.grid_coulomb_metric(interp_vecs, coords, weights)


def grid_cross_metric(left, right, coords, weights):
    """Grid Coulomb metric between two sets of interpolation vectors
    (python/isdf.py:76-92), on the GPU."""
    from . import mp_isdf
    return mp_isdf.grid_cross_metric(left, right, coords, weights)


def evaluate_isdf_jk(result, dm):
    """Return (J, K) for a real symmetric AO density matrix (python/isdf.py:130-151).

    ``result["X1"]`` holds the anchor AO values X[p, mu] and
    ``result["coulomb_metric"]`` the anchor metric Z (both anchor weights
    included); (mu nu|lambda sigma) ~ sum_pq X[p,mu] X[p,nu] Z[p,q] X[q,lambda] X[q,sigma].
    The contractions run on the GPU.
    """
    from . import mp_isdf
    x = np.asarray(result["X1"], dtype=np.float64)
    z = result["coulomb_metric"]
    if z is None:
        raise ValueError("a Coulomb metric is required to evaluate J and K")
    density = np.asarray(dm, dtype=np.float64)
    if density.shape != (x.shape[1], x.shape[1]) or not np.isfinite(density).all():
        raise ValueError("dm must be a finite AO density matrix of shape (nao,nao)")
    if not np.allclose(density, density.T, rtol=1e-12, atol=1e-12):
        raise ValueError("dm must be symmetric")
    if np.shape(z) != (len(x), len(x)):
        raise ValueError("Coulomb metric has the wrong anchor dimensions")
    return mp_isdf.isdf_jk(x, np.asarray(z, dtype=np.float64), density)
