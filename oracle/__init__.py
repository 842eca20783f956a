"""CPU ORACLE -- test infrastructure, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package, and only as the checker (or the timed CPU baseline).
The product package ``mp_isdf_amd`` never imports it.

``liboracle.so`` (built from ``ao_oracle.cpp`` and ``select_oracle.cpp`` by
``oracle/Makefile``) restates the reference's C++ CPU path in FP64; this module
wraps it with ctypes and adds NumPy restatements of the reference's Python
driver maths:

* ``physical_pair_refit``  python/isdf.py:238-263 (SVD solve, 1e-12 cutoff)
* ``pair_space_error_norms``  python/isdf.py:154-173
* ``stream_pair_space_error_norms``  python/isdf.py:176-191
* ``audit_projection``  cpp/src/anchor_select_randomized_id.cpp:19-88 (pinned
  against LAPACK's dgeqp3 pivots through SciPy and an orthogonal-projector
  residual from NumPy's QR, tests/test_audit_oracle.py)

Reference build: not possible here (CMake FetchContent of Kokkos, LAPACK++,
libcint, pybind11 plus an ILP64 LAPACK), so there is no ``oracle/_ref``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_lib = None

AXIS_CODES = {"ao": 0, "dx": 1, "dy": 2, "dz": 3, "occ": 4}
ALGORITHMS = {"qrcp_dense": 0, "qrcp_online": 1, "rand_sketch": 2, "density": 4, "density_col_norm": 5}


def build(force: bool = False) -> Path:
    """Compile liboracle.so with the committed Makefile (g++ only)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.oracle_last_error.restype = ctypes.c_char_p
        _lib.oracle_select_last_error.restype = ctypes.c_char_p
        _lib.oracle_eri_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, selection=False):
    if rc == 0:
        return
    if selection == "eri":
        msg = lib().oracle_eri_last_error().decode()
    else:
        msg = (lib().oracle_select_last_error() if selection else lib().oracle_last_error()).decode()
    if rc == 1:
        raise ValueError(msg)
    if rc == 2:
        raise OverflowError(msg)
    raise RuntimeError(msg)


def _tables(atm, bas, env):
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    return atm, bas, env


def c2s(l: int, cart: bool = False) -> np.ndarray:
    ncart = (l + 1) * (l + 2) // 2
    ncomp = ncart if cart else 2 * l + 1
    out = np.empty((ncomp, ncart))
    _check(lib().oracle_c2s(ctypes.c_int(l), ctypes.c_int(int(cart)), _p(out)))
    return out


def nao(atm, bas, env, cart=False) -> int:
    atm, bas, env = _tables(atm, bas, env)
    n = ctypes.c_int64()
    _check(lib().oracle_nao(_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)),
                            _p(env), ctypes.c_int64(len(env)), ctypes.c_int(int(cart)), ctypes.byref(n)))
    return n.value


def ao_eval(atm, bas, env, coords, deriv=0, cart=False) -> np.ndarray:
    """deriv 0: (n, nao) values; 1: (3, n, nao) gradient; 2: (4, n, nao) both."""
    atm, bas, env = _tables(atm, bas, env)
    xyz = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    m = nao(atm, bas, env, cart)
    comps = (1, 3, 4)[deriv]
    out = np.empty((comps, len(xyz), m))
    _check(lib().oracle_ao_eval(_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)),
                                _p(env), ctypes.c_int64(len(env)), ctypes.c_int(int(cart)),
                                _p(xyz), ctypes.c_int64(len(xyz)), ctypes.c_int(deriv), _p(out)))
    return out[0] if deriv == 0 else out


def _axes(axes):
    if isinstance(axes, (str, int)):
        axes = [axes]
    return np.asarray([AXIS_CODES[a] if isinstance(a, str) else int(a) for a in axes], dtype=np.int32)


def axis_eval(atm, bas, env, coords, axes, mo_coeff=None, nocc=0, scales=None, cart=False):
    atm, bas, env = _tables(atm, bas, env)
    xyz = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
    codes = _axes(axes)
    mo = None if mo_coeff is None else np.ascontiguousarray(mo_coeff, dtype=np.float64)
    sc = None if scales is None else np.ascontiguousarray(scales, dtype=np.float64)
    width = ctypes.c_int64()
    args = (_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)), _p(env),
            ctypes.c_int64(len(env)), ctypes.c_int(int(cart)))
    _check(lib().oracle_axis_eval(*args, _p(xyz), ctypes.c_int64(0), _p(codes), ctypes.c_int(len(codes)),
                                  _p(mo), ctypes.c_int64(nocc), None, None, ctypes.byref(width)))
    out = np.empty((len(xyz), width.value))
    _check(lib().oracle_axis_eval(*args, _p(xyz), ctypes.c_int64(len(xyz)), _p(codes),
                                  ctypes.c_int(len(codes)), _p(mo), ctypes.c_int64(nocc), _p(sc),
                                  _p(out), ctypes.byref(width)))
    return out


def geqp3(a: np.ndarray, kmax: int | None = None):
    """dgeqp3 semantics on a copy of ``a``: returns (packed QR, 0-based pivots, tau).
    ``kmax`` stops after that many Householder steps (pivots 0..kmax-1 are the
    full factorisation's)."""
    m, n = a.shape
    f = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    jpvt = np.empty(n, dtype=np.int64)
    tau = np.empty(min(m, n))
    if kmax is None:
        _check(lib().oracle_geqp3(ctypes.c_int64(m), ctypes.c_int64(n), _p(f), ctypes.c_int64(m),
                                  _p(jpvt), _p(tau)), selection=True)
    else:
        _check(lib().oracle_geqp3_steps(ctypes.c_int64(m), ctypes.c_int64(n), _p(f), ctypes.c_int64(m),
                                        _p(jpvt), _p(tau), ctypes.c_int64(kmax)), selection=True)
    return f, jpvt - 1, tau


def packed_pairs(features, col):
    """The audit's packed weighted pair matrix (m = d(d+1)/2 rows, one column per
    point): col_g f_mu f_nu for mu <= nu, off-diagonal rows times sqrt(2)
    (cpp/src/anchor_select_randomized_id.cpp:35-46)."""
    f = np.asarray(features, dtype=np.float64)
    mu, nu = np.triu_indices(f.shape[1])
    fac = np.where(mu == nu, 1.0, np.sqrt(2.0))
    return (f[:, mu] * f[:, nu] * fac * np.asarray(col, dtype=np.float64)[:, None]).T.copy()


def projection_error(a, pivots):
    """Relative residual of every column of ``a`` after projection on the span
    of the pivot columns (cpp/src/anchor_select_randomized_id.cpp:49-81): a
    basis by two passes of modified Gram-Schmidt that skips a column left with
    at most 32 eps of its norm, then two projection passes per column."""
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    q = []
    for g in pivots:
        v = a[:, g].copy()
        original = np.linalg.norm(v)
        for _ in range(2):
            for qj in q:
                v -= (qj @ v) * qj
        norm = np.linalg.norm(v)
        if norm <= 32 * np.finfo(float).eps * original:
            continue
        q.append(v / norm)
    total = float(np.sum(a * a))
    r = a.copy()
    for _ in range(2):
        for qj in q:
            r -= np.outer(qj, qj @ r)
    return np.sqrt(float(np.sum(r * r)) / total) if total else 0.0


def audit_projection(features, col, selected):
    """Dense audit of sketch pivots (cpp/src/anchor_select_randomized_id.cpp:19-88):
    (residual of the selected pivots, residual of the first len(selected)
    dense QRCP pivots) on the packed weighted pair matrix."""
    a = packed_pairs(features, col)
    _, piv, _ = geqp3(a)
    return projection_error(a, list(selected)), projection_error(a, list(piv[:len(selected)]))


def omega(d: int, ell: int, seed: int = 7):
    """u, v as (d, ell) row-major arrays (u[a*ell + s])."""
    u = np.empty((d, ell))
    v = np.empty((d, ell))
    _check(lib().oracle_omega(ctypes.c_int64(d), ctypes.c_int64(ell), ctypes.c_uint64(seed),
                              _p(u), _p(v)), selection=True)
    return u, v


def sketch(phi, u, v, sqrt_weights):
    """Y as (count, ell): Y[g, s] = ell^-1/2 sw_g (phi_g.u_s)(phi_g.v_s)."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    count, d = phi.shape
    ell = u.shape[1]
    sw = np.ascontiguousarray(sqrt_weights, dtype=np.float64)
    y = np.empty((count, ell))
    _check(lib().oracle_sketch(_p(phi), ctypes.c_int64(count), ctypes.c_int64(d),
                               _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(v)),
                               ctypes.c_int64(ell), _p(sw), _p(y)), selection=True)
    return y


def select(atm, bas, env, points, weights, axes="ao", tolerance=1e-10, max_rank=0,
           algorithm="qrcp_dense", mo_coeff=None, nocc=0, axis_weights=None,
           batch_size=128, sketch_size=128, oversampling=16, seed=7, cart=False,
           return_stats=False, point_importance=None, ao_screening_threshold=1e-24):
    atm, bas, env = _tables(atm, bas, env)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    codes = _axes(axes)
    mo = None if mo_coeff is None else np.ascontiguousarray(mo_coeff, dtype=np.float64)
    aw = None if axis_weights is None or len(axis_weights) == 0 else \
        np.ascontiguousarray(axis_weights, dtype=np.float64)
    if aw is not None and len(aw) != len(codes):
        raise ValueError("axis weights must match axes")
    cap = max(1, len(pts))
    out = np.empty(cap, dtype=np.int64)
    count = ctypes.c_int64()
    stats = np.zeros(4, dtype=np.int64)
    alg = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    imp = None if point_importance is None else np.ascontiguousarray(point_importance, dtype=np.float64)
    if imp is not None and imp.shape != (len(pts),):
        raise ValueError("point_importance must have shape (ngrid,)")
    _check(lib().oracle_select(
        _p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)), _p(env),
        ctypes.c_int64(len(env)), ctypes.c_int(int(cart)), _p(pts), _p(w), ctypes.c_int64(len(pts)),
        _p(codes), ctypes.c_int(len(codes)), _p(mo), ctypes.c_int64(nocc), _p(aw),
        ctypes.c_double(tolerance), ctypes.c_int64(max_rank), ctypes.c_int(alg),
        ctypes.c_int64(batch_size), ctypes.c_int64(sketch_size), ctypes.c_int64(oversampling),
        ctypes.c_uint64(seed), _p(imp), ctypes.c_double(ao_screening_threshold),
        _p(out), ctypes.c_int64(cap), ctypes.byref(count), _p(stats)),
        selection=True)
    result = out[:count.value].tolist()
    return (result, stats) if return_stats else result


# --------------------------------------------------------------------------
# NumPy restatements of the driver maths (python/isdf.py)
# --------------------------------------------------------------------------

def physical_pair_refit(anchor_pairs, anchor_weights, ao, batch_size):
    """python/isdf.py:238-263: SVD solve C X ~= A, theta = diag(w_J)^-1 X."""
    u, s, vh = np.linalg.svd(anchor_pairs.T, full_matrices=False)
    retained = s > (1e-12 * s[0] if len(s) else 0)
    physical_transform = (vh[retained].T / s[retained]) @ u[:, retained].T
    transform = physical_transform / anchor_weights[:, None]
    theta = np.empty((len(anchor_weights), len(ao)), dtype=np.float64)
    batch_size = min(batch_size, max(1, 8_000_000 // anchor_pairs.shape[1]))
    coefficient_max = 0.0
    for first in range(0, len(ao), batch_size):
        block = ao[first:first + batch_size]
        pairs = (block[:, :, None] * block[:, None, :]).reshape(len(block), -1)
        physical_coeff = physical_transform @ pairs.T
        coefficient_max = max(coefficient_max, float(np.max(np.abs(physical_coeff))))
        theta[:, first:first + len(block)] = physical_coeff / anchor_weights[:, None]
    diagnostics = {
        "retained_rank": int(np.count_nonzero(retained)),
        "condition": float(s[0] / s[-1]) if len(s) and s[-1] > 0 else float("inf"),
        "singular_values": s,
        "max_abs_physical_coefficient": coefficient_max,
    }
    return theta, transform, diagnostics


def pair_space_error_norms(pair_grid, pair_anchor, anchor_weights, theta, grid_weights,
                           block_size=256):
    """python/isdf.py:154-173."""
    weighted_anchor = pair_anchor * anchor_weights[:, None]
    residual_sq = exact_sq = 0.0
    for first in range(0, len(grid_weights), block_size):
        last = min(first + block_size, len(grid_weights))
        exact = pair_grid[first:last]
        reconstructed = theta[:, first:last].T @ weighted_anchor
        weight = np.abs(grid_weights[first:last])
        residual_sq += np.einsum("g,gi,gi->", weight, exact - reconstructed,
                                 exact - reconstructed, optimize=True)
        exact_sq += np.einsum("g,gi,gi->", weight, exact, exact, optimize=True)
    absolute = np.sqrt(residual_sq)
    relative = absolute / np.sqrt(exact_sq) if exact_sq else (0.0 if absolute == 0 else np.inf)
    return {"l2": absolute, "relative": relative}


def stream_pair_space_error_norms(left, right, pair_anchor, anchor_weights, theta, grid_weights):
    """python/isdf.py:176-191 (one AO-pair matrix of scratch per point)."""
    anchor = (pair_anchor * anchor_weights[:, None]).reshape(
        len(anchor_weights), left.shape[1], right.shape[1])
    residual_sq = exact_sq = 0.0
    for g in range(len(grid_weights)):
        exact = np.outer(left[g], right[g])
        reconstructed = np.einsum("k,kij->ij", theta[:, g], anchor, optimize=True)
        weight = abs(grid_weights[g])
        residual_sq += weight * np.sum((exact - reconstructed) ** 2)
        exact_sq += weight * np.sum(exact ** 2)
    absolute = np.sqrt(residual_sq)
    relative = absolute / np.sqrt(exact_sq) if exact_sq else (0.0 if absolute == 0 else np.inf)
    return {"l2": absolute, "relative": relative}


def threads() -> int:
    """Cores the OpenMP oracle uses (OMP_NUM_THREADS or the affinity mask)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    return len(os.sched_getaffinity(0))


# --------------------------------------------------------------------------
# Grid Coulomb metric and J/K (python/isdf.py:48-67, :76-92, :130-151)
# --------------------------------------------------------------------------

def grid_cross_metric(left, right, coords, weights):
    """python/isdf.py:76-92 (C loop in coulomb_oracle.cpp)."""
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    n = len(weights)
    assert left.shape[1] == n and right.shape[1] == n and coords.shape == (n, 3)
    out = np.zeros((left.shape[0], right.shape[0]))
    rc = lib().oracle_grid_cross_metric(_p(left), ctypes.c_int64(left.shape[0]), _p(right),
                                        ctypes.c_int64(right.shape[0]), _p(coords), _p(weights),
                                        ctypes.c_int64(n), _p(out))
    if rc != 0:
        raise ValueError("invalid grid metric sizes")
    return out


def grid_coulomb_metric(interp_vecs, coords, weights):
    """python/isdf.py:48-67: the cross metric of interp_vecs with itself, symmetrised."""
    metric = grid_cross_metric(interp_vecs, interp_vecs, coords, weights)
    return 0.5 * (metric + metric.T)


def evaluate_isdf_jk(x, z, dm):
    """python/isdf.py:130-151 (validation left to the callers' tests)."""
    projected_density = x @ dm @ x.T
    j = x.T @ ((z @ np.diag(projected_density))[:, None] * x)
    k = x.T @ (z * projected_density) @ x
    return 0.5 * (j + j.T), 0.5 * (k + k.T)


def s_functions(atm, bas, env):
    """Contracted s functions of an s-only basis table (PySCF/libcint slots:
    bas = [atom, l, nprim, nctr, kappa, ptr_exp, ptr_coeff, 0], atm =
    [charge, ptr_coord, ...]).  Returns [(centre[3], exps[nprim], coefs[nprim])]
    with the s angular factor 1/sqrt(4 pi) folded into the coefficients, so
    phi(r) = sum_p coefs[p] exp(-exps[p] |r - centre|^2)."""
    out = []
    for row in np.asarray(bas):
        atom, l, nprim, nctr, _, pexp, pcoef = (int(v) for v in row[:7])
        if l != 0:
            raise ValueError("s_functions handles s shells only")
        centre = np.asarray(env[int(atm[atom][1]):int(atm[atom][1]) + 3], dtype=np.float64)
        exps = np.asarray(env[pexp:pexp + nprim], dtype=np.float64)
        for c in range(nctr):
            coefs = np.asarray(env[pcoef + c * nprim:pcoef + (c + 1) * nprim]) / np.sqrt(4.0 * np.pi)
            out.append((centre, exps, coefs))
    return out


def _boys0(t):
    """F0(t) = (1/2) sqrt(pi / t) erf(sqrt t), F0(0) = 1."""
    import math
    return 1.0 - t / 3.0 if t < 1e-12 else 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))


def s_overlap(funcs):
    """Analytic overlap of contracted s Gaussians: (pi/p)^(3/2) exp(-ab/p |AB|^2)."""
    nao = len(funcs)
    s = np.zeros((nao, nao))
    for u, (ca, ea, wa) in enumerate(funcs):
        for v, (cb, eb, wb) in enumerate(funcs):
            r2 = float(np.sum((ca - cb) ** 2))
            for a, x in zip(ea, wa):
                for b, y in zip(eb, wb):
                    p = a + b
                    s[u, v] += x * y * (np.pi / p) ** 1.5 * np.exp(-a * b / p * r2)
    return s


def s_eri(funcs):
    """Analytic two-electron integrals (uv|wx) of contracted s Gaussians:
    primitive (ab|cd) = 2 pi^(5/2) / (p q sqrt(p + q)) K_ab K_cd F0(pq/(p+q) |P - Q|^2)
    (Boys 1950; Szabo and Ostlund eq. A.41)."""
    prims = []  # per pair (u, v): list of (p, P, prefactor)
    nao = len(funcs)
    for u, (ca, ea, wa) in enumerate(funcs):
        row = []
        for v, (cb, eb, wb) in enumerate(funcs):
            r2 = float(np.sum((ca - cb) ** 2))
            lst = []
            for a, x in zip(ea, wa):
                for b, y in zip(eb, wb):
                    p = a + b
                    lst.append((p, (a * ca + b * cb) / p, x * y * np.exp(-a * b / p * r2)))
            row.append(lst)
        prims.append(row)
    eri = np.zeros((nao,) * 4)
    for u in range(nao):
        for v in range(nao):
            for w in range(nao):
                for x in range(nao):
                    total = 0.0
                    for p, pc, kp in prims[u][v]:
                        for q, qc, kq in prims[w][x]:
                            t = p * q / (p + q) * float(np.sum((pc - qc) ** 2))
                            total += 2 * np.pi ** 2.5 / (p * q * np.sqrt(p + q)) * kp * kq * _boys0(t)
                    eri[u, v, w, x] = total
    return eri


def select_cvt(points, weights, cvt_weights, max_rank, seed=7, cvt_max_iter=100, cvt_tol=1e-5):
    """cpp/src/anchor_select_cvt.cpp:17-204: sorted distinct grid indices."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    imp = np.ascontiguousarray(cvt_weights, dtype=np.float64)
    if len(w) != len(pts) or len(imp) != len(pts):
        raise ValueError("CVT points, quadrature weights and CVT weights must have equal lengths")
    out = np.empty(max(1, int(max_rank)), dtype=np.int64)
    count = ctypes.c_int64()
    _check(lib().oracle_select_cvt(_p(pts), _p(w), _p(imp), ctypes.c_int64(len(pts)), ctypes.c_int64(max_rank),
                                   ctypes.c_uint64(seed), ctypes.c_int64(cvt_max_iter), ctypes.c_double(cvt_tol),
                                   _p(out), ctypes.byref(count)), selection=True)
    return out[:count.value].tolist()


def rbrp_ao(ao, weights, batch_size=128, block_size=16, tolerance=1e-10, max_rank=0, seed=7, tau_block=None,
            rank_rtol=1e-12, refresh_every=8, svd_rtol=1e-12, compute_id=False, branch="host"):
    """cpp/src/anchor_select_rbrp.cpp:186-318 (RBRP) and :135-184 (the ID),
    with LAPACK's dgeqp3/dorgqr (scipy.linalg.qr) and gesdd (numpy.linalg.svd)
    as the reference calls them, and libstdc++'s discrete_distribution.

    branch="device" restates the reference's device branch instead
    (cpp/src/anchor_select_rbrp_device.cpp:327-491), which the driver takes on
    a GPU build (python/isdf.py:572-576): each draw is
    uniform_real_distribution(0, total) over the prefix sums of the unmasked
    residuals and takes the first positive, unmasked index whose prefix exceeds
    it (the first positive one if rounding leaves none, :129-160); a block asks
    for min(block, cap - rank) draws and records how many it got; the new
    basis is an unpivoted QR of the accepted rows, projected and factorised
    twice (:431-451).  Prefix sums run left to right here."""
    import scipy.linalg
    ao = np.ascontiguousarray(ao, dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    n, nao = ao.shape
    if (not n or not nao or batch_size <= 0 or block_size <= 0 or not np.isfinite(tolerance) or tolerance < 0
            or not np.isfinite(rank_rtol) or rank_rtol <= 0 or not np.isfinite(svd_rtol) or svd_rtol < 0
            or (tau_block is not None and (not np.isfinite(tau_block) or not 0 <= tau_block < 1))):
        raise ValueError("invalid RBRP dimensions, batch size or tolerances")
    if not np.all(np.isfinite(w)) or np.any(w == 0):
        raise ValueError("RBRP requires finite nonzero weights")
    if not np.all(np.isfinite(ao)):
        raise ValueError("RBRP requires finite real AOs")
    p = nao * (nao + 1) // 2
    cap = min(p, n, max_rank if max_rank else min(n, p))
    tau = (1.0 / block_size if block_size > 1 else 0.0) if tau_block is None else tau_block
    sw = np.sqrt(np.abs(w))
    iu, ju = np.triu_indices(nao)
    fac = np.where(iu == ju, 1.0, np.sqrt(2.0))
    rows = sw[:, None] * ao[:, iu] * ao[:, ju] * fac
    d = np.abs(w) * np.sum(ao * ao, axis=1) ** 2
    total = float(d.sum())
    target = tolerance * tolerance * total
    q = np.zeros((0, p))
    l = np.zeros((n, 0))
    chosen = np.zeros(n, dtype=bool)
    indices, history = [], []
    status = "rank_cap"
    if branch not in ("host", "device"):
        raise ValueError("branch must be 'host' or 'device'")
    lib().oracle_rng_create.restype = ctypes.c_void_p
    lib().oracle_rng_discrete.restype = ctypes.c_int64
    lib().oracle_rng_uniform.restype = ctypes.c_double
    lib().oracle_rng_uniform.argtypes = [ctypes.c_void_p, ctypes.c_double]

    def device_draw(mask):
        live = np.where(mask, 0.0, d)
        prefix = np.cumsum(live)  # sequential left-to-right sums
        tot = float(prefix[-1])
        if not (tot > 0) or not np.isfinite(tot):
            return n
        target = lib().oracle_rng_uniform(ctypes.c_void_p(rng), tot)
        ok = (~mask) & (d > 0)
        hit = np.flatnonzero(ok & (prefix > target))
        if hit.size:
            return int(hit[0])
        first = np.flatnonzero(ok)
        return int(first[0]) if first.size else n
    rng = lib().oracle_rng_create(ctypes.c_uint64(seed))

    def residual():
        e = rows - l @ q
        r = np.sum(e * e, axis=1)
        r[chosen] = 0
        return r
    try:
        iteration = 0
        while True:
            old = len(indices)
            if old and (d.sum() <= target or (refresh_every and iteration % refresh_every == 0)):
                d = residual()
            if d.sum() <= target:
                status = "tolerance"
                break
            if old >= cap:
                break
            if branch == "device":
                mask = np.zeros(n, dtype=bool)
                sample = []
                for _ in range(min(block_size, cap - old)):
                    idx = device_draw(mask)
                    if idx == n:
                        break
                    mask[idx] = True
                    sample.append(idx)
                if not sample:
                    status = "no_candidates"
                    break
                b = len(sample)
            else:
                active = [i for i in range(n) if d[i] > 0 and not chosen[i]]
                if not active:
                    status = "no_candidates"
                    break
                b = min(block_size, cap - old, len(active))
                sample = []
                for _ in range(b):
                    wl = np.ascontiguousarray(d[active])
                    pos = lib().oracle_rng_discrete(ctypes.c_void_p(rng), _p(wl), ctypes.c_int64(len(wl)))
                    sample.append(active.pop(pos))
            v = rows[sample].T.copy()
            original2 = float(np.sum(v * v))
            for _ in range(2):
                for j in range(b):
                    for t in range(old):
                        v[:, j] -= (v[:, j] @ q[t]) * q[t]
            qf, rf, piv = scipy.linalg.qr(v, mode="economic", pivoting=True)
            tails = np.zeros(b)
            for r in range(b - 1, -1, -1):
                tails[r] = (tails[r + 1] if r + 1 < b else 0.0) + float(np.sum(rf[r, r:] ** 2))
            if not (tails[0] > 0) or not np.isfinite(tails[0]):
                status = "numerical_rank"
                break
            keep = 0
            for i in range(b):
                if tails[i] >= tau * tails[0]:
                    keep = i + 1
            cutoff = rank_rtol * np.sqrt(original2)
            for i in range(keep):
                if abs(rf[i, i]) <= cutoff:
                    keep = i
                    break
            if not keep:
                status = "numerical_rank"
                break
            if branch == "device":
                # fresh rows of the accepted samples; (project twice, unpivoted QR) twice
                acc = rows[[sample[piv[j]] for j in range(keep)]].T.copy()
                for _ in range(2):
                    for _ in range(2):
                        if old:
                            acc -= q.T @ (q @ acc)
                    acc = scipy.linalg.qr(acc, mode="economic")[0]
                qb = acc.T.copy()
            else:
                qb = qf[:, :keep].T.copy()
                for j in range(keep):
                    col = qb[j]
                    for _ in range(2):
                        for t in range(old):
                            col -= (q[t] @ col) * q[t]
                        for t in range(j):
                            col -= (qb[t] @ col) * qb[t]
                    norm = np.linalg.norm(col)
                    if not (norm > 0) or not np.isfinite(norm):
                        raise RuntimeError("RBRP block lost orthogonality")
                    qb[j] = col / norm
            dots = rows @ qb.T
            dots[chosen] = 0
            d = np.where(chosen, 0.0, np.maximum(0.0, d - np.sum(dots * dots, axis=1)))
            l = np.hstack([l, dots])
            q = np.vstack([q, qb])
            for j in range(keep):
                idx = sample[piv[j]]
                indices.append(idx)
                chosen[idx] = True
                d[idx] = 0
            history.append([len(indices), b, keep, np.sqrt(d.sum() / total) if total else 0.0])
            iteration += 1
    finally:
        lib().oracle_rng_destroy(ctypes.c_void_p(rng))
    d = residual()
    relative = float(np.sqrt(d.sum() / total)) if total else 0.0
    if relative <= tolerance:
        status = "tolerance"
    elif status == "tolerance":
        status = "roundoff_limited"
    out = {"indices": indices, "relative_residual": relative, "status": status, "condition": 0.0,
           "retained_rank": 0, "history": np.array(history).reshape(-1, 4)}
    k = len(indices)
    if not compute_id or not k:
        return out
    a = l[indices]
    u, s, vt = np.linalg.svd(a)
    out["condition"] = float(s[0] / s[-1]) if s[-1] else float("inf")
    r = int(np.count_nonzero(s > svd_rtol * s[0]))
    if not r:
        raise RuntimeError("RBRP anchor block has zero numerical rank")
    out["retained_rank"] = r
    inv = (vt[:r].T / s[:r]) @ u[:, :r].T
    j = np.asarray(indices)
    theta = ((l @ inv) * (sw[j] / w[j])[None, :] / sw[:, None]).T
    theta[:, j] = np.diag(1.0 / w[j])
    packed = inv.T @ q  # [a, packed]
    packed = packed * (sw[j] / w[j])[:, None]
    packed = packed / fac[None, :]
    transform = np.zeros((k, nao, nao))
    transform[:, iu, ju] = packed
    transform[:, ju, iu] = packed
    out["interp_vecs"] = theta
    out["transform"] = transform.reshape(k, nao * nao)
    return out


# ---- integral-direct Cholesky Coulomb metric (eri_oracle.cpp) -------------

_CHOLESKY_ALGORITHMS = {"unblocked_greedy": 0, "unblocked_uniform": 1}


def _table_args(atm, bas, env, cart):
    atm, bas, env = _tables(atm, bas, env)
    keep = (atm, bas, env)
    return keep, (_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)), _p(env),
                  ctypes.c_int64(len(env)), ctypes.c_int(int(cart)))


def eri_prim(powers, exps, centres) -> float:
    """Primitive cartesian ERI (ab|cd) of unnormalised Gaussians
    (x - A_x)^{a_x} (y - A_y)^{a_y} (z - A_z)^{a_z} exp(-alpha |r - A|^2):
    powers (4, 3), exps (4,), centres (4, 3)."""
    pw = np.ascontiguousarray(powers, dtype=np.int32).reshape(12)
    ex = np.ascontiguousarray(exps, dtype=np.float64).reshape(4)
    xyz = np.ascontiguousarray(centres, dtype=np.float64).reshape(12)
    out = ctypes.c_double()
    _check(lib().oracle_eri_prim(_p(pw), _p(ex), _p(xyz), ctypes.byref(out)), "eri")
    return out.value


def boys(nmax: int, t: float) -> np.ndarray:
    out = np.empty(nmax + 1)
    _check(lib().oracle_boys(ctypes.c_int(nmax), ctypes.c_double(t), _p(out)), "eri")
    return out


def eri_packed(atm, bas, env, cart=False) -> np.ndarray:
    """Packed ERI matrix (mu nu | la si) over mu >= nu, la >= si, pair index mu (mu + 1) / 2 + nu."""
    keep, args = _table_args(atm, bas, env, cart)
    m = nao(atm, bas, env, cart)
    npair = m * (m + 1) // 2
    out = np.zeros((npair, npair))
    _check(lib().oracle_eri_packed(*args, _p(out)), "eri")
    return out


def cholesky_eri(atm, bas, env, tolerance=1e-12, max_rank=0, algorithm="unblocked_greedy", seed=0, cart=False):
    """(factors[rank, npair], pivots, residual[npair]) as mp_isdf.cholesky_eri returns them
    (cpp/src/cholesky_eri_greedy.cpp:53-175)."""
    if algorithm not in _CHOLESKY_ALGORITHMS:
        raise ValueError("unknown Cholesky algorithm")
    keep, args = _table_args(atm, bas, env, cart)
    m = nao(atm, bas, env, cart)
    npair = m * (m + 1) // 2
    capacity = max_rank if 0 < max_rank <= npair else npair
    factors = np.zeros((capacity, npair))
    pivots = np.zeros(capacity, dtype=np.int64)
    residual = np.zeros(npair)
    rank = ctypes.c_int64()
    _check(lib().oracle_cholesky_eri(*args, ctypes.c_double(tolerance), ctypes.c_int64(max_rank),
                                     ctypes.c_int(_CHOLESKY_ALGORITHMS[algorithm]), ctypes.c_uint64(seed),
                                     ctypes.c_int64(capacity), _p(factors), _p(pivots), _p(residual),
                                     ctypes.byref(rank)), "eri")
    r = rank.value
    return factors[:r].copy(), pivots[:r].copy(), residual


def coulomb_metric(atm, bas, env, transform, algorithm="unblocked_greedy", tolerance=1e-12, max_rank=0,
                   batch_size=64, seed=0, cart=False):
    """transform[nvec, nao * nao] -> (nvec, nvec) (cpp/src/coulomb_metric.cpp:20-98)."""
    if algorithm not in _CHOLESKY_ALGORITHMS:
        raise ValueError("unknown Cholesky algorithm")
    keep, args = _table_args(atm, bas, env, cart)
    m = nao(atm, bas, env, cart)
    t = np.ascontiguousarray(transform, dtype=np.float64)
    if t.ndim != 2 or t.shape[1] != m * m:
        raise ValueError("expected transform[nvec,nao*nao]")
    out = np.zeros((t.shape[0], t.shape[0]))
    _check(lib().oracle_coulomb_metric(*args, _p(t), ctypes.c_int64(t.shape[0]),
                                       ctypes.c_int(_CHOLESKY_ALGORITHMS[algorithm]), ctypes.c_double(tolerance),
                                       ctypes.c_int64(max_rank), ctypes.c_int64(batch_size), ctypes.c_uint64(seed),
                                       _p(out)), "eri")
    return out
