"""ORACLE — TEST INFRASTRUCTURE ONLY.

GCRO-DR (Parks, de Sturler, Mackey, Johnson, Maiti, SIAM J. Sci. Comput.
28(5), 2006) restated from the reference's solver
(crates/sonium-solvers/crates/sonium-solvers-iterative/src/gcro_dr.rs) in
numpy, step for step, including its small dense eigensolver
(gmres_dr.rs:1306-1660, 1857-1950) and that solver's behaviour on complex
input: the Householder reduction to Hessenberg form takes a real alpha and
then zeroes everything below the subdiagonal (gmres_dr.rs:1880-1950), so on
complex data the Ritz problem it solves is not exactly similar to the
augmented matrix.  Only tests/ import this module.

Where this restatement departs from the reference, on purpose:
- numerically real trailing 2x2 blocks with a complex-conjugate eigenvalue
  pair: the reference switches to a real Francis double-shift step and a
  quasi-triangular eigenvector extraction (gmres_dr.rs:1360-1400,1661-1760);
  here the complex single-shift iteration continues (a complex Wilkinson
  shift resolves the pair).  Only operators with real-valued augmented
  matrices reach that branch; Burton-Miller matrices are complex.
- the small LU solve of G_top^H w = e_l uses numpy's LAPACK solve; the
  reference's damping fallback (diagonal + 1e-12) is kept for singular G_top.
"""

from __future__ import annotations

import numpy as np


def _norm(x):
    return float(np.linalg.norm(x))


def givens_rotation(a, b):
    """gmres_common.rs:91-105."""
    if abs(b) < 1e-30:
        return 1.0 + 0j, 0j
    if abs(a) < 1e-30:
        return 0j, 1.0 + 0j
    r = np.sqrt(abs(a) ** 2 + abs(b) ** 2)
    return a / r, b / r


def solve_upper_triangular(h, g, k):
    """gmres_common.rs:108-127: back substitution on the rotated Hessenberg,
    skipping pivots below 1e-30."""
    y = np.zeros(k, dtype=np.complex128)
    for i in range(k - 1, -1, -1):
        s = g[i] - np.dot(h[i, i + 1:k], y[i + 1:k])
        if abs(h[i, i]) > 1e-30:
            y[i] = s / h[i, i]
    return y


def solve_upper_triangular_dense(r, g, k):
    """gcro_dr.rs:862-880."""
    y = np.zeros(k, dtype=np.complex128)
    for i in range(k - 1, -1, -1):
        s = g[i] - np.dot(r[i, i + 1:k], y[i + 1:k])
        if abs(r[i, i]) > 1e-30:
            y[i] = s / r[i, i]
    return y


def unapply_givens(h_rot, cs, sn, steps):
    """gcro_dr.rs:884-905: undo the rotations to recover the raw Hessenberg."""
    h = h_rot.copy()
    for j in range(steps - 1, -1, -1):
        if j >= len(cs):
            continue
        c, s = cs[j], sn[j]
        hj = h[j, j:steps].copy()
        hj1 = h[j + 1, j:steps].copy()
        h[j, j:steps] = c * hj - np.conj(s) * hj1
        h[j + 1, j:steps] = np.conj(s) * hj + np.conj(c) * hj1
    return h


def dense_to_hessenberg(a):
    """gmres_dr.rs:1857-1950: Householder reduction with a real alpha =
    -sign(Re x0) ||x||, then every entry below the subdiagonal set to zero."""
    n = a.shape[0]
    h = a.astype(np.complex128).copy()
    q = np.eye(n, dtype=np.complex128)
    for j in range(max(n - 2, 0)):
        x = h[j + 1:, j]
        x_norm = np.sqrt(np.sum(np.abs(x) ** 2))
        if x_norm < 1e-30:
            continue
        x0 = x[0]
        alpha = -x_norm if x0.real >= 0.0 else x_norm
        v = x.copy()
        v[0] = x0 - alpha
        v_sq = np.sum(np.abs(v) ** 2)
        if v_sq < 1e-30:
            continue
        beta = 2.0 / v_sq
        # left: rows j+1.., columns j..
        dots = np.conj(v) @ h[j + 1:, j:]
        h[j + 1:, j:] -= beta * np.outer(v, dots)
        # right: all rows, columns j+1..
        dots = h[:, j + 1:] @ v
        h[:, j + 1:] -= beta * np.outer(dots, np.conj(v))
        dots = q[:, j + 1:] @ v
        q[:, j + 1:] -= beta * np.outer(dots, np.conj(v))
    for j in range(n):
        h[j + 2:, j] = 0.0
    return h, q


def wilkinson_shift(a, b, c, d):
    """gmres_dr.rs:1589-1615 (complex square root: no NaN branch)."""
    trace_half = (a + d) * 0.5
    det_part = ((a - d) * 0.5) * ((a - d) * 0.5) + b * c
    disc = np.sqrt(complex(det_part))
    l1, l2 = trace_half + disc, trace_half - disc
    return l1 if abs(l1 - d) < abs(l2 - d) else l2


def schur_back_substitute(t):
    """gmres_dr.rs:1618-1650: eigenvectors of the triangular Schur form, with
    the 1e-10 regularisation of near-degenerate pivots."""
    m = t.shape[0]
    vecs = np.zeros((m, m), dtype=np.complex128)
    for j in range(m):
        vecs[j, j] = 1.0
        for i in range(j - 1, -1, -1):
            s = np.dot(t[i, i + 1:j + 1], vecs[i + 1:j + 1, j])
            diag = t[i, i] - t[j, j]
            if abs(diag) > 1e-30:
                vecs[i, j] = -s / diag
            else:
                vecs[i, j] = -s / (diag + 1e-10)
    return vecs


def hessenberg_qr_eigen(h):
    """gmres_dr.rs:1306-1460: shifted QR on the Hessenberg matrix with the
    reference's deflation rule (the largest i with |t[i,i-1]| below
    1e-14 (|t[i,i]| + |t[i-1,i-1]|) ends the active window at i), Wilkinson
    shifts and Givens sweeps; eigenvalues from the diagonal, eigenvectors by
    back substitution, mapped back by the accumulated rotations."""
    m = h.shape[0]
    t = h.astype(np.complex128).copy()
    q = np.eye(m, dtype=np.complex128)
    p = m
    for _ in range(200 * m):
        if p <= 1:
            break
        deflated = False
        for i in range(p - 1, 0, -1):
            sub = abs(t[i, i - 1])
            ds = abs(t[i, i]) + abs(t[i - 1, i - 1])
            if sub < 1e-14 * (ds if ds > 0.0 else 1.0):
                t[i, i - 1] = 0.0
                p = i
                deflated = True
                break
        if deflated:
            continue
        sigma = wilkinson_shift(t[p - 2, p - 2], t[p - 2, p - 1], t[p - 1, p - 2], t[p - 1, p - 1])
        for i in range(p):
            t[i, i] -= sigma
        gc, gs = [], []
        for i in range(p - 1):
            c, s = givens_rotation(t[i, i], t[i + 1, i])
            gc.append(c)
            gs.append(s)
            a = t[i, i:].copy()
            b = t[i + 1, i:].copy()
            t[i, i:] = np.conj(c) * a + np.conj(s) * b
            t[i + 1, i:] = -s * a + c * b
        for i in range(p - 1):
            c, s = gc[i], gs[i]
            rows = min(p, i + 3)
            a = t[:rows, i].copy()
            b = t[:rows, i + 1].copy()
            t[:rows, i] = a * c + b * s
            t[:rows, i + 1] = a * (-np.conj(s)) + b * np.conj(c)
            a = q[:, i].copy()
            b = q[:, i + 1].copy()
            q[:, i] = a * c + b * s
            q[:, i + 1] = a * (-np.conj(s)) + b * np.conj(c)
        for i in range(p):
            t[i, i] += sigma
    eigenvalues = np.diag(t).copy()
    vecs = q @ schur_back_substitute(t)
    return eigenvalues, vecs


def _solve_g_top_h(g_top, e):
    """gcro_dr.rs:1060-1090: G_top^H w = e by LU; a singular G_top is damped by 1e-12."""
    a = np.conj(g_top.T).copy()
    try:
        return np.linalg.solve(a, e)
    except np.linalg.LinAlgError:
        a = a + 1e-12 * np.eye(a.shape[0])
        return np.linalg.solve(a, e)


def harmonic_ritz_selection(r_mat, b_block, h_rot, cs, sn, cycle_steps, k_in, k_target):
    """gcro_dr.rs:922-1000: the selected coefficient vectors y (columns,
    length k_in + cycle_steps) of the harmonic Ritz pairs with the smallest
    |theta|, in ascending order; empty when there is nothing to select."""
    if k_target == 0 or cycle_steps == 0:
        return np.zeros((k_in + cycle_steps, 0), dtype=np.complex128)
    total = k_in + cycle_steps
    h_raw = unapply_givens(h_rot, cs, sn, cycle_steps)
    rho = h_raw[cycle_steps, cycle_steps - 1] if cycle_steps < h_raw.shape[0] else 0.0
    g_top = np.zeros((total, total), dtype=np.complex128)
    g_top[:k_in, :k_in] = r_mat[:k_in, :k_in]
    g_top[:k_in, k_in:] = b_block[:k_in, :cycle_steps]
    g_top[k_in:, k_in:] = h_raw[:cycle_steps, :cycle_steps]
    rho_sq = rho * np.conj(rho)
    e = np.zeros(total, dtype=np.complex128)
    e[-1] = 1.0
    w = _solve_g_top_h(g_top, e)
    m_aug = g_top.copy()
    m_aug[:, total - 1] += rho_sq * w
    hh, qh = dense_to_hessenberg(m_aug)
    vals, vecs_h = hessenberg_qr_eigen(hh)
    vecs = qh @ vecs_h
    order = sorted(range(total), key=lambda i: abs(vals[i]))  # stable, as sort_by
    take = min(k_target, total)
    return vecs[:, order[:take]]


def _lift_and_orthonormalise(Y, U, V, k_in, cycle_steps):
    """gcro_dr.rs:1000-1040: psi = U y_u + V y_v, modified Gram-Schmidt
    against the accepted vectors, kept when its norm exceeds 1e-12."""
    new_u = []
    for s in range(Y.shape[1]):
        y = Y[:, s]
        psi = np.zeros(V.shape[1], dtype=np.complex128)
        for i in range(k_in):
            psi += y[i] * U[i]
        for j in range(cycle_steps):
            psi += y[k_in + j] * V[j]
        for prev in new_u:
            psi = psi - np.vdot(prev, psi) * prev
        nrm = _norm(psi)
        if nrm > 1e-12:
            new_u.append(psi / nrm)
    return new_u


def build_c_r(apply, precond, U):
    """gcro_dr.rs:770-860: C = orth(M^-1 A U) by CGS with an unconditional
    second pass, R upper triangular; truncates at the first column whose
    projected norm is below 1e-12.  Returns (C, R, used, spmv)."""
    k_in = len(U)
    C = []
    R = np.zeros((k_in, k_in), dtype=np.complex128)
    spmv = 0
    for j in range(k_in):
        z = precond(apply(U[j]))
        spmv += 1
        if C:
            Cm = np.stack(C)
            c1 = np.conj(Cm) @ z
            R[:len(C), j] = c1
            z = z - Cm.T @ c1
            c2 = np.conj(Cm) @ z
            R[:len(C), j] += c2
            z = z - Cm.T @ c2
        nz = _norm(z)
        if nz < 1e-12:
            break
        R[j, j] = nz
        C.append(z / nz)
    return C, R, len(C), spmv


class Subspace:
    """GcroDrSubspace (gcro_dr.rs:63-110): the recycled vectors U and the target size k."""

    def __init__(self, k):
        self.k = int(k)
        self.u = []

    def copy(self):
        s = Subspace(self.k)
        s.u = [v.copy() for v in self.u]
        return s


def gcro_dr(A, b, subspace: Subspace, *, x0=None, restart=50, max_iterations=100, tol=1e-6, precond=None):
    """gcro_dr_preconditioned (gcro_dr.rs:199-290) -> gcro_dr_core
    (:292-760) -> finalize_gmres_solution (gmres.rs:159-250).  A is a matrix
    or a callable; precond a callable M^-1 (None = identity).  The subspace is
    replaced only when the solve succeeds.  Returns (x, info)."""
    apply = A if callable(A) else (lambda v: A @ v)
    M = precond if precond is not None else (lambda v: v.copy())
    b = np.asarray(b, dtype=np.complex128)
    n = b.shape[0]
    if max_iterations == 0 or restart == 0 or not np.isfinite(tol) or tol <= 0:
        raise ValueError("invalid GCRO-DR configuration")
    if not np.all(np.isfinite(b)):
        raise ValueError("right-hand side contains NaN or infinity")
    if x0 is not None and not np.all(np.isfinite(x0)):
        raise ValueError("initial guess contains NaN or infinity")
    work = subspace.copy()
    x, info = _core(apply, M, b, x0, work, restart, max_iterations, tol)
    # finalize_gmres_solution; a non-finite result is an error and the
    # subspace is left as it was (gcro_dr.rs:250-270)
    ax = apply(x)
    info["spmv_count"] += 1
    if not (np.all(np.isfinite(x)) and np.all(np.isfinite(ax))):
        raise FloatingPointError("GCRO-DR: operator or preconditioner produced a non-finite result")
    rn = _norm(b - ax)
    raw_b = _norm(b)
    true_res = rn / raw_b if raw_b > 0 else rn
    info["residual"] = true_res
    info["backward_error"] = rn / max(_norm(ax) + raw_b, 2.2250738585072014e-308)
    conv = true_res <= tol
    if conv:
        info["status"] = "converged"
    elif info["status"] == "converged":
        info["status"] = "max_iterations"
    info["converged"] = conv
    subspace.u = work.u
    return x, info


def _core(apply, M, b, x0, sub, m, max_iterations, tol):
    n = b.shape[0]
    x = np.zeros(n, dtype=np.complex128) if x0 is None else np.asarray(x0, dtype=np.complex128).copy()
    info = {"iterations": 0, "restarts": 0, "spmv_count": 0, "status": "converged", "recycled_in": 0,
            "recycled_out": 0}
    pb = M(b)
    b_norm = _norm(pb)
    true_b = _norm(b)
    public_b = 1.0 if true_b < 1e-15 else true_b
    if b_norm < 1e-15:
        info["recycled_out"] = len(sub.u)
        return x, info
    C, R, k = [], np.zeros((0, 0), dtype=np.complex128), 0
    if sub.u:
        C, R, used, sp = build_c_r(apply, M, sub.u)
        info["spmv_count"] += sp
        if used == 0:
            sub.u = []
        elif used < len(sub.u):
            sub.u = sub.u[:used]
        k = len(C)
    info["recycled_in"] = k
    Cm = np.stack(C) if k else None
    V = np.zeros((m + 1, n), dtype=np.complex128)
    total, restarts = 0, 0

    def exact_rel():
        info["spmv_count"] += 1
        return _norm(b - apply(x)) / public_b

    def finish(status, res, conv):
        info.update(iterations=total, restarts=restarts, status=status, residual=res, converged=conv)
        info["recycled_out"] = len(sub.u)
        return x, info

    H = np.zeros((m + 1, m), dtype=np.complex128)
    Bk = np.zeros((max(k, 1), m), dtype=np.complex128)
    cs, sn = [], []
    for _outer in range(max_iterations):
        info["spmv_count"] += 1
        w = M(b - apply(x))
        if k:
            c = np.conj(Cm) @ w
            w = w - Cm.T @ c
            z = solve_upper_triangular_dense(R, c, k)
            for i in range(k):
                x = x + z[i] * sub.u[i]
        beta = _norm(w)
        rel_initial = beta / b_norm
        if rel_initial < tol:
            t = exact_rel()
            if t < tol or t < 1e-20:
                return finish("converged", t, True)
        V[0] = w / beta
        H[:] = 0
        if k:
            Bk[:] = 0
        cs, sn = [], []
        g = np.zeros(m + 1, dtype=np.complex128)
        g[0] = beta
        inner_converged = False
        converged_at = None
        for j in range(m):
            total += 1
            info["spmv_count"] += 1
            w = M(apply(V[j]))
            if k:
                c = np.conj(Cm) @ w
                Bk[:k, j] = c
                w = w - Cm.T @ c
                c2 = np.conj(Cm) @ w
                Bk[:k, j] += c2
                w = w - Cm.T @ c2
            before = _norm(w)
            h = np.conj(V[:j + 1]) @ w
            H[:j + 1, j] = h
            w = w - V[:j + 1].T @ h
            wn = _norm(w)
            if wn < 0.707 * before:
                h2 = np.conj(V[:j + 1]) @ w
                H[:j + 1, j] += h2
                w = w - V[:j + 1].T @ h2
                wn = _norm(w)
            H[j + 1, j] = wn
            if wn < 1e-20:
                inner_converged = True
            else:
                V[j + 1] = w / wn
            for i in range(j):
                temp = np.conj(cs[i]) * H[i, j] + np.conj(sn[i]) * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = temp
            c_, s_ = givens_rotation(H[j, j], H[j + 1, j])
            cs.append(c_)
            sn.append(s_)
            H[j, j] = np.conj(c_) * H[j, j] + np.conj(s_) * H[j + 1, j]
            H[j + 1, j] = 0.0
            temp = np.conj(c_) * g[j] + np.conj(s_) * g[j + 1]
            g[j + 1] = -s_ * g[j] + c_ * g[j + 1]
            g[j] = temp
            abs_res = abs(g[j + 1])
            if abs_res / b_norm < tol or abs_res < 1e-20 or inner_converged:
                converged_at = j
                break
        steps = m if converged_at is None else converged_at + 1
        y = solve_upper_triangular(H, g, steps)
        x = x + V[:steps].T @ y
        if k:
            by = -(Bk[:k, :steps] @ y)
            zu = solve_upper_triangular_dense(R, by, k)
            for i in range(k):
                x = x + zu[i] * sub.u[i]
        final_abs = abs(g[steps]) if steps else 0.0
        rel_after = final_abs / b_norm
        if converged_at is not None or rel_after < tol:
            t = exact_rel()
            if t < tol or t < 1e-20:
                Y = harmonic_ritz_selection(R, Bk, H, cs, sn, steps, k, sub.k)
                new_u = _lift_and_orthonormalise(Y, sub.u, V, k, steps)
                if new_u:
                    sub.u = new_u
                return finish("converged", t, True)
            if inner_converged:
                return finish("breakdown", t, False)
            restarts += 1
            continue
        restarts += 1
    Y = harmonic_ritz_selection(R, Bk, H, cs, sn, m, k, sub.k)
    new_u = _lift_and_orthonormalise(Y, sub.u, V, k, m)
    if new_u:
        sub.u = new_u
    t = exact_rel()
    return finish("max_iterations", t, False)
