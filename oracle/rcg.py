"""Test infrastructure: CPU restatement of recycling conjugate gradients (RCG), the checker of the
device solver.  Only tests import it.

Reference: the RCGStrategy of cf3/math/LSS/Trilinos/RCGStrategy.cpp:59-171 hands the system to
Belos::RCGSolMgr (Trilinos; the reference pins no version, cmake/DetectSoftware.cmake:88), with the
parameters "Num Blocks" (directions per cycle), "Num Recycled Blocks", "Maximum Iterations" and
"Convergence Tolerance" (:80-86), and keeps the solver object (so its recycle space) from one solve to
the next (:97-135).  The semi-implicit solver alternates two such strategies between the first and the
later pressure solves of each time step (ns_semi_implicit/implementation/NavierStokesSemiImplicit.cpp:
158-166, 532-537).  Belos is absent from the reference tree; its published algorithm is restated:

* deflated preconditioned CG (Saad, Yeung, Erhel and Guyomarc'h, SIAM J. Sci. Comput. 21, 2000):
  x += W E^-1 W^T r; every search direction p = z + beta p - W E^-1 (AW)^T z is A-orthogonal to W;
* the recycle space (Wang, de Sturler and Paulino, IJNME 69, 2007): after each cycle and each solve,
  W = the k harmonic Ritz vectors of A over span [W, P] with the smallest harmonic Ritz values,
  (AV)^T (AV) y = theta (V^T A V) y, solved here by scipy.linalg.eigh.

M is point Jacobi.  The reference preconditions with an ML smoothed-aggregation hierarchy (:69-79),
which is not restated: iteration counts differ from the reference's, the converged solution does not.

Pinned by tests/test_rcg.py: without recycling the iterates are scipy's CG; with recycling the
solutions equal a direct solve, AW = A W and E = W^T A W hold, every direction is A-orthogonal to W,
b = A w with w in span W converges without iterations, and the smallest Ritz value of W approaches the
smallest eigenvalue of A."""
import numpy as np
import scipy.linalg


class Space:
    """A recycle space: W, AW = A W, E = W^T A W (n x k, n x k, k x k) and the matrix it belongs to."""

    def __init__(self):
        self.W = None
        self.AW = None
        self.E = None
        self.A = None

    @property
    def k(self):
        return 0 if self.W is None else self.W.shape[1]


def _harmonic_ritz(V, AV, kmax):
    """The kmax harmonic Ritz vectors with the smallest values over span V: (W, AW, E) or None."""
    G = AV.T @ AV
    B = V.T @ AV
    G = 0.5 * (G + G.T)
    B = 0.5 * (B + B.T)
    if np.any(np.diag(B) <= 0):
        return None
    try:
        theta, Y = scipy.linalg.eigh(G, B)   # ascending; Y^T B Y = I
    except np.linalg.LinAlgError:
        return None
    kn = min(kmax, V.shape[1])
    Y = Y[:, :kn]
    W = V @ Y
    AW = AV @ Y
    E = Y.T @ B @ Y
    return W, AW, 0.5 * (E + E.T)


def rcg_solve(A, b, x0, space, tol, max_iter, cycle=40, recycle=0, trace=None):
    """Solve A x = b (A symmetric positive definite, scipy sparse or dense) from x0 with the recycle
    space `space` (updated in place); without recycled vectors (recycle = 0) it is Jacobi CG with no
    cycles.  Returns (x, iterations, converged, rel_residual).  trace, if a list, receives every search
    direction."""
    n = b.shape[0]
    if recycle <= 0:   # nothing to recycle: one uninterrupted CG recurrence
        cycle = max_iter + 1
    d = np.asarray(A.diagonal()).ravel()
    dinv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 1.0)
    x = np.array(x0, dtype=float)
    if space.k > recycle:
        space.W = space.AW = space.E = None
    if space.k and space.A is not A:
        space.AW = A @ space.W
        E = space.W.T @ space.AW
        space.E = 0.5 * (E + E.T)
    space.A = A
    bn2 = float(b @ b)
    if bn2 == 0.0:
        return np.zeros(n), 0, True, 0.0
    thr2 = tol * tol * bn2

    def solve_e(v):
        return scipy.linalg.cho_solve(scipy.linalg.cho_factor(space.E), v)

    def project(x, r):
        if space.k:
            mu = solve_e(space.W.T @ r)
            x = x + space.W @ mu
            r = r - space.AW @ mu
        return x, r

    def direction(z, beta=0.0, p=None):
        q = z if p is None else z + beta * p
        if space.k:
            q = q - space.W @ solve_e(space.AW.T @ z)
        return q

    def update(P, Q):
        if not P or recycle <= 0:
            return
        V = np.column_stack(([space.W] if space.k else []) + [np.column_stack(P)])
        AV = np.column_stack(([space.AW] if space.k else []) + [np.column_stack(Q)])
        new = _harmonic_ritz(V, AV, recycle)
        if new is not None:
            space.W, space.AW, space.E = new

    r = b - A @ x
    it, restarts, broke = 0, 0, False
    while True:
        x, r = project(x, r)
        z = dinv * r
        rz, rr = float(r @ z), float(r @ r)
        stop = rr <= thr2
        p = direction(z)
        P, Q, j = [], [], 0
        while not stop and it < max_iter:
            q = A @ p
            pq = float(p @ q)
            if not pq > 0:
                broke = True
                break
            if trace is not None:
                trace.append(p.copy())
            alpha = rz / pq
            x = x + alpha * p
            r = r - alpha * q
            z = dinv * r
            if recycle > 0:
                P.append(p)
                Q.append(q)
            it += 1
            j += 1
            rzn, rr = float(r @ z), float(r @ r)
            beta = rzn / rz
            rz = rzn
            if rr <= thr2:
                break
            if j == cycle:   # end of a cycle: new space, fresh recurrence from the true residual
                update(P, Q)
                P, Q, j = [], [], 0
                r = b - A @ x
                x, r = project(x, r)
                z = dinv * r
                rz, rr = float(r @ z), float(r @ r)
                stop = rr <= thr2
                p = direction(z)
                continue
            p = direction(z, beta, p)
        update(P, Q)
        r = b - A @ x
        rr = float(r @ r)
        conv = rr <= thr2
        if conv or broke or it >= max_iter or restarts >= 3:
            return x, it, conv, np.sqrt(rr / bn2)
        restarts += 1
