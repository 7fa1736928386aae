"""CPU oracle of the Krylov outer solvers chive offers on the transport
operator (test infrastructure only).

chive builds its outer solver with io::CreateIterativeSolver
(src/io.cpp:179-230): "gmres" is mfem::GMRESSolver and "fgmres" is
mfem::FGMRESSolver, both with kdim 50 unless the deck sets it
(src/io.cpp:206-225).  chive calls it on T = I - D L^{-1} D^T Ms with
b = D L^{-1} q (exe/chive.cpp:702-721), with the DSA operator as the
preconditioner when one is given (exe/chive.cpp:715).  mfem is a third-party
dependency that the reference does not vendor or pin (CMakeLists.txt:46 takes
any installed MFEM).  This file restates the published algorithms of those
two classes (mfem linalg/solvers.cpp, releases 4.x):

* GMRES(m), left-preconditioned: r = M (b - A x), beta = ||r||; the stopping
  level is max(reltol * beta_0, abstol).  Arnoldi with modified Gram-Schmidt
  on w = M A v_i; Givens rotations on the Hessenberg column; the estimate
  |s_{i+1}| is compared with the stopping level after each step.  After m
  steps the iterate is updated, the residual is recomputed and the cycle
  restarts.
* FGMRES(m), right-preconditioned and flexible: z_i = M v_i, w = A z_i, the
  same Arnoldi and rotations on the true residual, x += Z y.

Both count one iteration per operator application inside a cycle and report
(iteration, norm) to a monitor, which chive prints as "transport iterations"
(exe/chive.cpp:40-75).

Pinned in tests/test_krylov_oracle.py: the iterate after k steps is the
minimum-residual point of the explicit Krylov space (least squares on
[r, Ar, ..., A^{k-1} r]); an n x n system is solved exactly in n steps; the
residual history equals scipy.sparse.linalg.gmres's on the same matrix;
restarting, the preconditioned forms and the transport operator (the GMRES
solution is source iteration's fixed point).
"""
import numpy as np


def _rotation(dx, dy):
    """Givens rotation (cs, sn) that zeroes dy against dx."""
    if dy == 0.0:
        return 1.0, 0.0
    if abs(dy) > abs(dx):
        t = dx / dy
        sn = 1.0 / np.sqrt(1.0 + t * t)
        return t * sn, sn
    t = dy / dx
    cs = 1.0 / np.sqrt(1.0 + t * t)
    return cs, t * cs


def _apply(dx, dy, cs, sn):
    return cs * dx + sn * dy, -sn * dx + cs * dy


def _solve_upper(H, s, k):
    """y = H[:k+1, :k+1]^{-1} s[:k+1] by back substitution."""
    y = s[:k + 1].copy()
    for i in range(k, -1, -1):
        y[i] /= H[i, i]
        y[:i] -= H[:i, i] * y[i]
    return y


def _krylov(A, b, M, kdim, abstol, reltol, maxit, x0, flexible, monitor):
    n = b.size
    x = np.zeros(n) if x0 is None else np.array(x0, float)

    r = b.copy() if x0 is None else b - A(x)
    if M is not None and not flexible:
        r = M(r)
    beta = float(np.linalg.norm(r))
    stop = max(reltol * beta, abstol)
    norms = [beta]
    if monitor:
        monitor(0, beta)
    if beta <= stop:
        return x, 0, np.array(norms), True
    m = kdim
    j = 1
    while j <= maxit:
        V = np.zeros((m + 1, n))
        Z = np.zeros((m, n)) if flexible else None
        H = np.zeros((m + 1, m))
        s = np.zeros(m + 1)
        cs = np.zeros(m)
        sn = np.zeros(m)
        V[0] = r / beta
        s[0] = beta
        i = 0
        while i < m and j <= maxit:
            if flexible:
                Z[i] = M(V[i]) if M is not None else V[i]
                w = A(Z[i])
            else:
                w = A(V[i])
                if M is not None:
                    w = M(w)
            for k in range(i + 1):
                H[k, i] = w @ V[k]
                w = w - H[k, i] * V[k]
            H[i + 1, i] = np.linalg.norm(w)
            V[i + 1] = w / H[i + 1, i]
            for k in range(i):
                H[k, i], H[k + 1, i] = _apply(H[k, i], H[k + 1, i], cs[k], sn[k])
            cs[i], sn[i] = _rotation(H[i, i], H[i + 1, i])
            H[i, i], H[i + 1, i] = _apply(H[i, i], H[i + 1, i], cs[i], sn[i])
            s[i], s[i + 1] = _apply(s[i], s[i + 1], cs[i], sn[i])
            res = abs(s[i + 1])
            norms.append(res)
            if monitor:
                monitor(j, res)
            if res <= stop:
                y = _solve_upper(H, s, i)
                x = x + (Z[:i + 1] if flexible else V[:i + 1]).T @ y
                return x, j, np.array(norms), True
            i += 1
            j += 1
        y = _solve_upper(H, s, i - 1)
        x = x + (Z[:i] if flexible else V[:i]).T @ y
        r = b - A(x)
        if M is not None and not flexible:
            r = M(r)
        beta = float(np.linalg.norm(r))
        if beta <= stop:
            # mfem reports the step counter here, one past the cycle's last step;
            # the recomputed residual is recorded at that index
            norms.append(beta)
            return x, j, np.array(norms), True
    return x, maxit, np.array(norms), False


def gmres(A, b, M=None, kdim=50, abstol=0.0, reltol=0.0, maxit=100, x0=None, monitor=None):
    """mfem::GMRESSolver::Mult: returns (x, iterations, norms, converged).
    norms[0] is ||M b|| (x0 = 0) and norms[j] the estimate after step j."""
    return _krylov(A, np.asarray(b, float), M, kdim, abstol, reltol, maxit, x0, False, monitor)


def fgmres(A, b, M=None, kdim=50, abstol=0.0, reltol=0.0, maxit=100, x0=None, monitor=None):
    """mfem::FGMRESSolver::Mult (right, flexible preconditioning)."""
    return _krylov(A, np.asarray(b, float), M, kdim, abstol, reltol, maxit, x0, True, monitor)


def transport_operator(pr, sigma_s):
    """T x = x - D L^{-1} D^T Ms x (TransportOperator::Mult, src/transport_op.hpp:23-33)."""
    return lambda x: x - pr.D(pr.sweep(pr.DT(pr.mass(sigma_s, x))))
