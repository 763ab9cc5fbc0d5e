"""Dense Krylov solvers of the steady-state system, for tests only.

Test infrastructure: the checker of the device steady-state solvers, never
called by the product path.

PHI solves the steady state with BiCGSTAB (BicgstabSteadyState,
src/hierarchy_integrator.cc) and with BiCGSTAB(ell) (BicgstabLSteadyState,
src/hierarchy_integrator.cc:3112-3489).  The functions here restate the
published algorithms on a dense complex matrix:

- bicgstab: H. A. van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992) 631.
- bicgstab_l: G. L. G. Sleijpen and D. R. Fokkema, ETNA 1 (1993) 11,
  Algorithm 3.1 (BiCG part, then the minimal-residual part by modified
  Gram-Schmidt), the order of the reference's loops.

Inner products are (a, b) = a^H b.  Both return the residual norms |b - A x|
after 0, 1, 2, ... iterations (bicgstab_l: after each cycle of ell BiCG
steps) and the final x.

Pins (tests/test_oracle_pins.py): bicgstab_l with ell = 1 takes
bicgstab's iterates; both converge to numpy.linalg.solve's solution on
well-conditioned random systems; the residual they report equals
|b - A x| recomputed from the returned x; mr_part leaves the least-squares
residual of r_0 against r_1..r_ell (numpy.linalg.lstsq) and applies the same
polynomial in A to u_0 and to r_0.
"""

import numpy as np


def bicgstab(A, b, x0, iters):
    x = np.array(x0, dtype=complex).reshape(-1).copy()
    r = b - A @ x
    rh = r.copy()
    p = np.zeros_like(r)
    v = np.zeros_like(r)
    rho_old = alpha = omega = 1.0 + 0j
    res = [np.linalg.norm(r)]
    for _ in range(iters):
        rho_new = np.vdot(rh, r)
        beta = (rho_new / rho_old) * (alpha / omega)
        p = r + beta * (p - omega * v)
        v = A @ p
        alpha = rho_new / np.vdot(rh, v)
        s = r - alpha * v
        t = A @ s
        omega = np.vdot(t, s) / np.vdot(t, t)
        x = x + alpha * p + omega * s
        r = s - omega * t
        rho_old = rho_new
        res.append(np.linalg.norm(r))
    return res, x


def bicgstab_l(A, b, x0, ell, cycles):
    x = np.array(x0, dtype=complex).reshape(-1).copy()
    m = x.size
    r = [np.zeros(m, dtype=complex) for _ in range(ell + 1)]
    u = [np.zeros(m, dtype=complex) for _ in range(ell + 1)]
    r[0] = b - A @ x
    rt = r[0].copy()
    rho0, alpha, omega = 1.0 + 0j, 0j, 1.0 + 0j
    res = [np.linalg.norm(r[0])]
    for _ in range(cycles):
        rho0 = -omega * rho0
        for j in range(ell):
            rho1 = np.vdot(rt, r[j])
            beta = alpha * rho1 / rho0
            rho0 = rho1
            for i in range(j + 1):
                u[i] = r[i] - beta * u[i]
            u[j + 1] = A @ u[j]
            alpha = rho0 / np.vdot(rt, u[j + 1])
            for i in range(j + 1):
                r[i] = r[i] - alpha * u[i + 1]
            r[j + 1] = A @ r[j]
            x = x + alpha * u[0]
        x, omega = mr_part(r, u, x, ell)
        res.append(np.linalg.norm(r[0]))
    return res, x


def mr_part(r, u, x, ell):
    """The minimal-residual part of a cycle on the lists r[0..ell],
    u[0..ell] (updated in place; r[1..ell] are orthogonalised): r[0]
    becomes r[0] - sum_j g_j r_j over the original r_j, with g minimising
    its norm, u[0] becomes u[0] - sum_j g_j u_j and x moves so that b - A x
    follows r[0].  Returns (x, omega = g_ell)."""
    tau = np.zeros((ell + 1, ell + 1), dtype=complex)
    sigma = np.zeros(ell + 1)
    gp = np.zeros(ell + 1, dtype=complex)
    for j in range(1, ell + 1):
        for i in range(1, j):
            tau[i, j] = np.vdot(r[i], r[j]) / sigma[i]
            r[j] = r[j] - tau[i, j] * r[i]
        sigma[j] = np.vdot(r[j], r[j]).real
        gp[j] = np.vdot(r[j], r[0]) / sigma[j]
    g = np.zeros(ell + 1, dtype=complex)
    g[ell] = gp[ell]
    omega = g[ell]
    for j in range(ell - 1, 0, -1):
        g[j] = gp[j] - sum(tau[j, i] * g[i] for i in range(j + 1, ell + 1))
    gpp = np.zeros(ell + 1, dtype=complex)
    for j in range(1, ell):
        gpp[j] = g[j + 1] + sum(tau[j, i] * g[i + 1] for i in range(j + 1, ell))
    x = x + g[1] * r[0]
    r[0] = r[0] - gp[ell] * r[ell]
    u[0] = u[0] - g[ell] * u[ell]
    for j in range(1, ell):
        u[0] = u[0] - g[j] * u[j]
        x = x + gpp[j] * r[j]
        r[0] = r[0] - gp[j] * r[j]
    return x, omega
