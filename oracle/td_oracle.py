"""Time-dependent (annealing) and Hilbert-space-truncated systems in the CPU
oracle: PHI's CubicSpline, HilbertSpaceTruncator and the RK4 loop with
UpdateTimeDependentHamiltonian, around the C++ oracle's RHS.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; the product
(phi_amd/) never imports this module.

Follows
  cubic_spline.cc:1-119                 spline on x_i = i/(m-1), knot
                                        derivatives from the (2,4,..,4,2; 1)
                                        system with PHI's right-hand side
  hilbert_space_truncator.cc:34-129     Truncate, Rotate, UnTruncate,
                                        UpdateWith (lowest NumStates
                                        eigenvectors, zheevr range 'I')
  hierarchy_integrator.cc:1617-1667     UpdateTimeDependentHamiltonian
                                        (MinimumFieldChange gate; last
                                        multipliers start at the largest
                                        float, :518-523)
  hierarchy_integrator.cc:2684-2826     RK4 order: update at t, t + dt/2,
                                        t + dt; rotations of the live
                                        buffers after each update
Eigenvector phases are not fixed by PHI (zheevr's output); here each
eigenvector's first largest-magnitude entry is made real and positive.
Quantities in the eigenbasis are therefore compared up to that convention;
full-space results do not depend on it.
"""

import numpy as np


class CubicSpline:
    def __init__(self, y):
        y = np.asarray(y, dtype=float)
        m = len(y)
        self.x = np.arange(m) * (1.0 / (m - 1))
        self.a = y.copy()
        b = np.zeros(m)
        b[0] = 3.0 * (y[1] - y[0])
        b[m - 1] = 3.0 * (y[m - 1] - y[m - 2])
        for i in range(1, m - 1):
            b[i] = 3.0 * (y[i + 1] - y[i])
        T = np.diag(np.full(m, 4.0)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)
        T[0, 0] = T[m - 1, m - 1] = 2.0
        self.b = np.linalg.solve(T, b)
        self.c = np.zeros(m)
        self.d = np.zeros(m)
        for i in range(m - 1):
            self.c[i] = 3.0 * (self.a[i + 1] - self.a[i]) - 2.0 * self.b[i] - self.b[i + 1]
            self.d[i] = 2.0 * (self.a[i] - self.a[i + 1]) + self.b[i] + self.b[i + 1]

    def __call__(self, x):
        m = len(self.x)
        if x > self.x[m - 1]:
            return self.a[m - 1]
        if x < self.x[0]:
            return self.a[0]
        i = min(int(np.searchsorted(self.x, x, side="right")) - 1, m - 2)
        t = (x - self.x[i]) / (self.x[i + 1] - self.x[i])
        return self.a[i] + self.b[i] * t + self.c[i] * t * t + self.d[i] * t * t * t


def lowest_eigenvectors(H, n):
    """n lowest eigenpairs of Hermitian H (ascending), phase-normalised."""
    w, V = np.linalg.eigh(np.asarray(H, dtype=complex), UPLO="L")
    V = V[:, :n].copy()
    for j in range(n):
        # first entry of largest magnitude, ties within a relative 1e-12
        k, mag = 0, -1.0
        for r, m in enumerate(np.abs(V[:, j])):
            if m > mag * (1 + 1e-12):
                k, mag = r, m
        V[:, j] *= np.conj(V[k, j]) / abs(V[k, j])
    return w[:n], V


class Hamiltonian:
    """Static or annealing H, optionally truncated to n states."""

    def __init__(self, n, H=None, H_initial=None, H_final=None, f_initial=(1.0, 0.0),
                 f_final=(0.0, 1.0), t_total=1.0, min_change=0.01, bath_ops=None):
        """bath_ops: full-space coupling operators (M x F x F) of a FullBath
        model; for a truncated system they are truncated once with the
        initial eigenvectors and then rotated with every basis change
        (hierarchy_integrator.cc:1170-1180, 1661-1667), as PHI does."""
        self.n = n
        self.ops = None if bath_ops is None else np.asarray(bath_ops, dtype=complex)
        self.td = H is None
        self.T = t_total
        self.min_change = min_change
        if self.td:
            self.Hi = np.asarray(H_initial, dtype=complex)
            self.Hf = np.asarray(H_final, dtype=complex)
            self.si = CubicSpline(f_initial)
            self.sf = CubicSpline(f_final)
            full = self.Hi.shape[0]
        else:
            self.Hs = np.asarray(H, dtype=complex)
            full = self.Hs.shape[0]
        self.full = full
        self.old_i = self.old_f = np.finfo(np.float64).max
        self.V = None
        self.R = None
        self._set(self._full_at(0.0), rotate=False)
        if self.ops is not None and self.truncated:
            self.ops = np.array([self.truncate(F) for F in self.ops])

    @property
    def truncated(self):
        return self.full > self.n

    def _full_at(self, alpha):
        if not self.td:
            return self.Hs
        return self.si(alpha) * self.Hi + self.sf(alpha) * self.Hf

    def _set(self, Hfull, rotate):
        if not self.truncated:
            self.H = Hfull.copy()
            return
        old = self.V
        _, self.V = lowest_eigenvectors(Hfull, self.n)
        self.H = self.truncate(Hfull)
        self.R = self.V.conj().T @ old if (rotate and old is not None) else np.eye(self.n)

    def update(self, t):
        if not self.td:
            return False
        alpha = t / self.T
        vi, vf = self.si(alpha), self.sf(alpha)
        # a multiplier that was 0 gives inf (IEEE division, as in PHI): update
        with np.errstate(divide="ignore", invalid="ignore"):
            di, df = abs(vi / self.old_i - 1.0), abs(vf / self.old_f - 1.0)
        if di < self.min_change and df < self.min_change:
            return False
        self.old_i, self.old_f = vi, vf
        self._set(self._full_at(alpha), rotate=True)
        if self.ops is not None and self.truncated:
            self.ops = self.R @ self.ops @ self.R.conj().T
        return True

    def truncate(self, A):
        if not self.truncated:
            return np.asarray(A, dtype=complex).copy()
        return self.V.conj().T @ A @ self.V

    def untruncate(self, a):
        if not self.truncated:
            return np.asarray(a, dtype=complex).copy()
        return self.V @ a @ self.V.conj().T


class TimeLocal:
    """PHI's time-local truncation matrices (hierarchy.cc:264-341,
    hierarchy_integrator.cc:1675-1731).  For bath m and exponential term k

        T_mk <- exp(-nu_k ts) P^dagger (T_mk + ts c_mk V_m) P,  P = 1 - i H ts,

    with H the current (truncated) Hamiltonian as stored (PHI does not
    divide it by hbar here), nu_0 = gamma_m, nu_k = 2 pi k kB T / hbar and
    c_mk the correlation constants; Q_m = sum_k T_mk.  V_m is the matrix
    with a one at element (m, m) for IndependentBath (the bath's own index,
    AddIndependentBathOperator), diag(b_m) for DiagonalBath and the current
    F_m for FullBath.  Steps shorter than 1e-9 are skipped (:1678).  All
    matrices start at zero."""

    def __init__(self, orc):
        prob = orc.prob
        c, _, _ = orc.constants()
        self.M, self.N = prob.M, prob.N
        self.K = len(c) // self.M
        self.c = c.reshape(self.M, self.K)
        nu = 2 * np.pi * prob.kb * prob.temperature / prob.hbar
        self.rate = np.array([[prob.gamma[m] if k == 0 else nu * k for k in range(self.K)]
                              for m in range(self.M)])
        self.model = prob.bath_model
        self.ops = prob.bath_ops
        self.T = np.zeros((self.M, self.K, self.N, self.N), dtype=complex)
        self.saved = self.T.copy()

    def _seed(self, m, ops):
        if self.model == "independent":
            V = np.zeros((self.N, self.N), dtype=complex)
            V[m, m] = 1
            return V
        if self.model == "diagonal":
            return np.diag(self.ops[m])
        return (self.ops if ops is None else ops)[m]

    def step(self, H, ts, ops=None):
        if ts < 1e-9:
            return
        P = np.eye(self.N) - 1j * np.asarray(H) * ts
        Pd = P.conj().T
        for m in range(self.M):
            V = self._seed(m, ops)
            for k in range(self.K):
                X = self.T[m, k] + (ts * self.c[m, k]) * V
                self.T[m, k] = np.exp(-self.rate[m, k] * ts) * (Pd @ X @ P)

    def Q(self):
        return self.T.sum(axis=1)

    def store(self):
        self.saved = self.T.copy()

    def restore(self):
        self.T = self.saved.copy()


def rk4(orc, ham, t_start, t_end, dt, tl=None):
    """PHI's RK4 loop with H updates (classic stage order); the oracle's
    state is advanced in place.  With time-local truncation (tl, a
    TimeLocal) Q advances by dt/2 after k0 (with H(t)) and after k2 (with
    H(t + dt/2)), before the H updates (hierarchy_integrator.cc:2721,
    2784).  Returns the number of steps."""
    rot = lambda x: ham.R @ x @ ham.R.conj().T if ham.truncated else x  # noqa: E731

    def push():
        orc.set_hamiltonian(ham.H)
        if ham.ops is not None:
            orc.set_bath("full", ham.ops)

    def tl_step(ts):
        if tl is not None:
            tl.step(ham.H, ts, ham.ops)
            orc.set_timelocal(tl.Q())

    push()
    if tl is not None:
        orc.set_timelocal(tl.Q())
    t = t_start
    steps = 0
    rho = orc.state()
    while t < t_end:
        if ham.update(t):
            push()
            if ham.truncated:
                rho = rot(rho)
        k0 = orc.rhs(rho)
        y = rho + k0 * (dt / 2)
        tl_step(dt / 2)
        if ham.update(t + dt / 2):
            push()
            if ham.truncated:
                rho, y, k0 = rot(rho), rot(y), rot(k0)
        k1 = orc.rhs(y)
        y = rho + k1 * (dt / 2)
        k2 = orc.rhs(y)
        y = rho + k2 * dt
        tl_step(dt / 2)
        if ham.update(t + dt):
            push()
            if ham.truncated:
                rho, y, k0, k1, k2 = rot(rho), rot(y), rot(k0), rot(k1), rot(k2)
        k3 = orc.rhs(y)
        rho = rho + (dt / 6) * (k0 + 2 * k1 + 2 * k2 + k3)
        t += dt
        steps += 1
    orc.set_state(rho)
    return steps


# PHI's RKF45 tableau (hierarchy_integrator.cc:1926-1946).
_A = [2. / 9, 1. / 3, 3. / 4, 1.0, 5. / 6]
_B = [[2. / 9.], [1. / 12., 1. / 4.], [69. / 128., -243. / 128., 135. / 64.],
      [-17. / 12., 27. / 4., -27. / 5, 16. / 15], [65. / 432., -5. / 16, 13. / 16., 4. / 27., 5. / 144.]]
_C4 = [1. / 9, 9. / 20, 16. / 45, 1. / 12]              # k0, k2, k3, k4
_C5 = [47. / 450, 12. / 25, 32. / 225, 1. / 30, 6. / 25]  # k0, k2, k3, k4, k5


def rkf45(orc, ham, t_start, t_end, dt0, tol, mindt, tl=None):
    """PHI's RKF45 loop (hierarchy_integrator.cc:1797-2573) with H updates
    at t and t + a_s dt before each stage, rotating rho, the stage input and
    the live k's (k1 is not rotated before the last stage, :2257-2263).
    The weights are multiplied by dt once and then rescaled in place by the
    same factors as dt, the end clamp replaces the growth factor, a step is
    accepted when err < tol or dt * 0.5 < mindt.  With time-local
    truncation (tl) Q advances before each H update of stages 1-4 by the
    stage-time increment (a1, a2 - a1, a3 - a2, a4 - a3), is restored to its
    value at t and advanced by a5 before stage 5, advances by dt - a5 and is
    saved after an accepted step, and is restored after a rejected one
    (:2028-2247, 2386-2389, 2476-2480).  Returns (attempts, accepted, final
    time)."""
    rot = lambda x: ham.R @ x @ ham.R.conj().T if ham.truncated else x  # noqa: E731

    def push():
        orc.set_hamiltonian(ham.H)
        if ham.ops is not None:
            orc.set_bath("full", ham.ops)

    def tl_step(ts):
        if tl is not None:
            tl.step(ham.H, ts, ham.ops)
            orc.set_timelocal(tl.Q())

    if not dt0 > 0:
        raise ValueError("the RKF45 timestep must be positive")
    push()
    if tl is not None:
        orc.set_timelocal(tl.Q())
    mydt = dt0
    dt_f = dt0
    dtup, dtdn = 1.1, 0.5
    a = [x * mydt for x in _A]
    b = [[x * dt0 for x in row] for row in _B]
    c4 = [x * dt0 for x in _C4]
    c5 = [x * dt0 for x in _C5]
    t = t_start
    attempts = accepted = 0
    rho = orc.state()
    while t < t_end:
        attempts += 1
        if ham.update(t):
            push()
            if ham.truncated:
                rho = rot(rho)
        k = [orc.rhs(rho)]
        for s in range(5):
            y = rho.copy()
            for j in range(s + 1):
                y = y + k[j] * b[s][j]
            if s < 4:
                tl_step(a[s] - a[s - 1] if s > 0 else a[0])
            elif tl is not None:
                tl.restore()
                tl_step(a[4])
            if ham.update(t + a[s]):
                push()
                if ham.truncated:
                    rho, y = rot(rho), rot(y)
                    k = [kk if (s == 4 and j == 1) else rot(kk) for j, kk in enumerate(k)]
            k.append(orc.rhs(y))
        y4 = rho + k[0] * c4[0] + k[2] * c4[1] + k[3] * c4[2] + k[4] * c4[3]
        y5 = rho + k[0] * c5[0] + k[2] * c5[1] + k[3] * c5[2] + k[4] * c5[3] + k[5] * c5[4]
        err = float(np.abs(y4 - y5).max())
        if err < tol or dt_f * dtdn < mindt:
            rho = y5
            if tl is not None:
                tl_step(mydt - a[4])
                tl.store()
            t += mydt
            accepted += 1
            update = err < tol / 10
            f = dtup
            if t + mydt > t_end:
                update = True
                dtup = (t_end - t) / mydt
                f = dtup
            if update:
                mydt *= f
                a = [x * f for x in a]
                b = [[x * f for x in row] for row in b]
                c4 = [x * f for x in c4]
                c5 = [x * f for x in c5]
            dt_f = mydt
        else:
            if tl is not None:
                tl.restore()
                orc.set_timelocal(tl.Q())
            mydt *= dtdn
            a = [x * dtdn for x in a]
            b = [[x * dtdn for x in row] for row in b]
            c4 = [x * dtdn for x in c4]
            c5 = [x * dtdn for x in c5]
            dt_f = mydt
    orc.set_state(rho)
    return attempts, accepted, t
