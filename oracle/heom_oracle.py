"""ctypes front end of the CPU oracle (oracle/heom_oracle.cc).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (phi_amd/) never
imports this module.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "heom_oracle.cc")
# HEOM_ORACLE_LIBRARY names another build of the oracle (the sanitizer
# build of phi_amd/build_sanitized.py); the default is the in-tree build.
LIB = os.environ.get("HEOM_ORACLE_LIBRARY") or os.path.join(HERE, "libheom_oracle.so")


def build(force=False):
    """Compile the oracle with g++ (FP64, no FMA contraction, pthreads)."""
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
           "-Wall", "-o", LIB + ".tmp", SRC, "-lpthread"]
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


class _Params(ctypes.Structure):
    _fields_ = [
        ("num_states", ctypes.c_int),
        ("num_baths", ctypes.c_int),
        ("num_terms", ctypes.c_int),
        ("hierarchy_truncation", ctypes.c_int),
        ("hbar", ctypes.c_double),
        ("kb", ctypes.c_double),
        ("temperature", ctypes.c_double),
        ("gamma", ctypes.POINTER(ctypes.c_double)),
        ("lambda_", ctypes.POINTER(ctypes.c_double)),
        ("bath_site", ctypes.POINTER(ctypes.c_int)),
        ("hamiltonian", ctypes.POINTER(ctypes.c_double)),
        ("rho0", ctypes.POINTER(ctypes.c_double)),
        ("normalized", ctypes.c_int),
        ("filter", ctypes.c_double),
        ("threads", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        # (re)built when missing or older than its source; a machine that
        # only runs the tests uses the library built beforehand
        if not os.path.exists(LIB) or (os.path.exists(SRC) and
                                       os.path.getmtime(LIB) < os.path.getmtime(SRC)):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [ctypes.POINTER(_Params)]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_num_nodes.argtypes = [vp]
        L.oracle_num_modes.argtypes = [vp]
        L.oracle_set_threads.argtypes = [vp, ctypes.c_int]
        L.oracle_get_index.argtypes = [vp, ip]
        L.oracle_get_constants.argtypes = [vp, dp, dp, dp]
        L.oracle_get_state.argtypes = [vp, dp]
        L.oracle_set_state.argtypes = [vp, dp]
        L.oracle_get_active.argtypes = [vp, ip]
        L.oracle_set_hamiltonian.argtypes = [vp, dp]
        L.oracle_set_bath.argtypes = [vp, ctypes.c_int, dp]
        L.oracle_set_timelocal.argtypes = [vp, dp]
        L.oracle_rhs.argtypes = [vp, dp, dp]
        L.oracle_steady_matvec.argtypes = [vp, dp, dp]
        L.oracle_rk4.restype = ctypes.c_longlong
        L.oracle_rk4.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, dp, dp, ctypes.c_int, ip]
        L.oracle_rkf45.argtypes = [vp] + [ctypes.c_double] * 6 + [dp, dp, ctypes.c_int, ip, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Problem:
    """A HEOM problem in PHI's units and conventions.

    H and rho0 are N x N complex (H[i, j] = element (i, j)); gamma, lambda_
    have one entry per bath; num_terms is PHI's num_matsubara_terms (the
    .prm MatsubaraTerms + 1); lh is PHI's HierarchyTruncation.  bath_model
    "independent" (site projectors), "diagonal" (bath_ops: M x N diagonal
    entries) or "full" (bath_ops: M x N x N, ops[m][i, j] = element (i, j)).
    """

    def __init__(self, H, rho0, gamma, lambda_, temperature, lh, num_terms=1,
                 hbar=5.3088354, kb=0.6950344, bath_site=None, normalized=True,
                 filter_tolerance=0.0, bath_model="independent", bath_ops=None):
        self.bath_model = bath_model
        self.bath_ops = None if bath_ops is None else np.asarray(bath_ops, dtype=complex)
        self.H = np.asarray(H, dtype=complex)
        self.rho0 = np.asarray(rho0, dtype=complex)
        self.N = self.H.shape[0]
        self.gamma = np.asarray(gamma, dtype=float)
        self.lambda_ = np.asarray(lambda_, dtype=float)
        self.M = len(self.gamma)
        self.temperature = float(temperature)
        self.lh = int(lh)
        self.num_terms = int(num_terms)
        self.hbar = float(hbar)
        self.kb = float(kb)
        self.bath_site = (np.arange(self.M, dtype=np.int32) if bath_site is None
                          else np.asarray(bath_site, dtype=np.int32))
        self.normalized = bool(normalized)
        self.filter_tolerance = float(filter_tolerance)

    @classmethod
    def from_prm(cls, p):
        """From an oracle.prm_oracle.PhiParams (IndependentBath only)."""
        return cls(H=p.time_indepenent_hamiltonian, rho0=p.initial_density_matrix,
                   gamma=p.gamma, lambda_=p.lambda_, temperature=p.temperature,
                   lh=p.hierarchy_truncation_level, num_terms=p.num_matsubara_terms,
                   hbar=p.hbar, kb=p.boltzmann_constant,
                   bath_site=p.multiple_independent_bath_indices,
                   normalized=p.is_rho_normalized, filter_tolerance=p.filter_tolerance)


class Oracle:
    """One oracle instance: hierarchy, prefactors and the ADO state."""

    def __init__(self, prob, threads=1):
        self.prob = prob
        L = lib()
        # Column-major interleaved complex arrays (PHI's storage).
        self._h = np.ascontiguousarray(prob.H.T).view(np.float64).reshape(-1).copy()
        self._r = np.ascontiguousarray(prob.rho0.T).view(np.float64).reshape(-1).copy()
        self._g = prob.gamma.copy()
        self._l = prob.lambda_.copy()
        self._s = prob.bath_site.copy()
        P = _Params()
        P.num_states = prob.N
        P.num_baths = prob.M
        P.num_terms = prob.num_terms
        P.hierarchy_truncation = prob.lh
        P.hbar = prob.hbar
        P.kb = prob.kb
        P.temperature = prob.temperature
        P.gamma = _dp(self._g)
        P.lambda_ = _dp(self._l)
        P.bath_site = self._s.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        P.hamiltonian = _dp(self._h)
        P.rho0 = _dp(self._r)
        P.normalized = 1 if prob.normalized else 0
        P.filter = prob.filter_tolerance
        P.threads = threads
        self.h = L.oracle_create(ctypes.byref(P))
        self.num_nodes = L.oracle_num_nodes(self.h)
        self.D = L.oracle_num_modes(self.h)
        self.N = prob.N
        if prob.bath_model != "independent":
            self.set_bath(prob.bath_model, prob.bath_ops)

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, t):
        lib().oracle_set_threads(self.h, t)

    def index(self):
        out = np.zeros((self.num_nodes, self.D), dtype=np.int32)
        lib().oracle_get_index(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def constants(self):
        c = np.zeros(2 * self.D)
        a = np.zeros(self.D)
        m = np.zeros(self.prob.M)
        lib().oracle_get_constants(self.h, _dp(c), _dp(a), _dp(m))
        return c.view(complex), a, m

    def state(self):
        """All ADOs, shape (nodes, N, N) with [n, i, j] = element (i, j)."""
        out = np.zeros(self.num_nodes * self.N * self.N * 2)
        lib().oracle_get_state(self.h, _dp(out))
        return out.view(complex).reshape(self.num_nodes, self.N, self.N).transpose(0, 2, 1).copy()

    def set_state(self, rho):
        flat = np.ascontiguousarray(np.asarray(rho, dtype=complex).transpose(0, 2, 1))
        lib().oracle_set_state(self.h, _dp(flat.view(np.float64).reshape(-1)))

    def set_bath(self, model, ops):
        """Coupling model and operators (see Problem)."""
        code = {"independent": 0, "diagonal": 1, "full": 2}[model]
        ops = np.asarray(ops, dtype=complex)
        if code == 2:
            ops = ops.transpose(0, 2, 1)  # column-major blocks
        self._ops = np.ascontiguousarray(ops).view(np.float64).reshape(-1).copy()
        lib().oracle_set_bath(self.h, code, _dp(self._ops))

    def set_hamiltonian(self, H):
        """Replace the N x N system Hamiltonian (H[i, j] = element (i, j))."""
        self._hnew = np.ascontiguousarray(np.asarray(H, dtype=complex).T).view(np.float64)
        lib().oracle_set_hamiltonian(self.h, _dp(self._hnew.reshape(-1)))

    def set_timelocal(self, Q):
        """Time-local truncation matrices Q (M, N, N) with Q[m, i, j] =
        element (i, j); None switches the term off."""
        if Q is None:
            lib().oracle_set_timelocal(self.h, None)
            return
        self._tl = np.ascontiguousarray(np.asarray(Q, dtype=complex).transpose(0, 2, 1))
        lib().oracle_set_timelocal(self.h, _dp(self._tl.view(np.float64).reshape(-1)))

    def active(self):
        out = np.zeros(self.num_nodes, dtype=np.int32)
        lib().oracle_get_active(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def _apply(self, fn, x, out=None):
        xin = np.ascontiguousarray(np.asarray(x, dtype=complex).transpose(0, 2, 1))
        if out is None:
            yo = np.zeros_like(xin)
        else:
            yo = np.ascontiguousarray(np.asarray(out, dtype=complex).transpose(0, 2, 1))
        fn(self.h, _dp(xin.view(np.float64).reshape(-1)), _dp(yo.view(np.float64).reshape(-1)))
        return yo.transpose(0, 2, 1).copy()

    def rhs(self, x, out=None):
        """HierarchyUpdater::Get over every node; x, result (nodes, N, N)."""
        return self._apply(lib().oracle_rhs, x, out)

    def steady_matvec(self, x):
        return self._apply(lib().oracle_steady_matvec, x)

    def rk4(self, t_end, dt, out_min, t_start=0.0, max_hist=None):
        """PHI's RK4 loop; returns (steps, times, rho_0 history)."""
        if max_hist is None:
            max_hist = int((t_end - t_start) / dt / max(1, np.ceil(out_min / dt))) + 4
        ht = np.zeros(max_hist)
        hr = np.zeros(max_hist * self.N * self.N * 2)
        nh = ctypes.c_int(0)
        steps = lib().oracle_rk4(self.h, t_start, t_end, dt, out_min, _dp(ht), _dp(hr),
                                 max_hist, ctypes.byref(nh))
        n = nh.value
        rho0 = hr.view(complex).reshape(max_hist, self.N, self.N).transpose(0, 2, 1)[:n].copy()
        return steps, ht[:n].copy(), rho0

    def rkf45(self, t_end, dt0, tol, mindt, out_min, t_start=0.0, max_hist=100000):
        ht = np.zeros(max_hist)
        hr = np.zeros(max_hist * self.N * self.N * 2)
        nh = ctypes.c_int(0)
        st = np.zeros(4)
        lib().oracle_rkf45(self.h, t_start, t_end, dt0, tol, mindt, out_min, _dp(ht),
                           _dp(hr), max_hist, ctypes.byref(nh), _dp(st))
        n = nh.value
        rho0 = hr.view(complex).reshape(max_hist, self.N, self.N).transpose(0, 2, 1)[:n].copy()
        return {"attempted": int(st[0]), "accepted": int(st[1]), "dt": st[2],
                "time": st[3], "times": ht[:n].copy(), "rho0": rho0}
