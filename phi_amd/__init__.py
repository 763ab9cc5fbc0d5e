"""Python bindings of the MI355X HEOM integrator (ctypes over libphi_hip.so).

Mirrors PHI's C++ entry points for this path:
  PhiParameters        struct PhiParameters + ReadCommandLine/ReadInputFile
                       (phi_parameters.h:32-122, phi_parameters.cc:93-888)
  HierarchyIntegrator  constructor, Launch(), density_matrix()
                       (hierarchy_integrator.h:25-33)
  Context              the device hierarchy: batched Get (rhs), RK4 steps,
                       RKF45 attempts, steady state (include/phi_hip.h)
Every call goes through the C ABI; there is no fallback path.  Importing the
module fails loudly when the library has not been built.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PHI_HIP_LIBRARY names another build of the same library (the host-code
# sanitizer build of phi_amd/build_sanitized.py); the default is the
# in-tree build.
LIB_PATH = os.environ.get("PHI_HIP_LIBRARY") or os.path.join(HERE, "libphi_hip.so")
EXE_PATH = os.path.join(HERE, "phi")

RUN_METHODS = ["BROKENINPUT", "NONE", "RK4", "RKF45", "RK4SPECTRUM", "RKF45SPECTRUM",
               "BICGSTABL", "BICGSTAB", "BICGSTABU", "PRINTHIERARCHY", "PRINTMEMORY"]

_lib = None


class PhiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("phi_hip error %d: %s" % (code, msg))
        self.code = code


def lib():
    """The loaded C-ABI library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("%s is missing: run `python phi_amd/build.py` (or "
                          "__graft_entry__.build()) first" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    sz, cs = ctypes.c_size_t, ctypes.c_char_p
    psz = ctypes.POINTER(ctypes.c_size_t)
    fp = ctypes.POINTER(ctypes.c_float)
    sigs = {
        "phi_hip_last_error": (cs, []),
        "phi_hip_version": (ctypes.c_int, []),
        "phi_hip_device_count": (ctypes.c_int, []),
        "phi_hip_main": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(cs)]),
        "phi_hip_params_new": (ctypes.c_int, [ctypes.POINTER(vp)]),
        "phi_hip_params_read": (ctypes.c_int, [cs, cs, ctypes.c_int, ctypes.POINTER(vp)]),
        "phi_hip_params_free": (None, [vp]),
        "phi_hip_params_get_double": (ctypes.c_int, [vp, cs, dp]),
        "phi_hip_params_set_double": (ctypes.c_int, [vp, cs, ctypes.c_double]),
        "phi_hip_params_get_int": (ctypes.c_int, [vp, cs, ip]),
        "phi_hip_params_set_int": (ctypes.c_int, [vp, cs, ctypes.c_int]),
        "phi_hip_params_get_array": (ctypes.c_int, [vp, cs, dp, sz, psz]),
        "phi_hip_params_set_array": (ctypes.c_int, [vp, cs, dp, sz]),
        "phi_hip_params_get_matrix": (ctypes.c_int, [vp, cs, dp, sz, psz]),
        "phi_hip_params_set_matrix": (ctypes.c_int, [vp, cs, dp, sz]),
        "phi_hip_params_get_string": (ctypes.c_int, [vp, cs, cs, sz]),
        "phi_hip_params_set_string": (ctypes.c_int, [vp, cs, cs]),
        "phi_hip_params_fix": (ctypes.c_int, [vp, cs]),
        "phi_hip_params_header": (ctypes.c_int, [vp, ctypes.c_int, cs, sz, psz]),
        "phi_hip_validate": (ctypes.c_int, [cs, cs, ip]),
        "phi_hip_integrator_create": (ctypes.c_int, [vp, ctypes.POINTER(vp)]),
        "phi_hip_integrator_launch": (ctypes.c_int, [vp]),
        "phi_hip_integrator_density_matrix": (ctypes.c_int, [vp, dp, sz]),
        "phi_hip_integrator_state": (ctypes.c_int, [vp, dp, sz]),
        "phi_hip_integrator_info": (ctypes.c_int, [vp, dp, sz]),
        "phi_hip_integrator_destroy": (None, [vp]),
        "phi_hip_create": (ctypes.c_int, [vp, ctypes.c_int, ctypes.POINTER(vp)]),
        "phi_hip_destroy": (None, [vp]),
        "phi_hip_num_nodes": (ctypes.c_int, [vp]),
        "phi_hip_num_modes": (ctypes.c_int, [vp]),
        "phi_hip_get_index": (ctypes.c_int, [vp, ip, sz]),
        "phi_hip_set_state": (ctypes.c_int, [vp, dp, sz]),
        "phi_hip_get_state": (ctypes.c_int, [vp, dp, sz]),
        "phi_hip_get_rho0": (ctypes.c_int, [vp, dp, sz]),
        "phi_hip_rhs": (ctypes.c_int, [vp, dp, dp, sz]),
        "phi_hip_rhs_noalloc": (ctypes.c_int, [vp, dp, dp, sz]),
        "phi_hip_rhs_nodes": (ctypes.c_int, [vp, ip, ctypes.c_int, dp, sz]),
        "phi_hip_rk4_steps": (ctypes.c_int, [vp, ctypes.c_longlong, ctypes.c_double]),
        "phi_hip_rk4_stage_times": (ctypes.c_int, [vp, ctypes.c_longlong, ctypes.c_double, dp]),
        "phi_hip_rk4_steps_timed": (ctypes.c_int, [vp, ctypes.c_longlong, ctypes.c_double, dp]),
        "phi_hip_rkf45_attempt": (ctypes.c_int, [vp, dp, dp]),
        "phi_hip_rkf45_accept": (ctypes.c_int, [vp]),
        "phi_hip_active_flags": (ctypes.c_int, [vp, ip, ctypes.c_int, ip]),
        "phi_hip_steady_state": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int, ip, dp]),
        "phi_hip_steady_state_l": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                  ip, dp]),
        "phi_hip_synchronize": (ctypes.c_int, [vp]),
        "phi_hip_set_launch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int]),
        "phi_hip_set_stage_launch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int]),
        "phi_hip_stage_launch": (ctypes.c_int, [vp, ip, ip]),
        "phi_hip_set_rk4_form": (ctypes.c_int, [vp, ctypes.c_int]),
        "phi_hip_set_graphs": (ctypes.c_int, [vp, ctypes.c_int]),
        "phi_hip_get_graphs": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int)]),
        "phi_hip_ensemble_create": (ctypes.c_int, [vp, ctypes.c_int, dp, ip, ctypes.c_int,
                                                   ctypes.POINTER(vp)]),
        "phi_hip_ensemble_members": (ctypes.c_int, [vp, ip, ctypes.c_int]),
        "phi_hip_ensemble_rho0": (ctypes.c_int, [vp, dp, sz]),
        "phi_hip_rk4_history": (ctypes.c_int, [vp, ctypes.c_longlong, ctypes.c_double,
                                               ctypes.c_int, dp, sz, ip]),
        "phi_hip_launch_counts": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_longlong),
                                                 ctypes.POINTER(ctypes.c_longlong)]),
        "phi_hip_rk4_history_timed": (ctypes.c_int, [vp, ctypes.c_longlong, ctypes.c_double,
                                                     ctypes.c_int, dp, sz, ip, dp]),
        "phi_hip_mlp_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(vp)]),
        "phi_hip_mlp_destroy": (None, [vp]),
        "phi_hip_mlp_param_count": (ctypes.c_longlong, [vp]),
        "phi_hip_mlp_set_params": (ctypes.c_int, [vp, fp, sz]),
        "phi_hip_mlp_get_params": (ctypes.c_int, [vp, fp, sz]),
        "phi_hip_mlp_forward": (ctypes.c_int, [vp, fp, ctypes.c_int, fp]),
        "phi_hip_mlp_gradients": (ctypes.c_int, [vp, fp, fp, ctypes.c_int, dp, fp, sz]),
        "phi_hip_mlp_train": (ctypes.c_int, [vp, fp, fp, ctypes.c_longlong, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_uint, dp]),
        "phi_hip_closure_samples": (ctypes.c_int, [vp, ctypes.c_int, fp, fp, ctypes.c_longlong,
                                                   ctypes.POINTER(ctypes.c_longlong)]),
        "phi_hip_closure_attach": (ctypes.c_int, [vp, vp]),
        "phi_hip_closure_fused": (ctypes.c_int, [vp, ctypes.c_int]),
        "phi_hip_closure_fallback": (ctypes.c_int, [vp, ctypes.c_double,
                                                    ctypes.POINTER(ctypes.c_longlong)]),
        "phi_hip_autotune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, dp, ip]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise PhiError(rc, lib().phi_hip_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _colmajor(m):
    """N x N complex (m[i, j] = element (i, j)) -> interleaved column-major."""
    m = np.asarray(m, dtype=complex)
    return np.ascontiguousarray(m.T).view(np.float64).reshape(-1).copy()


def _from_colmajor(flat, n):
    return np.asarray(flat).view(complex).reshape(n, n).T.copy()


class PhiParameters:
    """Handle on a PhiParameters instance (field names as in PHI)."""

    def __init__(self, handle=None):
        if handle is None:
            h = ctypes.c_void_p()
            _check(lib().phi_hip_params_new(ctypes.byref(h)))
            handle = h
        self._h = handle

    @classmethod
    def read(cls, path, method=None, threads=0, check=True):
        """ReadCommandLine(path, method, threads): parse a `.prm` file."""
        h = ctypes.c_void_p()
        rc = lib().phi_hip_params_read(os.fsencode(path),
                                       method.encode() if method else None, threads,
                                       ctypes.byref(h))
        obj = cls(h)
        if check:
            _check(rc)
        obj.read_status = rc
        return obj

    def __del__(self):
        try:
            if self._h:
                lib().phi_hip_params_free(self._h)
                self._h = None
        except Exception:
            pass

    def get_double(self, name):
        v = ctypes.c_double()
        _check(lib().phi_hip_params_get_double(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_double(self, name, v):
        _check(lib().phi_hip_params_set_double(self._h, name.encode(), float(v)))

    def get_int(self, name):
        v = ctypes.c_int()
        _check(lib().phi_hip_params_get_int(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_int(self, name, v):
        _check(lib().phi_hip_params_set_int(self._h, name.encode(), int(v)))

    def get_array(self, name):
        n = ctypes.c_size_t()
        _check(lib().phi_hip_params_get_array(self._h, name.encode(), None, 0, ctypes.byref(n)))
        out = np.zeros(n.value)
        if n.value:
            _check(lib().phi_hip_params_get_array(self._h, name.encode(), _dp(out), n.value, None))
        return out

    def set_array(self, name, v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        _check(lib().phi_hip_params_set_array(self._h, name.encode(), _dp(a), a.size))

    def get_matrix(self, name):
        """Complex matrix as an N x N array (m[i, j] = element (i, j)), or None."""
        n = ctypes.c_size_t()
        _check(lib().phi_hip_params_get_matrix(self._h, name.encode(), None, 0, ctypes.byref(n)))
        if n.value == 0:
            return None
        out = np.zeros(2 * n.value)
        _check(lib().phi_hip_params_get_matrix(self._h, name.encode(), _dp(out), n.value, None))
        k = int(round(np.sqrt(n.value)))
        return _from_colmajor(out, k)

    def set_matrix(self, name, m):
        flat = _colmajor(m)
        _check(lib().phi_hip_params_set_matrix(self._h, name.encode(), _dp(flat), flat.size // 2))

    def get_string(self, name):
        buf = ctypes.create_string_buffer(4096)
        _check(lib().phi_hip_params_get_string(self._h, name.encode(), buf, 4096))
        return buf.value.decode()

    def set_string(self, name, v):
        _check(lib().phi_hip_params_set_string(self._h, name.encode(), v.encode()))

    def fix(self, method=None):
        """run_method from the CLI word, then FixIncompatibleCombinations."""
        return lib().phi_hip_params_fix(self._h, method.encode() if method else None)

    @property
    def run_method(self):
        return RUN_METHODS[self.get_int("run_method")]

    def header(self, num_nodes):
        n = ctypes.c_size_t()
        _check(lib().phi_hip_params_header(self._h, num_nodes, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(lib().phi_hip_params_header(self._h, num_nodes, buf, n.value + 1, None))
        return buf.value.decode()


def device_count():
    """Visible HIP devices (0 without a GPU)."""
    return lib().phi_hip_device_count()


def validate(kind, text):
    v = ctypes.c_int()
    _check(lib().phi_hip_validate(kind.encode(), text.encode(), ctypes.byref(v)))
    return bool(v.value)


class HierarchyIntegrator:
    """PHI's HierarchyIntegrator on one MI355X: Launch() + density_matrix()."""

    def __init__(self, params):
        self.params = params
        h = ctypes.c_void_p()
        _check(lib().phi_hip_integrator_create(params._h, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        try:
            if self._h:
                lib().phi_hip_integrator_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def Launch(self, check=True):
        rc = lib().phi_hip_integrator_launch(self._h)
        if check:
            _check(rc)
        return rc

    def density_matrix(self):
        """rho_0 after Launch (full space for a truncated system)."""
        n = max(self.params.get_int("num_states"), self.params.get_int("hilbert_space_size"))
        out = np.zeros(2 * n * n)
        _check(lib().phi_hip_integrator_density_matrix(self._h, _dp(out), n * n))
        return _from_colmajor(out, n)

    def info(self):
        out = np.zeros(8)
        _check(lib().phi_hip_integrator_info(self._h, _dp(out), 8))
        keys = ["steps", "accepted", "final_time", "final_dt", "seconds",
                "bicgstab_iterations", "bicgstab_residual", "num_nodes"]
        return dict(zip(keys, out.tolist()))

    def state(self):
        info = self.info()
        n = self.params.get_int("num_states")
        nodes = int(info["num_nodes"])
        out = np.zeros(2 * nodes * n * n)
        _check(lib().phi_hip_integrator_state(self._h, _dp(out), nodes * n * n))
        return out.view(complex).reshape(nodes, n, n).transpose(0, 2, 1).copy()


class Context:
    """Device-resident hierarchy (phi_hip_create): the hot path."""

    def __init__(self, params, precision="fp64", _handle=None):
        self.params = params
        if _handle is None:
            h = ctypes.c_void_p()
            _check(lib().phi_hip_create(params._h, 1 if precision == "fp32" else 0,
                                        ctypes.byref(h)))
        else:
            h = _handle
        self._h = h
        self.num_nodes = lib().phi_hip_num_nodes(h)
        self.D = lib().phi_hip_num_modes(h)
        self.N = params.get_int("num_states")

    def rk4_history(self, nsteps, dt, every, timed=False):
        """nsteps RK4 steps; rho_0 (an ensemble: the mean of the members'
        rho_0) after every step i with i % every == 0.  Returns (k, N, N);
        with timed=True also the milliseconds between two HIP events on the
        context's stream around the steps: (history, ms)."""
        k_max = (int(nsteps) + int(every) - 1) // int(every)
        out = np.zeros(2 * max(k_max, 1) * self.N * self.N)
        count = ctypes.c_int()
        cap = max(k_max, 1) * self.N * self.N
        if timed:
            ms = np.zeros(1)
            _check(lib().phi_hip_rk4_history_timed(self._h, int(nsteps), float(dt), int(every),
                                                   _dp(out), cap, ctypes.byref(count), _dp(ms)))
        else:
            _check(lib().phi_hip_rk4_history(self._h, int(nsteps), float(dt), int(every),
                                             _dp(out), cap, ctypes.byref(count)))
        k = count.value
        hist = out[:2 * k * self.N * self.N].view(complex).reshape(k, self.N, self.N).transpose(
            0, 2, 1).copy()
        return (hist, float(ms[0])) if timed else hist

    def close(self):
        if getattr(self, "_h", None):
            lib().phi_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_complex(self):
        return self.num_nodes * self.N * self.N

    def index(self):
        out = np.zeros(self.num_nodes * self.D, dtype=np.int32)
        _check(lib().phi_hip_get_index(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       out.size))
        return out.reshape(self.num_nodes, self.D)

    def _to_flat(self, x):
        return np.ascontiguousarray(np.asarray(x, dtype=complex).transpose(0, 2, 1)).view(
            np.float64).reshape(-1)

    def _from_flat(self, f):
        return f.view(complex).reshape(self.num_nodes, self.N, self.N).transpose(0, 2, 1).copy()

    def set_state(self, rho):
        f = self._to_flat(rho)
        _check(lib().phi_hip_set_state(self._h, _dp(f), self.n_complex))

    def state(self):
        out = np.zeros(2 * self.n_complex)
        _check(lib().phi_hip_get_state(self._h, _dp(out), self.n_complex))
        return self._from_flat(out)

    def rho0(self):
        out = np.zeros(2 * self.N * self.N)
        _check(lib().phi_hip_get_rho0(self._h, _dp(out), self.N * self.N))
        return _from_colmajor(out, self.N)

    def rhs(self, x):
        f = self._to_flat(x)
        out = np.zeros_like(f)
        _check(lib().phi_hip_rhs(self._h, _dp(f), _dp(out), self.n_complex))
        return self._from_flat(out)

    def rhs_noalloc(self, x):
        """rhs() through the allocation-free entry point (phi_hip_rhs_noalloc)."""
        f = self._to_flat(x)
        out = np.zeros_like(f)
        _check(lib().phi_hip_rhs_noalloc(self._h, _dp(f), _dp(out), self.n_complex))
        return self._from_flat(out)

    def rhs_nodes(self, nodes):
        """PHI's per-node Get on the current state (phi_hip_rhs_nodes):
        an array (len(nodes), N, N)."""
        ids = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32).reshape(-1))
        n = len(ids) * self.N * self.N
        out = np.zeros(2 * n)
        _check(lib().phi_hip_rhs_nodes(self._h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       len(ids), _dp(out), n))
        return out.view(complex).reshape(len(ids), self.N, self.N).transpose(0, 2, 1).copy()

    def rk4_steps(self, nsteps, dt):
        _check(lib().phi_hip_rk4_steps(self._h, int(nsteps), float(dt)))

    def rk4_steps_timed(self, nsteps, dt):
        """rk4_steps between two HIP events on the context's stream;
        returns the milliseconds between them."""
        ms = np.zeros(1)
        _check(lib().phi_hip_rk4_steps_timed(self._h, int(nsteps), float(dt), _dp(ms)))
        return float(ms[0])

    def rk4_stage_times(self, nsteps, dt):
        """Milliseconds of each stage kernel (4 values) summed over nsteps
        back-to-back launches of that stage; the state is left unchanged."""
        ms = np.zeros(4)
        _check(lib().phi_hip_rk4_stage_times(self._h, int(nsteps), float(dt), _dp(ms)))
        return ms

    def rkf45_attempt(self, coefs):
        c = np.ascontiguousarray(np.asarray(coefs, dtype=np.float64))
        assert c.size == 24
        err = ctypes.c_double()
        _check(lib().phi_hip_rkf45_attempt(self._h, _dp(c), ctypes.byref(err)))
        return err.value

    def rkf45_accept(self):
        _check(lib().phi_hip_rkf45_accept(self._h))

    def active_flags(self):
        """Adaptive truncation flags (1 = active) of every node, PHI id order."""
        out = np.zeros(self.num_nodes, dtype=np.int32)
        cnt = ctypes.c_int()
        _check(lib().phi_hip_active_flags(self._h,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          self.num_nodes, ctypes.byref(cnt)))
        return out

    def steady_state(self, tol, max_iter):
        it = ctypes.c_int()
        res = ctypes.c_double()
        rc = lib().phi_hip_steady_state(self._h, tol, max_iter, ctypes.byref(it), ctypes.byref(res))
        return rc, it.value, res.value

    def steady_state_l(self, ell, tol, max_iter):
        """BiCGSTAB(ell) on the steady-state system: (status, BiCG steps,
        residual)."""
        it = ctypes.c_int()
        res = ctypes.c_double()
        rc = lib().phi_hip_steady_state_l(self._h, int(ell), tol, max_iter, ctypes.byref(it),
                                          ctypes.byref(res))
        return rc, it.value, res.value

    def synchronize(self):
        _check(lib().phi_hip_synchronize(self._h))

    def set_launch(self, gemm=-1, tile=0, threads=0, blocks=0):
        """Stage-kernel shape of every RK4 stage: gemm 0 = element path,
        1 = matrix cores, 2 = persistent matrix-core kernels, -1 = default;
        tile = nodes per block (0 = default); threads = threads per block
        (0 = 256, or one element per thread for the persistent kernels);
        blocks = grid of the persistent kernels (0 = all resident)."""
        _check(lib().phi_hip_set_launch(self._h, int(gemm), int(tile), int(threads),
                                        int(blocks)))

    def set_stage_launch(self, stage, threads=0, blocks=0):
        """Threads and blocks of RK4 stage `stage` (0..3) of the persistent
        kernels alone; the tile stays set_launch's, which resets it."""
        _check(lib().phi_hip_set_stage_launch(self._h, int(stage), int(threads), int(blocks)))

    def stage_launch(self):
        """((threads of stages 0..3), (blocks of stages 0..3)) in the
        context's precision."""
        th = (ctypes.c_int * 4)()
        bl = (ctypes.c_int * 4)()
        _check(lib().phi_hip_stage_launch(self._h, th, bl))
        return tuple(th), tuple(bl)

    def set_graphs(self, on):
        """RK4 step blocks replayed from a captured hipGraph (True) or
        launched kernel by kernel (False)."""
        _check(lib().phi_hip_set_graphs(self._h, 1 if on else 0))

    def graphs(self):
        """True when RK4 step blocks replay a captured hipGraph."""
        on = ctypes.c_int()
        _check(lib().phi_hip_get_graphs(self._h, ctypes.byref(on)))
        return bool(on.value)

    def launch_counts(self):
        """(graph steps, eager steps) of the last RK4 call."""
        g = ctypes.c_longlong()
        e = ctypes.c_longlong()
        _check(lib().phi_hip_launch_counts(self._h, ctypes.byref(g), ctypes.byref(e)))
        return g.value, e.value

    def autotune(self, dt, steps=5):
        """Pick the fastest stage-kernel shape for this problem, precision
        and RK4 form (state unchanged).  Returns (ms_per_step, shape)."""
        ms = ctypes.c_double()
        shape = (ctypes.c_int * 4)()
        _check(lib().phi_hip_autotune(self._h, int(steps), float(dt), ctypes.byref(ms), shape))
        return ms.value, tuple(shape)

    def set_rk4_form(self, form):
        """RK4 stage order: "classic" (PHI's k0..k3 with an accumulator),
        "nested" (Taylor/Horner form) or "increment" (the nested form's
        increments as stage inputs, the default)."""
        f = {"classic": 0, "nested": 1, "increment": 2}[form]
        _check(lib().phi_hip_set_rk4_form(self._h, f))


class Ensemble(Context):
    """Many hierarchies of one bath model in one device slab
    (phi_hip_ensemble_create): member m has Hamiltonian Hs[m] and PHI depth
    depths[m] (HierarchyTruncation); all start from params' initial density
    matrix.  RK4 calls of Context apply to every member at once."""

    def __init__(self, params, Hs, depths, precision="fp64"):
        Hs = np.asarray(Hs, dtype=complex)
        m = Hs.shape[0]
        flat = np.ascontiguousarray(Hs.transpose(0, 2, 1)).view(np.float64).reshape(-1)
        d = np.ascontiguousarray(np.asarray(depths, dtype=np.int32))
        assert d.size == m
        h = ctypes.c_void_p()
        _check(lib().phi_hip_ensemble_create(params._h, m, _dp(flat),
                                             d.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                             1 if precision == "fp32" else 0, ctypes.byref(h)))
        super().__init__(params, precision, _handle=h)
        self.members = m
        counts = np.zeros(m, dtype=np.int32)
        lib().phi_hip_ensemble_members(h, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), m)
        self.counts = counts

    def member_rho0(self):
        """rho_0 of every member, (members, N, N)."""
        out = np.zeros(2 * self.members * self.N * self.N)
        _check(lib().phi_hip_ensemble_rho0(self._h, _dp(out), self.members * self.N * self.N))
        return out.view(complex).reshape(self.members, self.N, self.N).transpose(0, 2, 1).copy()

    def index(self):
        """Multi-indices of the deepest member (every member's nodes are the
        first counts[m] of them)."""
        n = int(self.counts.max())
        out = np.zeros(n * self.D, dtype=np.int32)
        _check(lib().phi_hip_get_index(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       out.size))
        return out.reshape(n, self.D)

    def closure_samples(self, shallow_depth):
        """Training rows for hierarchies of depth `shallow_depth` (PHI's
        HierarchyTruncation) from the current state: (X, T) float32."""
        N, D = self.N, self.D
        nb = int(np.sum(self.index().sum(axis=1) == shallow_depth - 1))
        cap = max(1, self.members * nb)
        X = np.zeros((cap, 2 * N * N + D), dtype=np.float32)
        T = np.zeros((cap, 2 * N * N * D), dtype=np.float32)
        rows = ctypes.c_longlong()
        _check(lib().phi_hip_closure_samples(self._h, int(shallow_depth), _fp(X), _fp(T), cap,
                                             ctypes.byref(rows)))
        return X[:rows.value], T[:rows.value]

    def attach_closure(self, net):
        """Couple a trained closure network (Mlp); resets the state."""
        _check(lib().phi_hip_closure_attach(self._h, net._h))
        self._net = net
        self.num_nodes = lib().phi_hip_num_nodes(self._h)

    def set_closure_fused(self, on):
        """Closure predictions by the fused pass (True) or unfused kernels."""
        _check(lib().phi_hip_closure_fused(self._h, 1 if on else 0))

    def closure_fallback(self, ymax=0.0):
        """Set the closure's fallback threshold (ymax > 0; outputs above it,
        or not finite, zero their boundary node's predictions) and return
        how many boundary-node predictions fell back since the attach."""
        n = ctypes.c_longlong()
        _check(lib().phi_hip_closure_fallback(self._h, float(ymax), ctypes.byref(n)))
        return n.value


def _fp(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The truncation-closure network on the device (phi_hip_mlp_*): 3-layer
    perceptron n_in -> n_hidden -> n_hidden -> n_out with GELU, bfloat16 GEMM
    operands and float32 accumulation; flat float32 parameters W1, b1, W2, b2,
    W3, b3 (weights (out, in) row-major)."""

    def __init__(self, n_in, n_hidden, n_out, max_batch=4096):
        h = ctypes.c_void_p()
        _check(lib().phi_hip_mlp_create(int(n_in), int(n_hidden), int(n_out), int(max_batch),
                                        ctypes.byref(h)))
        self._h = h
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        self.max_batch = max_batch
        self.param_count = lib().phi_hip_mlp_param_count(h)

    def close(self):
        if getattr(self, "_h", None):
            lib().phi_hip_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        f = np.ascontiguousarray(np.asarray(flat, dtype=np.float32))
        _check(lib().phi_hip_mlp_set_params(self._h, _fp(f), f.size))

    def get_params(self):
        f = np.zeros(self.param_count, dtype=np.float32)
        _check(lib().phi_hip_mlp_get_params(self._h, _fp(f), f.size))
        return f

    def forward(self, x):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        _check(lib().phi_hip_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
        t = np.ascontiguousarray(np.asarray(t, dtype=np.float32))
        g = np.zeros(self.param_count, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().phi_hip_mlp_gradients(self._h, _fp(x), _fp(t), x.shape[0],
                                           ctypes.byref(loss), _fp(g), g.size))
        return loss.value, g

    def train(self, X, T, batch=256, epochs=1, lr=1e-3, seed=0, lr_final=0.0):
        """Adam over `epochs` shuffled passes; the learning rate goes
        geometrically from lr to lr_final (0: constant).  Per-epoch losses."""
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float32))
        T = np.ascontiguousarray(np.asarray(T, dtype=np.float32))
        losses = np.zeros(epochs)
        _check(lib().phi_hip_mlp_train(self._h, _fp(X), _fp(T), X.shape[0], int(batch),
                                       int(epochs), float(lr), float(lr_final), int(seed),
                                       _dp(losses)))
        return losses


def params_from_problem(H, rho0, gamma, lambda_, temperature, lh, num_terms=1,
                        hbar=None, kb=None, dt=0.001, t_end=1.0, out_min=0.01,
                        method="rk4", output=None, precision="fp64", verbose=-2,
                        full_precision_output=None):
    """PhiParameters set field by field, as PHI's own tests do
    (hierarchy_integrator_test.cc:27-44), then FixIncompatibleCombinations."""
    p = PhiParameters()
    H = np.asarray(H, dtype=complex)
    N = H.shape[0]
    p.set_int("verbose", verbose)
    if hbar is not None:
        p.set_double("hbar", hbar)
    if kb is not None:
        p.set_double("boltzmann_constant", kb)
    p.set_int("num_states", N)
    p.set_int("hilbert_space_size", N)
    p.set_int("hilbert_space_num_elements", N * N)
    p.set_int("num_bath_couplings", len(gamma))
    p.set_int("num_matsubara_terms", num_terms)
    p.set_int("hierarchy_truncation_level", lh)
    p.set_double("temperature", temperature)
    p.set_array("gamma", gamma)
    p.set_array("lambda", lambda_)
    p.set_matrix("time_indepenent_hamiltonian", H)
    p.set_matrix("initial_density_matrix", rho0)
    p.set_double("timestep", dt)
    p.set_double("integration_time", t_end)
    p.set_double("output_minimum_timestep", out_min)
    p.set_string("output_filename", output or "")
    if full_precision_output:
        p.set_string("full_precision_output", full_precision_output)
    p.set_int("bulk_precision", 1 if precision == "fp32" else 0)
    rc = p.fix(method)
    if rc != 0:
        raise PhiError(rc, lib().phi_hip_last_error().decode())
    return p


def run_cli(args, cwd=None, timeout=None):
    """Run the `phi` program (returns CompletedProcess)."""
    import subprocess
    return subprocess.run([EXE_PATH] + list(args), cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=timeout)
