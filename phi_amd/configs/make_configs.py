"""Generates the `.prm` inputs of the benchmark configurations.

Random draws use numpy's PCG64 with fixed seeds and are written at %.17g, so
the files are exact and both this integrator and PHI read identical values.
Units follow PHI: energies in cm^-1, gamma in ps^-1, times in ps; PHI's
HierarchyTruncation = depth + 1 and MatsubaraTerms = K (stored as K + 1).

  c1  dimer, H = [[100,100],[100,0]], lambda 35, gamma 20, T 300 K,
      MatsubaraTerms=1, depth 8 (Lh 9): 495 nodes; RK4 dt 1 fs to 1 ps
  c2  7-site exciton, eps ~ U[12200,12600], J ~ N(0,40) clipped |J|<=120,
      seed 7001; lambda 35, gamma 20, T 77 K, K 0, depth 8 (Lh 9): 6435 nodes
  c3  24-site ring, NN couplings alternating 300/230, NNN -30,
      eps ~ N(12500,100) seed 7003; lambda 50, gamma 10, T 300 K, K 0,
      depth 4 (Lh 5): 20475 nodes; RKF45 tol 1e-8 (and RK4 dt 1 fs)
  c4  ensemble: c2 with eps += N(0,100) (seed 7004+r), depth U{4..8}
  c5  closure set: random c2-type Hamiltonians (seeds 50000+i)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def g(x):
    return "%.17g" % x


def prm_text(H, lam, gam, T, lh, kt, rho0=None, dt=0.001, t=1.0, out_min=0.01, extra=()):
    N = H.shape[0]
    if rho0 is None:
        rho0 = np.zeros((N, N))
        rho0[0, 0] = 1.0
    lines = ["NumStates=%d" % N, "HierarchyTruncation=%d" % lh, "MatsubaraTerms=%d" % kt,
             "Temperature=%s" % g(T), "Timestep=%s" % g(dt), "Time=%s" % g(t),
             "OutputMinimumTimestep=%s" % g(out_min), "Hamiltonian:"]
    lines += [",".join(g(v) for v in row) for row in H]
    lines += ["Lambda:", ",".join(g(lam) for _ in range(N)),
              "Gamma:", ",".join(g(gam) for _ in range(N)),
              "InitialDensityMatrix:"]
    lines += [",".join(g(v) for v in row) for row in np.real(rho0)]
    lines += list(extra)
    return "\n".join(lines) + "\n"


def c2_hamiltonian(rng):
    N = 7
    eps = rng.uniform(12200.0, 12600.0, size=N)
    H = np.diag(eps)
    J = np.clip(rng.normal(0.0, 40.0, size=(N, N)), -120.0, 120.0)
    iu = np.triu_indices(N, 1)
    H[iu] = J[iu]
    H[(iu[1], iu[0])] = J[iu]
    return H


def c1():
    H = np.array([[100.0, 100.0], [100.0, 0.0]])
    return prm_text(H, 35.0, 20.0, 300.0, 9, 1, extra=["OutputFile=c1.out"])


def c2():
    H = c2_hamiltonian(np.random.Generator(np.random.PCG64(7001)))
    return prm_text(H, 35.0, 20.0, 77.0, 9, 0,
                    extra=["OutputFile=c2.out", "bicgstab_tolerance=1e-10"])


def c3_hamiltonian():
    rng = np.random.Generator(np.random.PCG64(7003))
    N = 24
    H = np.diag(rng.normal(12500.0, 100.0, size=N))
    for i in range(N):
        j = (i + 1) % N
        H[i, j] = H[j, i] = 300.0 if i % 2 == 0 else 230.0
        k = (i + 2) % N
        H[i, k] = H[k, i] = -30.0
    return H


def c3():
    return prm_text(c3_hamiltonian(), 50.0, 10.0, 300.0, 5, 0,
                    extra=["OutputFile=c3.out", "RKF45tolerance=1e-08"])


def c4_member(r):
    """Realisation r of the static-disorder ensemble: (H, Lh)."""
    base = c2_hamiltonian(np.random.Generator(np.random.PCG64(7001)))
    rng = np.random.Generator(np.random.PCG64(7004 + r))
    H = base + np.diag(rng.normal(0.0, 100.0, size=7))
    lh = int(rng.integers(5, 10))   # depth 4..8
    return H, lh


def c5_hamiltonian(i):
    return c2_hamiltonian(np.random.Generator(np.random.PCG64(50000 + i)))


def write_all(outdir=HERE):
    for name, fn in (("c1", c1), ("c2", c2), ("c3", c3)):
        with open(os.path.join(outdir, name + ".prm"), "w") as f:
            f.write(fn())


if __name__ == "__main__":
    write_all(sys.argv[1] if len(sys.argv) > 1 else HERE)
