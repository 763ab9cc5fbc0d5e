"""The reference's own run: example/one_qubit.prm (annealing qubit,
DiagonalBath with CorrelatedCouplingTerms, HierarchyTruncation=3, RKF45 to
T = 1000) and the output file PHI wrote for it, example/one_qubit_100.txt.

CPU: the oracle's annealing RKF45 loop reproduces PHI's attempted and
accepted step counts (one_qubit_100.txt:444-445) and the final density
matrix at PHI's printed digits (:447).
GPU: the `phi` program on the same file writes an output file that equals
PHI's: the header byte for byte, the dynamics lines (t = 0, the first
accepted step, the final line) at their printed precision, the step-count
lines exactly.  Only wall-clock lines (#Info ..., #Total integration time)
are skipped.
"""

import os
import re
import shutil

import numpy as np
import pytest

import phi_amd as P
from oracle import heom_oracle as HO
from oracle import prm_oracle as PO
from oracle import td_oracle as TD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "tests", "golden", "reference_data")
EXAMPLE = os.path.join(REF, "one_qubit.prm")
OUTPUT = os.path.join(REF, "one_qubit_100.txt")

_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
_CPLX = re.compile(r"\((%s),(%s)\)" % (_NUM, _NUM))


def dynamics_line(line):
    """(time, [complex, ...]) of a PHI dynamics line `t (re,im) (re,im) ...`."""
    t = float(line.split()[0])
    vals = [complex(float(a), float(b)) for a, b in _CPLX.findall(line)]
    return t, vals


def printed_tolerance(text):
    """Half a unit in the last printed digit of a number as C++ ostream
    prints it (default or setprecision(p) general format); 0 -> 1e-12,
    the level below which a printed 0 cannot be told from rounding noise."""
    v = float(text)
    if v == 0:
        return 1e-12
    mant = text.lower().split("e")[0].lstrip("+-")
    decimals = len(mant.split(".")[1]) if "." in mant else 0
    exp = int(text.lower().split("e")[1]) if "e" in text.lower() else 0
    return 0.5 * 10.0 ** (exp - decimals)


def assert_line_matches(got, want):
    """A dynamics line equal to PHI's at the precision PHI printed."""
    tg, vg = dynamics_line(got)
    tw, vw = dynamics_line(want)
    assert abs(tg - tw) <= printed_tolerance(want.split()[0]), (got, want)
    assert len(vg) == len(vw), (got, want)
    for (re_w, im_w), zg in zip(_CPLX.findall(want), vg):
        assert abs(zg.real - float(re_w)) <= printed_tolerance(re_w), (got, want)
        assert abs(zg.imag - float(im_w)) <= printed_tolerance(im_w), (got, want)


def example_oracle():
    op = PO.read_command_line(EXAMPLE, "rkf45")
    ham = TD.Hamiltonian(op.num_states, H_initial=op.initial_annealing_hamiltonian,
                         H_final=op.final_annealing_hamiltonian,
                         f_initial=op.initial_hamiltonian_field,
                         f_final=op.final_hamiltonian_field, t_total=op.integration_time,
                         min_change=op.minimum_field_change)
    prob = HO.Problem(H=ham.H, rho0=op.initial_density_matrix, gamma=op.gamma,
                      lambda_=op.lambda_, temperature=op.temperature,
                      lh=op.hierarchy_truncation_level, num_terms=op.num_matsubara_terms,
                      hbar=op.hbar, kb=op.boltzmann_constant, bath_model="diagonal",
                      bath_ops=[op.bath_coupling_op])
    return op, ham, HO.Oracle(prob, threads=1)


def reference_lines():
    return open(OUTPUT).read().splitlines()


def test_oracle_reproduces_reference_example_run():
    """13046 attempted / 12412 accepted steps and the final line of
    one_qubit_100.txt:444-447 at PHI's printed digits."""
    op, ham, orc = example_oracle()
    attempts, accepted, t = TD.rkf45(orc, ham, 0.0, op.integration_time, op.timestep,
                                     op.rkf45_tolerance, op.rkf45_minimum_timestep)
    lines = reference_lines()
    assert lines[443] == "#Total number of integration steps = %d" % attempts
    assert lines[444] == "#Total number of accepted integration steps = %d" % accepted
    assert (attempts, accepted) == (13046, 12412)
    r = orc.state()[0]
    got = "%g " % t + " ".join("(%.17g,%.17g)" % (z.real, z.imag) for z in r.T.reshape(-1))
    assert_line_matches(got, lines[446])


def test_printed_tolerance_rules():
    assert printed_tolerance("0.9984") == pytest.approx(5e-5)
    assert printed_tolerance("-0.0001018") == pytest.approx(5e-8)
    assert printed_tolerance("5.408e-12") == pytest.approx(5e-16)
    assert printed_tolerance("1000") == pytest.approx(0.5)
    assert printed_tolerance("0") == 1e-12
    # a planted error of one printed digit fails
    with pytest.raises(AssertionError):
        assert_line_matches("1000 (0.9985,0)", "1000 (0.9984,0)")


def _skip_wall_clock(lines):
    return [l for l in lines
            if not l.startswith("#Info") and not l.startswith("#Total integration time")]


@pytest.mark.gpu
def test_phi_program_reproduces_reference_example_output(tmp_path):
    """`phi one_qubit.prm rkf45` on the GPU writes PHI's output file."""
    print("case: phi one_qubit.prm rkf45", flush=True)
    shutil.copy(EXAMPLE, tmp_path / "one_qubit.prm")
    r = P.run_cli(["one_qubit.prm", "rkf45"], cwd=str(tmp_path), timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _skip_wall_clock((tmp_path / "one_qubit_100.txt").read_text().splitlines())
    want = _skip_wall_clock(reference_lines())
    # header (48 lines) byte for byte
    assert got[:48] == want[:48]
    assert len(got) == len(want), "\n".join(got[48:])
    for g, w in zip(got[48:], want[48:]):
        if w.startswith("#"):
            assert g == w
        else:
            assert_line_matches(g, w)
