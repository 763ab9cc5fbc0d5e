"""Restart files on PHI's wall-clock cadence (RestartOutputTimeSecs,
hierarchy_integrator.cc:2512-2526, 2881-2895; restart_helper.cc:30-84).

  * a long RK4 run of the `phi` program dumps its state while it runs; the
    run is terminated after a dump, resumed from the file, and the resumed
    run equals an uninterrupted one to the same end time;
  * a restart file in PHI's own format (default ostream precision, 6
    significant digits) resumes to the oracle's run from the same rounded
    state.
"""

import os
import shutil
import subprocess
import time

import numpy as np
import pytest

import phi_amd as P
from oracle import heom_oracle as HO
from oracle import prm_oracle as PO

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C1 = os.path.join(ROOT, "phi_amd", "configs", "c1.prm")


def c1_text(extra, t_end):
    lines = [l for l in open(C1).read().splitlines()
             if not l.startswith("Time=") and not l.startswith("OutputFile=")]
    return "\n".join(lines + ["Time=%.17g" % t_end] + list(extra)) + "\n"


def read_restart(path):
    with open(path) as f:
        t = float(f.readline())
        dt = float(f.readline())
        rows = [[complex(*map(float, v.strip("()").split(","))) for v in line.split()]
                for line in f]
    return t, dt, np.array(rows)


def test_rk4_dump_during_run_kill_and_resume(tmp_path):
    """The `phi` program writes restart files while it runs; stopped after
    a dump and resumed from it, it reaches the uninterrupted run's state."""
    (tmp_path / "long.prm").write_text(
        c1_text(["OutputFile=long.out", "RestartFile=dump.txt", "RestartOutputTimeSecs=0.5"],
                1e6))
    print("case: phi long.prm rk4 (terminated after a restart dump)", flush=True)
    proc = subprocess.Popen([P.EXE_PATH, "long.prm", "rk4"], cwd=str(tmp_path),
                            stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    dump = tmp_path / "dump.txt"
    t_k = None
    try:
        deadline = time.time() + 90
        while time.time() < deadline and proc.poll() is None:
            if dump.exists():
                t_k = float(dump.read_text().split("\n", 1)[0])
                if t_k > 0:
                    break
            time.sleep(0.05)
    finally:
        proc.terminate()
        proc.wait(timeout=60)
    assert t_k is not None and t_k > 0, "no restart file was written while the run lasted"
    shutil.copy(dump, tmp_path / "resume.txt")
    t, dt, rows = read_restart(tmp_path / "resume.txt")
    assert dt == 0.001 and rows.shape == (495, 4)
    print("dumped at t = %.17g" % t, flush=True)
    t_end = t + 0.05
    # PHI opens RestartInput relative to the working directory; this process
    # runs elsewhere, so the path is absolute
    (tmp_path / "resume.prm").write_text(
        c1_text(["RestartInput=%s" % (tmp_path / "resume.txt")], t_end))
    a = P.HierarchyIntegrator(P.PhiParameters.read(str(tmp_path / "resume.prm"), "rk4"))
    a.Launch()
    (tmp_path / "whole.prm").write_text(c1_text([], t_end))
    b = P.HierarchyIntegrator(P.PhiParameters.read(str(tmp_path / "whole.prm"), "rk4"))
    b.Launch()
    assert a.info()["final_time"] == b.info()["final_time"]
    assert a.info()["steps"] + round(t / dt) == pytest.approx(b.info()["steps"], abs=1)
    d = np.abs(a.state() - b.state()).max()
    print("resumed vs uninterrupted max|d| = %.3e" % d)
    assert d <= 1e-12


def test_resume_from_phi_format_restart_file(tmp_path):
    """A restart file as PHI writes it (6 significant digits) resumes to the
    oracle's trajectory from the same rounded state and start time."""
    op = PO.read_command_line(C1, "rk4")
    orc = HO.Oracle(HO.Problem.from_prm(op), threads=4)
    orc.rk4(0.2, op.timestep, op.output_minimum_timestep)
    state = orc.state()
    # PHI: out << time << "\n" << timestep << "\n", then every node's N^2
    # column-major elements as "(re,im)" at the stream's default precision
    lines = ["%g" % 0.2, "%g" % op.timestep]
    for node in state:
        lines.append(" ".join("(%g,%g)" % (z.real, z.imag) for z in node.T.reshape(-1)))
    (tmp_path / "phi_restart.txt").write_text("\n".join(lines) + "\n")
    t, dt, rows = read_restart(tmp_path / "phi_restart.txt")
    rounded = rows.reshape(-1, 2, 2).transpose(0, 2, 1)
    assert np.abs(rounded - state).max() > 0  # the file did lose digits
    (tmp_path / "r.prm").write_text(
        c1_text(["RestartInput=%s" % (tmp_path / "phi_restart.txt")], 0.3))
    it = P.HierarchyIntegrator(P.PhiParameters.read(str(tmp_path / "r.prm"), "rk4"))
    it.Launch()
    orc.set_state(rounded)
    steps, _, _ = orc.rk4(0.3, dt, op.output_minimum_timestep, t_start=t)
    assert it.info()["steps"] == steps
    assert np.abs(it.state() - orc.state()).max() <= 1e-10
