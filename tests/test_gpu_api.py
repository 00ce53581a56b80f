"""The reference's API surface beyond solve() on the GPU: readInit
(sdpa_call.cpp:946-990), writeInitSparse (1199-1222), getDimacsError
(849-853, IO::computeDimacs sdpa_io.cpp:775-813), printResult* and the
command-line front end (sdpa_exe.cpp:224-343).  Known answers: the manual's
example1 (objective -41.9 both sides, pdOPT)."""
import os
import re
import subprocess

import numpy as np
import pytest

import sdpa_amd as sa

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_read_init_solves_manual_init_point(golden_dir):
    s = sa.Solver(0).read_input(os.path.join(golden_dir, "example1.dat-s"))
    s.read_init(os.path.join(golden_dir, "example1.ini-s"))
    r = s.solve()
    assert r.phase == "pdOPT"
    assert r.primal_obj == pytest.approx(-41.9, rel=1e-6)
    assert r.dual_obj == pytest.approx(-41.9, rel=1e-6)
    # the same point through set_init gives the same trajectory
    s2 = sa.Solver(0).read_input(os.path.join(golden_dir, "example1.dat-s"))
    s2.set_init(np.array([5.9, -1.375, -1.375, 1.0]), np.array([11.0, 0.0, 0.0, 9.0]),
                -np.array([0.0, -4.0, 0.0]))
    r2 = s2.solve()
    assert r.iterations == r2.iterations
    np.testing.assert_allclose(r.trace, r2.trace, rtol=1e-12, atol=1e-14)


def test_dimacs_and_printing(golden_dir, tmp_path):
    s = sa.Solver(0).read_input(os.path.join(golden_dir, "example1.dat-s"))
    r = s.solve()
    e = s.dimacs()
    assert e[0] == 0.0
    assert 0.0 <= e[1] < 1e-7 and 0.0 <= e[3] < 1e-7      # feasibility
    assert e[2] == 0.0 and e[4] == 0.0                    # X, Z positive definite
    assert abs(e[5]) < 1e-7 and 0.0 <= e[6] < 1e-6        # gap, complementarity
    xv = tmp_path / "xvec.txt"
    s.print_result(0, xv)
    vals = [float(t) for t in re.findall(r"[-+][0-9.]+e[-+][0-9]+", open(xv).read())]
    assert open(xv).read().startswith("{") and len(vals) == 3
    xm = tmp_path / "xmat.txt"
    s.print_result(1, xm)
    txt = open(xm).read()
    assert txt.startswith("{\n{ {") and txt.rstrip().endswith("}")
    ini = tmp_path / "final.ini-s"
    s.write_init(ini, "%+.17e")
    first = open(ini).readline().split()
    assert len(first) == 3


def test_cli_example1(golden_dir, tmp_path):
    exe = os.path.join(ROOT, "sdpa_amd", "sdpa")
    out = tmp_path / "example1.result"
    print("cli: example1.dat-s -dimacs", flush=True)
    r = subprocess.run([exe, "-ds", os.path.join(golden_dir, "example1.dat-s"), "-o", str(out),
                        "-p", os.path.join(golden_dir, "param.sdpa"), "-dimacs"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    res = open(out).read()
    assert "phase.value  = pdOPT" in res
    mp = re.search(r"objValPrimal = ([-+0-9.e]+)", res)
    md = re.search(r"objValDual   = ([-+0-9.e]+)", res)
    assert float(mp.group(1)) == pytest.approx(-41.9, rel=1e-6)
    assert float(md.group(1)) == pytest.approx(-41.9, rel=1e-6)
    for tag in ("xVec = ", "xMat = ", "yMat = ", "* DIMACS_ERRORS *", "** Parameters **"):
        assert tag in res, tag
    # the iteration table: one line per iteration plus the final one
    rows = [l for l in r.stdout.splitlines() if re.match(r"^\s*\d+ \d\.\de[-+]\d\d ", l)]
    assert len(rows) >= 10


def test_schur_dense_decision_matches_oracle():
    """The device loader's record of Chordal::ordering_bMat's dense tests
    (problem_info[9]) against the oracle, including the aggregate-pattern rule
    (sdpa_chordal.cpp:404-410)."""
    from oracle import oracle as O
    from tests.test_oracle_math import _two_block_problem
    from sdpa_amd import problems as PR
    cases = [_two_block_problem([(0, 90), (30, 120)]), _two_block_problem([(0, 60), (60, 120)]),
             PR.maxcut(100, 0.1, 1), PR.lovasz(40, 0.3, 2),
             PR.multi_block(n_sdp_blocks=8, nmin=3, nmax=25, lp=20, m=60, density=0.03, seed=5)]
    for d in cases:
        g = sa.GpuContext(0).load(d)
        assert bool(g.info[9]) == O.Problem(d).schur_dense, d.name


COMPUTE_TIME_ROWS = [
    "Predictor time", "Corrector time", "Make bMat time", "Make bDia time", "Make bF1  time",
    "Make bF2  time", "Make bF3  time", "Make bPRE time", "Make rMat time", "Make gVec Mul",
    "Make gVec time", "Cholesky bMat", "Ste Pre time", "Ste Cor time", "solve", "sumDz",
    "makedX", "symmetriseDx", "makedXdZ", "xMatTime", "zMatTime", "invzMatTime",
    "xMatzMatTime", "EigxMatTime", "EigzMatTime", "EigxMatzMatTime", "updateRes", "EigTime",
    "sub_total_bMat", "Main Loop", "File Check", "File Change", "File Read", "Total"]


def test_computation_time_table(golden_dir, tmp_path):
    # printComputationTime (ComputeTime::display, sdpa_parts.cpp:78-162): the
    # reference's rows in its order and format; with timing enabled the
    # device buckets are filled and add up
    s = sa.Solver(0).read_input(os.path.join(golden_dir, "example1.dat-s"))
    s.enable_timing(True)
    r = s.solve()
    assert r.phase == "pdOPT"
    out = tmp_path / "times.txt"
    s.print_result(3, out)
    lines = [l for l in open(out).read().splitlines() if "=" in l]
    got = {}
    for want, line in zip(COMPUTE_TIME_ROWS, lines):
        name, vals = line.split("=")
        assert name.strip() == want.strip(), (name, want)
        t, ratio = (float(v) for v in vals.split(","))
        got[want] = t
    assert len(lines) == len(COMPUTE_TIME_ROWS)
    for k in ("Make bMat time", "Cholesky bMat", "solve", "EigxMatzMatTime", "Main Loop"):
        assert got[k] > 0.0, k
    assert got["Make bF1  time"] + got["Make bF3  time"] == pytest.approx(got["Make bMat time"],
                                                                          rel=1e-5, abs=2e-6)
    assert got["Predictor time"] + got["Corrector time"] <= got["Main Loop"]
    assert got["sub_total_bMat"] == pytest.approx(got["Main Loop"] - got["Make bMat time"],
                                                  abs=2e-6)
