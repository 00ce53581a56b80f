"""Host side of the drop-in boundary, no GPU needed: the .dat-s writer
against the reader (SDPA::writeInputSparse, sdpa_call.cpp:1099-1147, and
IO::read, sdpa_io.cpp:253-606) and the command-line front end's usage path
(sdpa_exe.cpp:30-83)."""
import os
import subprocess

import numpy as np
import pytest

import sdpa_amd as sa
from sdpa_amd import dats

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _records(d):
    rec = sorted(zip(d.k.tolist(), d.l.tolist(), d.i.tolist(), d.j.tolist(), d.v.tolist()))
    return rec


@pytest.mark.parametrize("name", ["example1.dat-s", "example1.dat", "example2.dat"])
def test_write_input_round_trip(golden_dir, tmp_path, name):
    src = os.path.join(golden_dir, name)
    s = sa.Solver(0).read_input(src)
    out = tmp_path / "round.dat-s"
    s.write_input(out, "%+.17e")
    a, b = dats.read(src), dats.read(str(out))
    assert a.m == b.m
    assert list(a.block_struct) == list(b.block_struct)
    np.testing.assert_array_equal(a.c, b.c)
    ra = [r for r in _records(a) if r[4] != 0.0]
    assert ra == _records(b)


def test_write_input_noprint(golden_dir, tmp_path):
    s = sa.Solver(0).read_input(os.path.join(golden_dir, "example1.dat-s"))
    out = tmp_path / "np.dat-s"
    s.write_input(out, "NOPRINT")
    lines = open(out).read().split("\n")
    assert lines[:4] == ["3", "1", "2", "NOPRINT"]


def test_cli_usage_without_arguments():
    exe = os.path.join(ROOT, "sdpa_amd", "sdpa")
    if not os.path.exists(exe):
        pytest.skip("command-line front end not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Please assign data file and output file" in r.stdout
    assert "-dd : data dense" in r.stdout
