"""The C++ driver (bin/nflow_main) over the C ABI.  Without options it runs the
reference's shipped case (main.cpp:56-85: 10^3 cavity, MRT + LES with Cs = 0,
Re 1000, lid 0.1, 199 steps after initialisation, 3 x 3 x 3 tiles) and prints
the arF slice main.cpp:321-348 prints, checked here against the FP64 oracle;
its outputs (iteration.dat, result.csv, section.csv, VTU) and restart files."""
import os
import subprocess

import numpy as np
import pytest

import nflow
import oracle

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(nflow.__file__)), "bin", "nflow_main")
RULE = "---------------------------------"


def run(args, cwd):
    assert os.path.exists(BIN), "driver not built: run __graft_entry__.build()"
    print("nflow_main", " ".join(args), flush=True)
    p = subprocess.run([BIN] + list(args), cwd=cwd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def printed_slice(out):
    """Rows y = ny-1 .. 0 of direction 9 on the z = 2 plane, as printed."""
    lines = out.splitlines()
    i = lines.index(RULE)
    rows = []
    for line in lines[i + 1:]:
        if line.startswith(RULE):
            break
        if line.strip():
            rows.append([float(v) for v in line.split()])
    return np.array(rows)


def test_shipped_case_prints_the_oracle_slice(tmp_path):
    got = printed_slice(run([], tmp_path))
    p = oracle.default_params()  # the same main.cpp defaults
    L = oracle.Lattice(p, oracle.decomposition(10, 10, 10, 3, 3, 3))
    L.initialize()
    L.run(199)
    a = L.read_arF()
    want = np.array([[a[y + 1, 3, x + 1, 9] for x in range(10)] for y in range(9, -1, -1)])
    assert got.shape == (10, 10)
    # printed with 6 decimals; FP32 against FP64 over 199 steps
    assert np.abs(got - want).max() < 3e-6, np.abs(got - want).max()


def test_outputs(tmp_path):
    out = run(["--n", "12", "--steps", "60", "--residual-every", "20", "--csv", "--vtu", "f.vtu", "--no-print"],
              tmp_path)
    assert "ITER=60" in out
    it = (tmp_path / "iteration.dat").read_text().splitlines()
    assert it[0] == "iter, error" and [int(l.split()[0]) for l in it[1:]] == [20, 40, 60]
    assert all(float(l.split()[1]) > 0 for l in it[1:])
    res = (tmp_path / "result.csv").read_text().splitlines()
    assert res[0] == "x,y,z,rho,u,v,w" and len(res) == 1 + 12 ** 3
    sec = (tmp_path / "section.csv").read_text().splitlines()
    assert len(sec) == 1 + 12
    from test_io_cpu import parse_vtu

    npts, _, arrays = parse_vtu(tmp_path / "f.vtu")
    assert npts == 12 ** 3
    rho = np.array([float(l.split(",")[3]) for l in res[1:]])
    assert np.allclose(arrays["Density"], rho, atol=1e-6)


def test_restart_continues_the_run(tmp_path):
    full = printed_slice(run(["--n", "12", "--steps", "60"], tmp_path))
    run(["--n", "12", "--steps", "30", "--checkpoint", "c.ckp", "--no-print"], tmp_path)
    cont = printed_slice(run(["--n", "12", "--steps", "30", "--restart", "c.ckp"], tmp_path))
    assert np.array_equal(full, cont)


def test_pipe_case_writes_solid_flags(tmp_path):
    run(["--pipe", "4", "5", "5", "--nx", "20", "--ny", "12", "--nz", "12", "--steps", "40", "--vtu", "p.vtu",
         "--no-print"], tmp_path)
    from test_io_cpu import parse_vtu

    npts, _, arrays = parse_vtu(tmp_path / "p.vtu")
    y, z = np.meshgrid(np.arange(12), np.arange(12), indexing="ij")
    solid = ((y - 5) ** 2 + (z - 5) ** 2 > 16)[:, :, None].repeat(20, axis=2)  # [y][z][x]
    assert np.array_equal(arrays["Struct"].astype(bool), solid.ravel())
    assert not arrays["Density"][solid.ravel()].any()
    assert np.isfinite(arrays["Density"]).all()


def test_tuning_options_keep_the_results(tmp_path):
    """--block and --temporal change only the launches: the printed slice and
    the result file are the same bit for bit (a 256-wide cavity, where two
    steps per launch apply)."""
    base = ["--nx", "256", "--ny", "20", "--nz", "16", "--les", "0.17", "--tau", "0.55", "--steps", "9", "--csv"]
    outs = []
    for opts in (["--temporal", "0"], ["--temporal", "4", "--block", "128"], ["--temporal", "1"]):
        d = tmp_path / opts[1]
        d.mkdir()
        outs.append((printed_slice(run(base + opts, d)), (d / "result.csv").read_text()))
    for got in outs[1:]:
        assert np.array_equal(got[0], outs[0][0]) and got[1] == outs[0][1]
