"""On-disk formats (no GPU): restart files in turn with newest-complete selection, and the
Tecplot point file of a structured block."""
import os

import numpy as np

import cfdx


def flow_of(n, seed=1):
    rng = np.random.default_rng(seed)
    f = np.zeros(n, dtype=cfdx.FLOW_DT)
    for k in ("u", "v", "w", "rho", "p", "t", "e", "mu"):
        f[k] = rng.normal(size=n) * 10.0 ** rng.integers(-3, 3)
    return f


def test_restart_round_trip_exact(tmp_path):
    f = flow_of(37)
    cfdx.write_restart(str(tmp_path), 1, f, 100)
    g, it = cfdx.read_restart(str(tmp_path), 1, 37)
    assert it == 100
    for k in ("u", "v", "w", "rho", "p", "t", "e", "mu"):
        assert np.array_equal(g[k], f[k])


def test_restart_reference_layout(tmp_path):
    f = flow_of(3)
    cfdx.write_restart(str(tmp_path), 2, f, 7, digits=6)
    lines = open(os.path.join(tmp_path, "restart_file_2.neu")).read().splitlines()
    assert len(lines) == 3
    cols = lines[1].split("\t")
    assert len(cols) == 9 and cols[8] == "7"
    # columns u v w rho p t e mu, six significant digits as an ostream writes them
    want = [f[k][1] for k in ("u", "v", "w", "rho", "p", "t", "e", "mu")]
    assert cols[:8] == [f"{x:.6g}" for x in want]


def test_latest_complete_file(tmp_path):
    d = str(tmp_path)
    n = 20
    assert cfdx.restart_latest(d, n) == (0, -1)
    cfdx.write_restart(d, 1, flow_of(n, 1), 100)
    assert cfdx.restart_latest(d, n) == (1, 100)
    cfdx.write_restart(d, 2, flow_of(n, 2), 200)
    assert cfdx.restart_latest(d, n) == (2, 200)
    # an interrupted rewrite of slot 1: new iteration on the first lines only
    p1 = os.path.join(d, "restart_file_1.neu")
    cfdx.write_restart(d, 1, flow_of(n, 3), 300)
    lines = open(p1).read().splitlines()
    old = [ln.rsplit("\t", 1)[0] + "\t100" for ln in lines]
    open(p1, "w").write("\n".join(lines[:7] + old[7:]) + "\n")
    assert cfdx.restart_latest(d, n) == (2, 200)
    # a truncated slot 2 leaves nothing complete
    p2 = os.path.join(d, "restart_file_2.neu")
    open(p2, "w").write("\n".join(open(p2).read().splitlines()[:5]) + "\n")
    assert cfdx.restart_latest(d, n) == (0, -1)
    try:
        cfdx.read_restart(d, 2, n)
        assert False, "incomplete file accepted"
    except cfdx.CfdxError:
        pass


def test_tecplot_file(tmp_path):
    ni, nj, nk = 4, 3, 2
    case = cfdx.periodic_case(ni, nj, nk, grid="box", init="tgv", length=(1.0, 1.0, 1.0), mach=0.2)
    path = str(tmp_path / "nodefile_0.dat")
    cfdx.write_tecplot(path, ni, nj, nk, case["xyz"], case["FLOW"][1:], solution_time=0.5)
    lines = open(path).read().split("\n")
    assert lines[0] == "TITLE = Node file"
    assert lines[1] == "VARIABLES = X, Y, Z, U, V, W, RHO, P, T, e, mu, loc, Proc, DUCROS,"
    assert lines[2].startswith("ZONE NODES= 24 ELEMENTS= 6 DATAPACKING=POINT, ZONETYPE=FEBRICK, SOLUTIONTIME= 0.5")
    nodes = [ln.split("\t") for ln in lines[3:3 + 24]]
    assert all(len(c) == 14 for c in nodes)
    x = np.array([[float(c[0]), float(c[1]), float(c[2])] for c in nodes])
    assert np.array_equal(x, case["xyz"])
    assert float(nodes[5][6]) == case["FLOW"]["rho"][6]
    cells = [ln for ln in lines[3 + 24:] if ln.strip()]
    assert len(cells) == 6
    assert cells[0].split("\t") == ["1", "2", "6", "5", "13", "14", "18", "17"]
