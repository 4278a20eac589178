"""The drop-in command line (src/main.cu:255-306 contract)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import sis

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CLI = os.path.join(ROOT, "sparse-ising-solver_amd", "build", "sis")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")


def run(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([CLI] + [str(a) for a in args], capture_output=True, text=True, env=e,
                          timeout=600)


def test_usage_errors(tmp_path):
    r = run(["1", "x.mtx"])
    assert r.returncode == 255 and r.stdout.startswith("usage: ./program <default?>")
    r = run(["0", "x.mtx", tmp_path / "o.csv", "maxcut", "0", "1"])
    assert r.returncode == 255 and "<max_steps> <agents>" in r.stdout


def test_unknown_extension_does_nothing(tmp_path):
    out = tmp_path / "o.csv"
    r = run(["1", "graph.txt", out])
    assert r.returncode == 0 and r.stdout == "" and not out.exists() or out.read_text() == ""


def test_bad_input_file_fails_cleanly(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer\n2 2 1\n1 2 1\n")
    r = run(["0", p, tmp_path / "o.csv", "maxcut", "0", "1", "0", "0.5", "100", "64"])
    assert r.returncode == 1 and "neither symmetric nor general" in r.stderr


def _fields(line):
    return [f.strip() for f in line.rstrip("\n").split(",")]


@pytest.mark.gpu
@pytest.mark.parametrize("fused,knobs", [
    (1, {}), (0, {}),
    # packed form, whole-run graph, tuned tile and row split: same results
    (1, {"SIS_FORM": "packed", "SIS_GRAPH": "2", "SIS_SPLIT": "-1", "SIS_TUNE": "1"}),
    (1, {"SIS_SPLIT": "8"}),
])
def test_maxcut_line_matches_oracle(tmp_path, fused, knobs):
    J = sis.Csr.gen_er(300, 0.05, sis.W_PM1, 3)
    mtx = tmp_path / "g.mtx"
    J.write_mtx(mtx)
    out = tmp_path / "out.csv"
    A, T, dt, seed = 128, 1500, 0.5, 99
    r = run(["0", mtx, out, "maxcut", "1", fused, "1", dt, T, A], env={"SIS_SEED": str(seed), **knobs})
    assert r.returncode == 0, r.stderr
    f = _fields(out.read_text())
    assert f[6:] == [str(mtx), "300", str(J.nnz), str(T), "0.5", str(A)]
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    X, Y = orc.init_state(seed, orc.MAXCUT, J.n, A)
    prob = orc.MAXCUT if fused else orc.MAXCUT_UNFUSED
    res = orc.solve(pb, prob, orc.DSB, X, Y, T, 1, P=100, dt=dt, xi0=J.xi0())
    st = orc.cut_stats(orc.cut(pb, res["cur"]))
    assert [float(v) for v in f[1:6]] == list(st)
    assert "max cut from random spin assignment" in r.stdout
    assert r.stdout.startswith(f"file: {mtx}, time: ")


@pytest.mark.gpu
@pytest.mark.parametrize("form,w", [("q12", 12), ("q10", 10), ("q8", 8), ("q8s", 8)])
def test_maxcut_line_narrow_form_matches_oracle(tmp_path, form, w):
    """SIS_FORM=q12 / q10 / q8 / q8s (12-, 10- or 8-bit X/Y codes; q8s rounds
    Y stochastically): the line's cut statistics
    equal the oracle's narrow-state solve from the same seed."""
    J = sis.Csr.gen_er(300, 0.05, sis.W_PM1, 3)
    mtx = tmp_path / "g.mtx"
    J.write_mtx(mtx)
    out = tmp_path / "out.csv"
    A, T, dt, seed = 128, 1500, 0.5, 99
    r = run(["0", mtx, out, "maxcut", "1", "1", "1", dt, T, A],
            env={"SIS_SEED": str(seed), "SIS_FORM": form})
    assert r.returncode == 0, r.stderr
    f = _fields(out.read_text())
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    X, Y = orc.init_state(seed, orc.MAXCUT, J.n, A)
    res = orc.solve_qn(pb, X, Y, T, 1, P=100, K=25, dt=dt, xi0=J.xi0(), w=w, sr=form == "q8s")
    assert [float(v) for v in f[1:6]] == list(orc.cut_stats(orc.cut(pb, res["cur"])))


@pytest.mark.gpu
@pytest.mark.parametrize("knobs", [{}, {"SIS_STRUCTURED": "1", "SIS_DYNAMICS": "dsb"},
                                   {"SIS_TSP_XI0_RULE": "scaled"}])
def test_tsp_line_matches_oracle(tmp_path, knobs):
    """bSB through the QUBO's CSR (the reference's TSP); dSB through the
    structured product, whose trajectory is bit-identical to the CSR's; and
    the coupling 0.5 / (sigma sqrt N) in place of 0.002."""
    src = os.path.join(INPUTS, "euc8.tsp")
    out = tmp_path / "out.csv"
    A, T, dt, seed = 256, 2000, 0.1, 5
    r = run(["0", src, out, "tsp", "0", "1", "0", dt, T, A], env={"SIS_SEED": str(seed), **knobs})
    assert r.returncode == 0, r.stderr
    f = _fields(out.read_text())
    t = sis.Tsp.read(src)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, field=h, ncities=8)
    X, Y = orc.init_state(seed, orc.TSP, 64, A)
    dyn = orc.DSB if knobs.get("SIS_DYNAMICS") == "dsb" else orc.BSB
    xi0 = t.xi0() if knobs.get("SIS_TSP_XI0_RULE") == "scaled" else 0.002
    res = orc.solve(pb, orc.TSP, dyn, X, Y, T, 0, P=T // 10, K=50, dt=dt, xi0=xi0,
                    coupling=1.005)
    _, _, d, _ = t.arrays()
    lens = np.array([orc.tsp_tour_length(8, d, res["fin"], a)[0] for a in range(A) if res["valid"][a] > 0])
    assert int(f[5]) == len(lens)
    if len(lens):
        assert float(f[1]) == lens.min() and float(f[2]) == lens.max()
    assert r.stdout.startswith("8\nEUC_2D\n\n")
    assert "traveling salesman distance from random spin assignment" in r.stdout


def test_tsp_xi0_rule_rejects_unknown_value(tmp_path):
    src = os.path.join(INPUTS, "euc8.tsp")
    r = run(["0", src, tmp_path / "o.csv", "tsp", "0", "1", "0", "0.1", "100", "64"],
            env={"SIS_TSP_XI0_RULE": "optimal"})
    assert r.returncode == 1 and "SIS_TSP_XI0_RULE must be fixed or scaled" in r.stderr
