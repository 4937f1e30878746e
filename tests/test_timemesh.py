"""Time-mesh preparation before the module (host only, CPU): TimeDirectSolver and CorrectTimeMesh.

The reference ships no time-mesh fixtures, so the oracle is pinned by hand-derived known
answers (a grid with three step intervals and a truncated last count; a decade correction; the
single-point decade whose mean step is 0/0 in the reference and whose points are dropped), and
the C ABI and the drop-in executables must write the same bytes as the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import refio, timemesh

PKG_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "a-e_formulations_3d_timedomain_em_horizontalgroundedwiresource_amd")
FILES_TDS = ["infite.0", "currentfunction", "deltafunction", "!infite.0", "!currentfunction", "!deltafunction"]
FILES_CTM = ["infite.0", "currentfunction", "deltafunction", "!!infite.0", "!!currentfunction", "!!deltafunction"]


def test_gen_reg_mesh_known_answer():
    # steps 1,1,1 then 7 > 3*1 closes [0, 11] with 11 steps of 1; 10 > 3*3 closes [11, 23] with
    # 4 steps of 3; the last point closes int((100 - 23)/9 + 1) = 9 steps of 9, ending at 104
    g = timemesh.gen_reg_mesh([0, 1, 2, 3, 10, 20, 30, 100], 3.0)
    want = list(range(12)) + [14, 17, 20, 23] + [23 + 9 * k for k in range(1, 10)]
    assert g == [float(v) for v in want]


def test_gen_reg_mesh_single_interval():
    # no step exceeds 2x the first: one interval closed at the last point with int(8/1 + 1) = 9
    assert timemesh.gen_reg_mesh([0, 1, 2, 4, 8], 2.0) == [float(v) for v in range(10)]


def test_correct_mesh_known_answer():
    # decades [0..4] (h=1), [6, 8] (h=2), [18, 28] (h=10); the second is too short (2*2 <= 5)
    # and continues h = (4 - 0)/4 = 1 up to 8; the third is kept (2*2 > 2)
    g = timemesh.correct_mesh([0, 1, 2, 3, 4, 6, 8, 18, 28], 2.0)
    assert g == [0.0, 1, 2, 3, 4, 5, 6, 7, 8, 18, 28]


def test_correct_mesh_single_point_decades():
    # decades [0,1,2], [4], [7]: [4] continues step 1 (3, 4); [7] follows a one-point decade,
    # whose mean step is 0/0 in the reference, so nothing is added and 7 is dropped
    assert timemesh.correct_mesh([0, 1, 2, 4, 7], 2.0) == [0.0, 1, 2, 3, 4]


def test_interpolate_known_answers():
    vt, vv = [0.0, 1.0, 3.0], [2.0, 4.0, 0.0]
    assert timemesh.interpolate(-1.0, vt, vv) == 2.0      # held before the table
    assert timemesh.interpolate(5.0, vt, vv) == 0.0       # and after it
    assert timemesh.interpolate(0.5, vt, vv) == 3.0
    assert timemesh.interpolate(2.0, vt, vv) == 2.0
    assert timemesh.interpolate(1.0, vt, vv) == 4.0       # a node: the interval below it


def test_written_grid_reads_back(tmp_path):
    g = timemesh.gen_reg_mesh([0, 1e-6, 2e-6, 5e-6, 2e-5, 1e-4, 1e-3], 2.0)
    p = tmp_path / "infite.0"
    p.write_text(timemesh.infite0_text(g))
    back = refio.read_infite0(str(p))
    assert len(back) == len(g)
    np.testing.assert_allclose(back, g, rtol=6e-8, atol=0)


@pytest.mark.parametrize("seed", range(6))
def test_gen_reg_mesh_properties(seed):
    """Steps are powers of coef times the first step, never shrink, and the grid reaches the
    last input time within one step."""
    rng = np.random.default_rng(seed)
    t = np.concatenate([[0.0], np.cumsum(1e-6 * np.exp(rng.random(40) * np.linspace(0, 8, 40)))])
    coef = float(rng.choice([1.5, 2.0, 3.0]))
    g = np.array(timemesh.gen_reg_mesh(t, coef))
    h = np.diff(g)
    k = np.log(h / (t[1] - t[0])) / np.log(coef)
    np.testing.assert_allclose(k, np.round(k), atol=1e-6)
    assert np.all(np.diff(np.round(k)) >= 0)
    assert g[-1] >= t[-1] - 1e-12 and g[-2] <= t[-1]


def _case(d, seed, correct=False):
    """User times (a few decades of growing steps, with noise), source tables and tcff."""
    rng = np.random.default_rng(seed)
    t, h = [0.0], 1e-6 * (1 + rng.random())
    for dec in range(5):
        n = int(rng.integers(2, 9)) if correct else int(rng.integers(3, 7))
        for _ in range(n):
            step = h if correct else h * (1 + 0.2 * rng.random())
            t.append(t[-1] + step)
        h *= float(rng.choice([2.0, 3.0, 5.0]))
    with open(os.path.join(d, "infite.0"), "w") as f:
        f.write(timemesh.infite0_text(t))
    with open(os.path.join(d, "tcff"), "w") as f:
        f.write("2\n" if correct else f"{float(rng.choice([1.5, 2.0, 3.0]))}\n")
    tc = np.linspace(-1e-6, t[-1] * 1.1, 37)
    with open(os.path.join(d, "currentfunction"), "w") as f:
        for a in tc:
            f.write(f"{a:.14e} {np.exp(-max(a, 0) / (t[-1] / 3)):.14e}\n")
    td = np.sort(rng.random(23) * t[-1] * 1.2)
    with open(os.path.join(d, "deltafunction"), "w") as f:
        f.write(f"{25.0:.14e}\n{len(td)}\n")
        for a in td:
            f.write(f"{a:.14e} {rng.standard_normal():.14e}\n")


def _same_files(a, b, names):
    for name in names:
        with open(os.path.join(a, name), "rb") as fa, open(os.path.join(b, name), "rb") as fb:
            assert fa.read() == fb.read(), name


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_time_direct_solver_abi_matches_oracle(emtd, tmp_path, seed):
    a, b = tmp_path / "oracle", tmp_path / "abi"
    a.mkdir()
    _case(str(a), seed)
    shutil.copytree(a, b, dirs_exist_ok=True)
    timemesh.time_direct_solver(str(a))
    assert emtd.lib().emtd_time_direct_solver(str(b).encode()) == 0
    _same_files(str(a), str(b), FILES_TDS)
    g = refio.read_infite0(str(b / "infite.0"))
    assert g[0] == 0.0 and np.all(np.diff(g) > 0)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_correct_time_mesh_abi_matches_oracle(emtd, tmp_path, seed):
    a, b = tmp_path / "oracle", tmp_path / "abi"
    a.mkdir()
    _case(str(a), seed, correct=True)
    shutil.copytree(a, b, dirs_exist_ok=True)
    timemesh.correct_time_mesh(str(a))
    assert emtd.lib().emtd_correct_time_mesh(str(b).encode()) == 0
    _same_files(str(a), str(b), FILES_CTM)


def test_time_mesh_executables_match_oracle(emtd, tmp_path):
    for exe, fn, names, correct in (("timedirectsolver", timemesh.time_direct_solver, FILES_TDS, False),
                                    ("correcttimemesh", timemesh.correct_time_mesh, FILES_CTM, True)):
        path = os.path.join(PKG_DIR, exe)
        if not os.path.exists(path):
            pytest.skip(f"{exe} not built")
        a, b = tmp_path / (exe + "_o"), tmp_path / (exe + "_x")
        a.mkdir()
        _case(str(a), 11, correct=correct)
        shutil.copytree(a, b, dirs_exist_ok=True)
        fn(str(a))
        r = subprocess.run([path, "-d", str(b)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        _same_files(str(a), str(b), names)


def test_time_mesh_missing_input_leaves_files(emtd, tmp_path):
    _case(str(tmp_path), 5)
    os.remove(tmp_path / "deltafunction")
    before = sorted(os.listdir(tmp_path))
    rc = emtd.lib().emtd_time_direct_solver(str(tmp_path).encode())
    assert rc == 0x0002  # EMTD_ERR_NOFILE
    assert sorted(os.listdir(tmp_path)) == before  # nothing renamed or written
