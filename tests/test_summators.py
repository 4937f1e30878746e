"""The output chain after the module (host only, CPU): SumNonstat3DOnly and SummatorEDMN.

The oracle is pinned by known answers of the MN-line integration (a constant field gives
E.(B - A); the midpoint rule is exact for a linear field; ddV of a constant field is zero), and
the C ABI must write the same bytes as the oracle."""
import os
import shutil

import numpy as np
import pytest

from conftest import make_model
from oracle import calc, refio, summators


def _write_picket_file(path, pts, times, vals):
    """pickets at pts with rows (t, n, a, s): vals (npts, nt, 3)."""
    with open(path, "w") as f:
        for p, v in zip(pts, vals):
            f.write(f"{int(p[0])} {int(p[1])}\n{p[0]:.16e} {p[1]:.16e} {p[2]:.16e}\n1\n1 1 1 1\n1\nlegend\n")
            for t, row in zip(times, v):
                f.write(f"{t:.16e} {row[0]:.16e} {row[1]:.16e} {row[2]:.16e}\n")


def _mn_case(d, field, per=8, profile=3, place=1):
    """Two MN lines with `per` receiver points at the segment midpoints; field(p, t) -> E (3,)."""
    lines = [((-100.0, 20.0, 0.0), (60.0, 20.0, 0.0)), ((10.0, -50.0, 0.0), (40.0, 70.0, 0.0))]
    times = np.array([0.1, 0.2, 0.5])
    with open(os.path.join(d, "xyzmn"), "w") as f:
        f.write(f"{len(lines)}\n")
    with open(os.path.join(d, f"xyzmn{place + 1}_{profile}"), "w") as f:
        f.write(f"{len(lines)}\n")
        for a, b in lines:
            f.write(" ".join(f"{v:.16e}" for v in (*a, *b)) + "\n")
    pts = []
    for a, b in lines:
        a, b = np.array(a), np.array(b)
        for k in range(per):
            pts.append(a + (b - a) * (k + 0.5) / per)
    pts = np.array(pts)
    with open(os.path.join(d, f"xyzVectorElin{place + 1}_{profile}"), "w") as f:
        f.write(f"{len(pts)}\n")
        for p in pts:
            f.write(" ".join(f"{v:.16e}" for v in p) + "\n")
    # columns (normal, anomalous, total) = (1, 2, 3) x field
    E = np.array([[field(p, t) for t in times] for p in pts])  # (npts, nt, 3 comps)
    for c, name in enumerate("xyz"):
        vals = np.stack([E[:, :, c] * s for s in (1.0, 2.0, 3.0)], axis=2)
        _write_picket_file(os.path.join(d, f"e{name}all{profile}_{place}"), pts, times, vals)
    return lines, times, profile, place


def _read_out(path):
    rows = []
    with open(path) as f:
        for ln in f:
            p = ln.replace("\t", " ").split()
            if len(p) == 4 and ln.strip() != "1 1 1 1":  # data rows, not the picket header
                try:
                    rows.append([float(v) for v in p])
                except ValueError:
                    pass
    return np.array(rows)


@pytest.mark.parametrize("kind", ["constant", "linear"])
def test_summator_edmn_known_answers(tmp_path, kind):
    e0 = np.array([1.5e-3, -2.0e-3, 7.0e-4])
    c = 3.0e-5
    field = (lambda p, t: e0 * np.exp(-t)) if kind == "constant" else (lambda p, t: np.array([c * p[0], c * p[1], 0.0]) * (1 + t))
    d = str(tmp_path)
    lines, times, prof, place = _mn_case(d, field)
    summators.summator_edmn(d, True, prof, place)
    dv = _read_out(os.path.join(d, f"edmnall{prof}_{place}"))
    ddv = _read_out(os.path.join(d, f"edmoonall{prof}_{place}"))
    for i, (a, b) in enumerate(lines):
        a, b = np.array(a), np.array(b)
        m = 0.5 * (a + b)
        for l, t in enumerate(times):
            if kind == "constant":
                want = e0 @ (b - a) * np.exp(-t)
                want_dd = 0.0
            else:
                # integral of c (x dx + y dy) = c/2 |p|^2 between the ends
                f = lambda p: 0.5 * c * (p[0] ** 2 + p[1] ** 2) * (1 + t)  # noqa: E731
                want = f(b) - f(a)
                want_dd = (f(m) - f(a)) - (f(b) - f(m))
            row = dv[i * len(times) + l]
            assert row[0] == pytest.approx(t)
            np.testing.assert_allclose(row[1:], [want, 2 * want, 3 * want], rtol=1e-12, atol=1e-18)
            np.testing.assert_allclose(ddv[i * len(times) + l][1:], [want_dd, 2 * want_dd, 3 * want_dd],
                                       rtol=1e-12, atol=1e-15 * abs(want) + 1e-18)


def test_summator_edmn_abi_matches_oracle(emtd, tmp_path):
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    os.makedirs(a)
    _, _, prof, place = _mn_case(a, lambda p, t: np.array([np.sin(p[0] / 50), p[1] * 1e-3, 0.2]) * t)
    shutil.copytree(a, b)
    summators.summator_edmn(a, True, prof, place)
    assert emtd.lib().emtd_summator_edmn(b.encode(), 1, prof, place) == 0
    for name in (f"edmnall{prof}_{place}", f"edmoonall{prof}_{place}"):
        with open(os.path.join(a, name)) as f1, open(os.path.join(b, name)) as f2:
            assert f1.read() == f2.read(), name


def test_sum_nonstat_3d_only_abi_matches_oracle(emtd, mg, tmp_path):
    """Module tables (written by the oracle's table writer from random values, two sources)
    through SumNonstat3DOnly: same bytes from the C ABI and the oracle, and the component files
    hold (0, value, value) of the table's column."""
    d, spec = make_model(tmp_path / "a", cfg=1, cells=(4, 4, 3), steps=(3, 2))
    spec.extra_wires = [((-5.0, -5.0, 0.0), (5.0, 5.0, 0.0))]
    mg.write_model(spec, d)
    m = refio.read_model(d)
    first, last = calc.print_window(m)
    nt = last - first + 1
    rng = np.random.default_rng(1)
    eds = rng.standard_normal((nt, m.npls, len(m.recB), 3))
    bf = rng.standard_normal((nt, m.npls, len(m.recB), 3))
    ef = rng.standard_normal((nt, m.npls, len(m.recE), 3))
    refio.write_tables(d, m, first, last, eds, bf, ef, "A")
    b = str(tmp_path / "b")
    shutil.copytree(d, b)
    summators.sum_nonstat_3d_only(d)
    assert emtd.lib().emtd_sum_nonstat_3d_only(b.encode()) == 0
    names = [f"{n}.{k}" for n in summators.NAMES_B + summators.NAMES_F + summators.NAMES_E for k in (1, 2)]
    for name in names:
        with open(os.path.join(d, name)) as f1, open(os.path.join(b, name)) as f2:
            assert f1.read() == f2.read(), name
    # Exall.2 column 2 = 1000 x the second source's Ex at its own receivers (the second half of
    # the receiver list; 1000 is the table's scale), column 1 = 0
    rows = _read_out(os.path.join(b, "Exall.2"))
    nr = m.rec_per_src_E[1]
    want = ef[:, 1, m.rec_per_src_E[0]:m.rec_per_src_E[0] + nr, 0].T.reshape(-1) * 1000.0
    np.testing.assert_array_equal(rows[:, 1], 0.0)
    np.testing.assert_allclose(rows[:, 2], want, rtol=1e-15)
    np.testing.assert_array_equal(rows[:, 2], rows[:, 3])
