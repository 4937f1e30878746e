"""Restatement of the downstream output chain (oracle; test infrastructure only).

* ``sum_nonstat_3d_only``: ``M/SumNonstat3DOnly/main.cpp:171-384`` with the picket reader and
  writer of ``M/SumNonstat3DOnly/picket.cpp`` (``ReadPickets``/``WritePickets``,
  ``CopyValuesOnlyAnomal``) and ``GetNumberOfPlaces``/``read_time_mesh`` of
  ``M/SumNonstat3DOnly/utils_direct_solver.cpp:26-100``. The module's anomalous tables become
  one file per component whose columns are (t ms, normal = 0, anomalous, total = anomalous).
* ``summator_edmn``: ``M/SummatorEDMN/main.cpp:40-295`` with ``Picket::ReadPicket``/``WritePicket``
  of ``M/SummatorEDMN/Picket.h``: the E field integrated along each MN line,
  dV = sum_k E(p_k).(p_k - p_{k-1}) with the first segment 2 (p_0 - A) (midpoint rule when the
  points are the segment midpoints); with ddV the second half of the line enters with a minus
  sign.
"""
from __future__ import annotations

import os

import numpy as np

from .refio import read_infite0

LEG_B = ["t (ms)        Bx (mTl)     Bx (mTl)     Bx (mTl)", "t (ms)        By (mTl)     By (mTl)     By (mTl)",
         "t (ms)        Bz (mTl)     Bz (mTl)     Bz (mTl)"]
LEG_F = ["t (ms)      Emfx (mV)   Emfx (mV)   Emfx (mV)", "t (ms)      Emfy (mV)   Emfy (mV)   Emfy (mV)",
         "t (ms)      Emfz (mV)   Emfz (mV)   Emfz (mV)"]
LEG_E = ["t (ms)        Ex (V/m)     Ex (V/m)     Ex (V/m)", "t (ms)        Ey (V/m)     Ey (V/m)     Ey (V/m)",
         "t (ms)        Ez (V/m)     Ez (V/m)     Ez (V/m)"]
NAMES_B = ["Bxall", "Byall", "Bzall"]
NAMES_F = ["Edsxall", "Edsyall", "Edsall"]
NAMES_E = ["Exall", "Eyall", "Ezall"]
LEG_DV = "         t, ms                         dVp, mV                         dVs, mV                         dV, mV"
LEG_DDV = "         t, ms                         ddVp, mV                         ddVs, mV                         ddV, mV"


def _e(v):
    return f"{v:.16e}"


def read_pickets(path, nr, nt):
    """``ReadPickets``: per receiver (iPx, iPy, (Px, Py, Pz), (nt, 4) rows)."""
    with open(path) as f:
        lines = f.read().split("\n")
    out, k = [], 0
    for _ in range(nr):
        ix, iy = (int(v) for v in lines[k].split()[:2])
        p = tuple(float(v) for v in lines[k + 1].split()[:3])
        k += 6  # coordinates, "1", "1 1 1 1", "1", legend
        rows = np.array([[float(v) for v in lines[k + i].split()[:4]] for i in range(nt)]).reshape(nt, 4)
        k += nt
        out.append((ix, iy, p, rows))
    return out


def write_pickets(path, pickets, legend, t, comp):
    """``WritePickets`` with values (0, v, v) of component ``comp`` of each picket."""
    with open(path, "w") as f:
        for ix, iy, p, rows in pickets:
            f.write(f"{ix} {iy}\n{_e(p[0])} {_e(p[1])} {_e(p[2])}\n1\n1 1 1 1\n1\n{legend}\n")
            for i in range(len(t)):
                v = rows[i, 1 + comp]
                f.write(f"{_e(t[i])} {_e(0.0)} {_e(v)} {_e(v)}\n")


def sum_nonstat_3d_only(d: str) -> None:
    j = lambda name: os.path.join(d, name)  # noqa: E731
    with open(j("group")) as f:
        tok = f.read().split()
    nprof = int(tok[0])
    npls = sum(int(tok[3 + 3 * g]) - int(tok[2 + 3 * g]) + 1 for g in range(nprof))
    with open(j("timeintervalforprint")) as f:
        ftime, ltime = (float(v) for v in f.read().split()[:2])
    times = read_infite0(j("infite.0"))
    tw = np.array([tj * 1000.0 for tj in times if ftime <= tj <= ltime])
    nt = len(tw)
    for kind, counts_file in (("B", "recvsb"), ("E", "recvse")):
        with open(j(counts_file)) as f:
            counts = [int(v) for v in f.read().split()[:npls]]
        for ipls in range(npls):
            nall = counts[ipls]
            if kind == "B":
                pk = read_pickets(j(f"ball_anom.{ipls + 1}"), nall, nt)
                for c in range(3):
                    write_pickets(j(f"{NAMES_B[c]}.{ipls + 1}"), pk, LEG_B[c], tw, c)
                pk = read_pickets(j(f"edsall_anom.{ipls + 1}"), nall, nt)
                for c in range(3):
                    write_pickets(j(f"{NAMES_F[c]}.{ipls + 1}"), pk, LEG_F[c], tw, c)
            else:
                pk = read_pickets(j(f"eall_anom.{ipls + 1}"), nall, nt)
                for c in range(3):
                    write_pickets(j(f"{NAMES_E[c]}.{ipls + 1}"), pk, LEG_E[c], tw, c)


def _read_picket_file(path, count):
    """``Picket::ReadPicket`` repeated: (t, (n, a, s)) arrays per picket."""
    with open(path) as f:
        lines = [ln for ln in f.read().split("\n")]
    out, k = [], 0
    for _ in range(count):
        k += 6
        ts, vals = [], []
        while k < len(lines):
            p = lines[k].split()
            try:
                if len(p) < 4:
                    raise ValueError
                row = [float(v) for v in p[:4]]
            except ValueError:
                break
            ts.append(row[0])
            vals.append(row[1:])
            k += 1
        out.append((np.array(ts), np.array(vals).reshape(-1, 3)))
    return out


def summator_edmn(d: str, with_ddv: bool, profile: int, place: int) -> None:
    j = lambda name: os.path.join(d, name)  # noqa: E731
    with open(j("xyzmn")) as f:
        nlin = int(f.read().split()[0])
    if not nlin:
        return
    with open(j(f"xyzmn{place + 1}_{profile}")) as f:
        tok = f.read().split()
    q = int(tok[0])
    ab = np.array([float(v) for v in tok[1:1 + 6 * q]]).reshape(q, 6)
    with open(j(f"xyzVectorElin{place + 1}_{profile}")) as f:
        tok = f.read().split()
    ne = int(tok[0])
    pts = np.array([float(v) for v in tok[1:1 + 3 * ne]]).reshape(ne, 3)
    per = ne // q
    E = [_read_picket_file(j(f"e{c}all{profile}_{place}"), q * per) for c in "xyz"]
    dv, ddv = [], []
    m = 0
    for ii in range(q):
        A = ab[ii, :3]
        acc = None
        acc2 = None
        prev = None
        for jj in range(per):
            pe = pts[m]
            vab = (pe - A) * 2.0 if jj == 0 else pe - prev
            prev = pe
            val = E[0][m][1] * vab[0] + E[1][m][1] * vab[1] + E[2][m][1] * vab[2]
            if jj == 0:
                acc = val.copy()
                acc2 = val.copy()
                times = E[0][m][0]
            else:
                acc = acc + val
                acc2 = acc2 + (-val if jj >= per // 2 else val)
            m += 1
        mid = 0.5 * (ab[ii, :3] + ab[ii, 3:])
        dv.append((mid, times, acc))
        ddv.append((mid, times, acc2))

    def write(path, rows, leg):
        with open(path, "w") as f:
            for mid, times, vals in rows:
                f.write(f"{int(mid[0])} {int(mid[1])}\n{_e(mid[0])} {_e(mid[1])} {_e(mid[2])}\n1\n1 1 1 1\n1\n{leg}\n")
                for i in range(len(times)):
                    f.write(f"{_e(times[i])} \t{_e(vals[i, 0])} \t{_e(vals[i, 1])} \t{_e(vals[i, 2])}\n")

    write(j(f"edmnall{profile}_{place}"), dv, LEG_DV)
    if with_ddv:
        write(j(f"edmoonall{profile}_{place}"), ddv, LEG_DDV)
