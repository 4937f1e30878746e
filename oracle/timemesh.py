"""Restatement of the time-mesh preparation modules (oracle; test infrastructure only).

* ``time_direct_solver``: ``M/TimeDirectSolver/main.cpp:98-219`` with ``TimeStep::GenRegMesh``
  (``M/TimeDirectSolver/TimeStep.cpp:76-143``) and ``TimeStep::WriteTimes``
  (``TimeStep.cpp:144-171``): the times of infite.0 become a grid of piecewise constant steps
  growing by ``tcff``; currentfunction and deltafunction are resampled on it with ``Interpolate``
  (``main.cpp:43-96``: bisection, linear inside, end values outside).
* ``correct_time_mesh``: ``M/CorrectTimeMesh/main.cpp:87-266`` (``GenerateDecInfo``,
  ``SaveTimeMesh``, main): a decade whose length times ``tcff`` does not exceed the previous
  decade's length continues the previous decade's mean step; deltafunction is resampled and
  written to deltafunction and currentfunction.

Pinned by hand-derived known answers (tests/test_timemesh.py): a grid with three step
intervals and a truncated last count, a decade correction, and the single-point-decade case in
which the reference's mean step is 0/0 and the decade's points are dropped. Arithmetic follows
the reference's expression order (IEEE double, no fused multiply-add).
"""
from __future__ import annotations

import math
import os

from .refio import read_infite0


def _numbers(path):
    with open(path) as f:
        return [float(v) for v in f.read().split()]


def gen_reg_mesh(t, coef):
    """``TimeStep::GenRegMesh``: intervals of constant step, each closed at the first point
    whose step exceeds coef x the current step (or at the last point) with
    int((t_k - start) / step + 1) steps; the step grows by coef per interval."""
    t = [float(v) for v in t]
    step_old = t[1] - t[0]
    bounds, h, nsub = [t[0]], [step_old], []
    for k in range(2, len(t)):
        step = t[k] - t[k - 1]
        if step > coef * step_old or k + 1 == len(t):
            nsub.append(int((t[k] - bounds[-1]) / step_old + 1))
            bounds.append(bounds[-1] + nsub[-1] * step_old)
            step_old *= coef
            h.append(step_old)
    g = [t[0]]
    for i in range(len(bounds) - 1):
        for j in range(nsub[i]):
            g.append(bounds[i] + (j + 1) * h[i])
    while len(g) >= 2 and g[-2] > t[-1]:
        g.pop()
    return g


def find_interval(vec, elem):
    """``FindIntervalInDoubleVec``: the bisection of the reference, step for step."""
    i, k = 0, len(vec) - 1
    if not k or elem < vec[0] or elem > vec[k]:
        return -1
    j = (i + k) // 2
    while True:
        if elem > vec[j]:
            i = j
        else:
            k = j
        j = (i + k) // 2
        if j == i:
            return i


def interpolate(t, vt, vv):
    """``Interpolate`` with ``GetValueFromLinear``: ((t2 - t) v1 + (t - t1) v2) / (t2 - t1)."""
    if t < vt[0]:
        return vv[0]
    if t > vt[-1]:
        return vv[-1]
    it = find_interval(vt, t)
    t1, t2 = vt[it], vt[it + 1]
    return ((t2 - t) * vv[it] + (t - t1) * vv[it + 1]) / (t2 - t1)


def decade_starts(t):
    """``GenerateDecInfo``: a new decade where |1 - h_prev / h_cur| > 1e-3; closed by ntime."""
    ig = [0]
    hpre = t[1] - t[0]
    for i in range(2, len(t)):
        hcur = t[i] - t[i - 1]
        if abs(1.0 - hpre / hcur) > 1e-3:
            ig.append(i)
            hpre = hcur
    ig.append(len(t))
    return ig


def correct_mesh(t, tcff):
    """The corrected grid of CorrectTimeMesh's main (``main.cpp:231-263``)."""
    t = [float(v) for v in t]
    ig = decade_starts(t)
    out = []
    for d in range(len(ig) - 1):
        n = ig[d + 1] - ig[d]
        if n > 1 and (d == 0 or n * tcff > ig[d] - ig[d - 1]):
            out.extend(t[ig[d]:ig[d + 1]])
            continue
        if d == 0:
            raise ValueError("error in first decade")
        base = t[ig[d] - 1]
        num, den = base - t[ig[d - 1]], float(ig[d] - ig[d - 1] - 1)
        hpre = num / den if den != 0 else (math.nan if num == 0 else math.copysign(math.inf, num))
        i = 1
        while True:
            tcur = base + i * hpre
            if not (tcur - 0.5 * hpre < t[ig[d + 1] - 1]):
                break
            out.append(tcur)
            i += 1
    return out


def infite0_text(g):
    """``TimeStep::WriteTimes`` / SaveTimeMesh's infite.0: std::scientific, 7 digits."""
    n = len(g)
    s = (f"    ktime=  {n};  ntstop=  {n};   kiter=    1;   ntime=    1;   niter=    1;\n"
         "   kprogm=    1;  kpropi=    1;  kpropt=   -1;  kitrel=  250;              ;\n"
         "       u0=   0.00000e+000\n"
         " T I M E :\n")
    for i in range(0, n, 5):
        s += "".join(f" {v:.7e};" for v in g[i:i + 5]) + "\n"
    return s


def _table(t, v):
    return "".join(f"{a:.14e} {b:.14e}\n" for a, b in zip(t, v))


def _read_delta(path):
    x = _numbers(path)
    n = int(x[1])
    return x[0], x[2:2 + 2 * n:2], x[3:3 + 2 * n:2]


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)


def time_direct_solver(d):
    j = lambda name: os.path.join(d, name)  # noqa: E731
    tcff = _numbers(j("tcff"))[0]
    g = gen_reg_mesh(read_infite0(j("infite.0")), tcff)
    cf = _numbers(j("currentfunction"))
    ct, cv = cf[0:len(cf) - len(cf) % 2:2], cf[1::2][:len(cf) // 2]
    recrad, dtt, dtv = _read_delta(j("deltafunction"))
    cnew = [interpolate(x, ct, cv) for x in g]
    dnew = [interpolate(x, dtt, dtv) for x in g]
    os.replace(j("infite.0"), j("!infite.0"))
    _write(j("infite.0"), infite0_text(g))
    os.replace(j("currentfunction"), j("!currentfunction"))
    _write(j("currentfunction"), _table(g, cnew))
    os.replace(j("deltafunction"), j("!deltafunction"))
    _write(j("deltafunction"), f"{recrad:.14e}\n{len(g)}\n" + _table(g, dnew))


def correct_time_mesh(d):
    j = lambda name: os.path.join(d, name)  # noqa: E731
    tcff = _numbers(j("tcff"))[0]
    g = correct_mesh(read_infite0(j("infite.0")), tcff)
    recrad, dtt, dtv = _read_delta(j("deltafunction"))
    v = [interpolate(x, dtt, dtv) for x in g]
    for name in ("currentfunction", "infite.0", "deltafunction"):
        if os.path.exists(j(name)):
            os.replace(j(name), j("!!" + name))
    _write(j("deltafunction"), f"{recrad:.14e}\n{len(g)}\n" + _table(g, v))
    _write(j("currentfunction"), _table(g, v))
    _write(j("infite.0"), infite0_text(g))
