"""The request loop on the GPU: results alone vs in the mix, admission by memory, statistics.

Requests are small models of the batched/serving generator (random mesh sizes, layers,
blocks, A/E and 2-/3-point). Each request's tables must be the same when it runs alone and
when it runs concurrently with others (the arithmetic is per request and deterministic, so
the bar is bit-exact), and one of them is checked against the oracle at the module's FP64 bar.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import modelgen
from oracle import calc, refio

pytestmark = pytest.mark.gpu

TABLES = ("eall_anom.1", "edsall_anom.1", "ball_anom.1")


def _write_requests(base, n, seed=11, axis_range=(6, 10), steps=(4, 3)):
    mg = modelgen()
    dirs = []
    for i in range(n):
        spec = mg.request_spec(seed, i, axis_range=axis_range, steps=steps)
        spec.times = mg.time_grid(1e-3, steps[0], steps[1])
        tq = mg.quantize_times(spec.times)
        spec.print_window = (tq[2] if spec.scheme == 3 else tq[1], tq[-1])
        d = os.path.join(base, f"req{i}")
        os.makedirs(d)
        mg.write_model(spec, d)
        dirs.append((d, spec))
    return dirs


def _read(path):
    with open(path) as f:
        return f.read()


def test_server_alone_and_in_the_mix(emtd, tmp_path):
    reqs = _write_requests(str(tmp_path), 5)
    # alone: one request at a time through the drop-in entry point
    alone = {}
    for d, spec in reqs:
        a = d + "_alone"
        shutil.copytree(d, a)
        print("alone", a, spec.formulation, spec.scheme, flush=True)
        emtd.calc_nonstat_3d_line(a)
        alone[d] = {t: _read(os.path.join(a, t)) for t in TABLES}
    # in the mix: all submitted at once, four at a time on the device
    srv = emtd.Server(max_active=4)
    ids = [srv.submit(d) for d, _ in reqs]
    for rid in ids:
        inf = srv.wait(rid)
        assert inf.status == emtd.REQ_DONE and inf.steps > 0
    st = srv.stats()
    print("completed", st.completed, "peak running", st.peak_running, "p50", st.p50_latency_s, "p99",
          st.p99_latency_s)
    assert st.completed == len(reqs) and st.failed == 0
    assert st.p50_latency_s <= st.p99_latency_s <= st.max_latency_s
    srv.close()
    for d, _ in reqs:
        for t in TABLES:
            assert _read(os.path.join(d, t)) == alone[d][t], (d, t)
    # one request against the oracle (FP64 bar 1e-6 on the receiver tables)
    d, spec = reqs[0]
    m = refio.read_model(d)
    ref = calc.run(m, spec.formulation, spec.scheme)
    od = os.path.join(str(tmp_path), "oracle0")
    os.makedirs(od)
    refio.write_tables(od, m, ref.first, ref.last, ref.eds, ref.bfield, ref.efield, spec.formulation)
    for t in TABLES:
        a = _table(os.path.join(d, t))
        b = _table(os.path.join(od, t))
        assert a.shape == b.shape
        scale = max(np.abs(b[:, 1:]).max(), 1e-300)
        assert np.abs(a[:, 1:] - b[:, 1:]).max() / scale < 1e-6, t


def _table(path):
    rows = []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        if lines[i].startswith("      t (ms)"):
            i += 1
            while i < len(lines) and "\t" in lines[i]:
                rows.append([float(v) for v in lines[i].split("\t")])
                i += 1
        else:
            i += 1
    return np.array(rows)


def test_server_admission_by_memory(emtd, tmp_path):
    reqs = _write_requests(str(tmp_path), 3, seed=12)
    # learn the footprints with a generous budget, one at a time
    srv = emtd.Server(max_active=1)
    fps = [srv.wait(srv.submit(d)).footprint_bytes for d, _ in reqs]
    srv.close()
    print("footprints", fps)
    # a budget that holds the largest request but not two of them: never two at once
    budget = max(fps) + 0.5 * min(fps)
    srv = emtd.Server(hbm_budget_bytes=budget, max_active=3)
    for d, _ in reqs:
        srv.submit(d)
    srv.drain()
    st = srv.stats()
    assert st.completed == 3 and st.failed == 0
    assert st.peak_running == 1 and st.peak_bytes <= budget
    srv.close()
    # a budget below every footprint: the request fails with EMTD_ERR_NOMEM, cleanly
    srv = emtd.Server(hbm_budget_bytes=0.5 * min(fps), max_active=2)
    rid = srv.submit(reqs[0][0])
    inf = srv.wait(rid, check=False)
    assert inf.status == emtd.REQ_FAILED and inf.rc == 0x0001
    srv.close()
