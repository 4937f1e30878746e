"""Shared test setup: markers, package import and small model fixtures."""
import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG_NAME = "a-e_formulations_3d_timedomain_em_horizontalgroundedwiresource_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running check")


def pkg():
    return importlib.import_module(PKG_NAME)


def modelgen():
    return importlib.import_module(PKG_NAME + ".modelgen")


@pytest.fixture(scope="session")
def emtd():
    return pkg()


@pytest.fixture(scope="session")
def mg():
    return modelgen()


def make_model(tmpdir, cfg=1, cells=(8, 8, 6), steps=(4, 3), t_end=1e-3, current=None, form=None,
               scheme=None, seed=None):
    """Write a small model of a named config into tmpdir and return (dir, spec)."""
    m = modelgen()
    spec = m.config_spec(cfg, small=cells, seed=seed)
    spec.times = m.time_grid(t_end, steps[0], steps[1])
    if form is not None:
        spec.formulation = form
    if scheme is not None:
        spec.scheme = scheme
    if current is not None:
        spec.current = current(len(spec.times))
    tq = m.quantize_times(spec.times)
    spec.print_window = (tq[2], tq[-1]) if spec.scheme == 3 else (tq[1], tq[-1])
    d = str(tmpdir)
    m.write_model(spec, d)
    return d, spec


# Sweep strategies: "split" is the default (levels with fewer than 1024 fronts split over many
# workgroups, which at these sizes is every level); "small" turns the split sweeps off so that
# the levels whose fronts have at most 64 pivots run one wavefront per front and the others one
# workgroup per front; "classic" turns both off (one workgroup per front everywhere).
SWEEPS = {"split": {}, "small": {"EMTD_SPLIT_LEVEL_FRONTS": "0"},
          "classic": {"EMTD_SPLIT_LEVEL_FRONTS": "0", "EMTD_SMALL_FRONT_SWEEPS": "0"}}


@pytest.fixture(params=list(SWEEPS))
def sweeps(request, monkeypatch):
    for k, v in SWEEPS[request.param].items():
        monkeypatch.setenv(k, v)
    print("sweeps", request.param, flush=True)
    return request.param
