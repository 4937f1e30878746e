"""The anomalous-edge export before a primary/secondary run (host only, CPU): UnloadAnomalNonstat.

The oracle is pinned by the geometry of its output (every flagged edge lies in an element of
anomalous conductivity and below the depth limit; on these axis-aligned meshes the tangents are
unit axis vectors and the midpoints are edge centres), by the depth limit moving with z_sig_2d,
and by agreement with the generator's own mtrsizeforE0; the C ABI and the drop-in executable
must write the same bytes as the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import modelgen
from oracle import anomal, refio

PKG_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "a-e_formulations_3d_timedomain_em_horizontalgroundedwiresource_amd")
FILES_E = ["isElemAnomal", "mtrsizeforE0", "xyzVectorE0", "TgCompE0"]
FILES_A = ["isElemAnomal_A", "mtrsizeforA0", "xyzVectorA0", "TgCompA0"]


def _model(d, z_top=0.0, mu_block=None):
    """A layered model with one conductive block (and optionally one permeable material),
    primary/secondary inputs, and z_sig_2d whose last depth is z_top."""
    mg = modelgen()
    spec = mg.config_spec(1, small=(8, 8, 6))
    spec.blocks = [mg.Block(-20.0, 15.0, -15.0, 20.0, -28.0, -8.0, 1.5)]
    spec.layer_tops = [0.0, 12.0]
    spec.layer_sigma = [0.02, 0.1]
    spec.times = mg.time_grid(1e-3, 4, 3)
    spec.primary_secondary = f"emtdanom{os.getpid()}"
    info = mg.write_model(spec, d)
    with open(os.path.join(d, "z_sig_2d"), "w") as f:
        f.write(f"2\n{-100.0:.6f} 0.1\n{z_top:.6f} 0.02\n")
    if mu_block is not None:
        # give the last material a relative permeability of mu_block in the 3D model
        rows = [ln.split() for ln in open(os.path.join(d, "mu3d")).read().split("\n") if ln.strip()]
        rows[-1][1] = repr(float(mu_block))
        with open(os.path.join(d, "mu3d"), "w") as f:
            f.write("".join(" ".join(r) + "\n" for r in rows))
    return info


def _same(a, b, names):
    for name in names:
        with open(os.path.join(a, name), "rb") as fa, open(os.path.join(b, name), "rb") as fb:
            assert fa.read() == fb.read(), name


def test_oracle_matches_generator_rule(tmp_path):
    d = str(tmp_path)
    _model(d)
    shutil.copy(os.path.join(d, "mtrsizeforE0"), os.path.join(d, "mtrsizeforE0.gen"))
    anomal.unload_anomal_nonstat(d)
    with open(os.path.join(d, "mtrsizeforE0"), "rb") as a, open(os.path.join(d, "mtrsizeforE0.gen"), "rb") as b:
        assert a.read() == b.read()


def test_oracle_geometry(tmp_path):
    d = str(tmp_path)
    _model(d)
    anomal.unload_anomal_nonstat(d)
    m = refio.read_model(d)
    flags = np.array([int(v) for v in open(os.path.join(d, "mtrsizeforE0")).read().split()], dtype=bool)
    assert flags.sum() > 0 and len(flags) == len(m.edges)
    elem_sig = np.abs(m.sigma0[m.nvkat] - m.sigma[m.nvkat]) > 1e-6
    in_anom = np.zeros(len(m.edges), dtype=bool)
    in_anom[m.ed[elem_sig].reshape(-1)] = True
    assert np.all(in_anom[flags])                                    # only edges of anomalous elements
    assert np.all(m.xyz[m.edges[flags]][:, :, 2] <= 1e-3)          # none above the depth limit
    tok = open(os.path.join(d, "xyzVectorE0")).read().split()
    assert int(tok[0]) == flags.sum()
    mid = np.array([float(v) for v in tok[1:]]).reshape(-1, 3)
    np.testing.assert_array_equal(mid, 0.5 * (m.xyz[m.edges[flags, 0]] + m.xyz[m.edges[flags, 1]]))
    tg = np.array([float(v) for v in open(os.path.join(d, "TgCompE0")).read().split()]).reshape(-1, 3)
    assert np.all(np.sort(np.abs(tg), axis=1) == [0.0, 0.0, 1.0])  # unit axis vectors
    elem = np.array([int(v) for v in open(os.path.join(d, "isElemAnomal")).read().split()], dtype=bool)
    assert np.all(elem[elem_sig]) and np.all(elem == (elem_sig | flags[m.ed].any(axis=1)))
    assert not os.path.exists(os.path.join(d, "mtrsizeforA0"))    # no permeability anomaly


def test_depth_limit_follows_z_sig_2d(tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    _model(str(a), z_top=0.0)
    _model(str(b), z_top=-12.0)   # the limit below the block's top rows of edges
    anomal.unload_anomal_nonstat(str(a))
    anomal.unload_anomal_nonstat(str(b))
    fa = np.array([int(v) for v in open(a / "mtrsizeforE0").read().split()], dtype=bool)
    fb = np.array([int(v) for v in open(b / "mtrsizeforE0").read().split()], dtype=bool)
    m = refio.read_model(str(b))
    assert fb.sum() < fa.sum() and np.all(fa[fb])
    assert np.all(m.xyz[m.edges[fb]][:, :, 2] <= -12.0 + 1e-3)


@pytest.mark.parametrize("mu_block", [None, 50.0])
def test_unload_anomal_abi_matches_oracle(emtd, tmp_path, mu_block):
    a, b = tmp_path / "oracle", tmp_path / "abi"
    a.mkdir()
    _model(str(a), mu_block=mu_block)
    shutil.copytree(a, b, dirs_exist_ok=True)
    anomal.unload_anomal_nonstat(str(a))
    assert emtd.lib().emtd_unload_anomal_nonstat(str(b).encode()) == 0
    _same(str(a), str(b), FILES_E + (FILES_A if mu_block else []))
    assert os.path.exists(b / "mtrsizeforA0") == (mu_block is not None)


def test_unload_anomal_executable_matches_oracle(emtd, tmp_path):
    path = os.path.join(PKG_DIR, "unloadanomalnonstat")
    if not os.path.exists(path):
        pytest.skip("unloadanomalnonstat not built")
    a, b = tmp_path / "oracle", tmp_path / "exe"
    a.mkdir()
    _model(str(a), mu_block=4.0)
    shutil.copytree(a, b, dirs_exist_ok=True)
    anomal.unload_anomal_nonstat(str(a))
    r = subprocess.run([path, "-d", str(b)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    _same(str(a), str(b), FILES_E + FILES_A)


def test_unload_anomal_missing_layers_file(emtd, tmp_path):
    _model(str(tmp_path))
    os.remove(tmp_path / "z_sig_2d")
    assert emtd.lib().emtd_unload_anomal_nonstat(str(tmp_path).encode()) == 0x0002  # EMTD_ERR_NOFILE
