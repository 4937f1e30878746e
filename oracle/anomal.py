"""Restatement of the anomalous-edge export before a primary/secondary run (oracle; test
infrastructure only).

``unload_anomal_nonstat``: ``M/UnloadAnomalNonstat/TaskSP.cpp:104-127`` (z_sig_2d) and
``Time_approx_for_vfem::Read_data``, ``Unload_anomal_nodes``, ``Unload_anomal_nodesA0``
(``M/UnloadAnomalNonstat/time_approx.cpp:118-363``):

* the 12 edges of every element with |sigma0 - sigma| > 1e-6 are flagged, then unflagged when
  a node lies above the last depth of z_sig_2d plus 1e-3; an element is marked when its
  conductivity is anomalous or it has a flagged edge;
* isElemAnomal and mtrsizeforE0 hold one 0/1 per element and per edge; xyzVectorE0 the count
  and the midpoints of the flagged edges, TgCompE0 their unit tangents (node 1 - node 0), both
  with std::scientific and 14 digits;
* when a material's relative permeability differs from the normal one (|mu - mu0| / MU_0 >
  1e-6), the elements with |mu0 - mu| > 1e-6 in H/m get the same export without the depth
  limit (isElemAnomal_A, mtrsizeforA0, xyzVectorA0, TgCompA0), the last two in the stream's
  default format (six significant digits, ``%g``).

Pinned by properties that fix the geometry independently of the rule's restatement (flagged
edges lie in anomalous elements and below the depth limit; on the axis-aligned meshes the
tangents are unit axis vectors and the midpoints are edge centres) and by agreement with the
generator's own mtrsizeforE0 (tests/test_anomal.py).
"""
from __future__ import annotations

import math
import os

import numpy as np

from .refio import read_inftry, read_material_table

MU_0 = 1.25663706143591729e-6  # M/UnloadAnomalNonstat/vfem_const.h:21


def _mesh(d):
    j = lambda name: os.path.join(d, name)  # noqa: E731
    kuz, kpar, _ = read_inftry(j("inftry.dat"))
    xyz = np.fromfile(j("xyz.dat"), dtype="<f8", count=3 * kuz).reshape(kuz, 3)
    nvkat = np.fromfile(j("nvkat.dat"), dtype="<i4", count=kpar) - 1
    with open(j("tsize3d_.dat")) as f:
        n = int(f.read().split()[1])
    edges = np.fromfile(j("nodesforedges.dat"), dtype="<i4", count=2 * n).reshape(n, 2) - 1
    ed = np.fromfile(j("edges.dat"), dtype="<i4", count=25 * kpar).reshape(kpar, 25)[:, :12] - 1
    sig, sig0 = read_material_table(j("sig3d"))
    mur, mur0 = read_material_table(j("mu3d"))
    return xyz, nvkat, edges, ed, sig, sig0, mur * MU_0, mur0 * MU_0


def _export(d, xyz, edges, elem, edge, names, fmt):
    j = lambda name: os.path.join(d, name)  # noqa: E731
    with open(j(names[0]), "w", newline="") as f:
        f.write("".join("1\n" if e else "0\n" for e in elem))
    with open(j(names[1]), "w", newline="") as f:
        f.write("".join("1\n" if e else "0\n" for e in edge))
    mid, tg = [f"{int(np.count_nonzero(edge))}\n"], []
    for i in np.nonzero(edge)[0]:
        a, c = xyz[edges[i, 0]], xyz[edges[i, 1]]
        m = [0.5 * (float(a[k]) + float(c[k])) for k in range(3)]
        p = [float(c[k]) - float(a[k]) for k in range(3)]
        ln = math.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2])
        mid.append(" ".join(fmt % v for v in m) + "\n")
        tg.append(" ".join(fmt % (v / ln) for v in p) + "\n")
    with open(j(names[2]), "w", newline="") as f:
        f.write("".join(mid))
    with open(j(names[3]), "w", newline="") as f:
        f.write("".join(tg))


def unload_anomal_nonstat(d):
    with open(os.path.join(d, "z_sig_2d")) as f:
        lines = f.read().split("\n")
    nzl = int(lines[0].split()[0])
    z0 = float(lines[nzl].split()[0]) + 1e-3
    xyz, nvkat, edges, ed, sig, sig0, mu, mu0 = _mesh(d)
    elem = np.abs(sig0[nvkat] - sig[nvkat]) > 1e-6
    edge = np.zeros(edges.shape[0], dtype=bool)
    edge[ed[elem].reshape(-1)] = True
    above = (xyz[edges[:, 0], 2] > z0) | (xyz[edges[:, 1], 2] > z0)
    edge &= ~above
    elem = elem | edge[ed].any(axis=1)
    _export(d, xyz, edges, elem, edge, ("isElemAnomal", "mtrsizeforE0", "xyzVectorE0", "TgCompE0"), "%.14e")
    if np.count_nonzero(np.abs(mu - mu0) / MU_0 > 1e-6):
        elem_a = np.abs(mu0[nvkat] - mu[nvkat]) > 1e-6
        edge_a = np.zeros(edges.shape[0], dtype=bool)
        edge_a[ed[elem_a].reshape(-1)] = True
        _export(d, xyz, edges, elem_a, edge_a, ("isElemAnomal_A", "mtrsizeforA0", "xyzVectorA0", "TgCompA0"), "%g")
