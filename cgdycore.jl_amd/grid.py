"""Cubed-sphere topology and the spectral-element global numbering ``Glob``.

Host-side setup (runs once, not on the hot path).  It restates the
reference's single-rank grid pipeline so that the local-to-global map fed to
the GPU kernels is the reference's own:

* node / edge / face generation:  ``src/Grids/CubedGrid.jl:1-513``
  (``CubePoint`` :377-391, ``InsertFaceEdge`` :393-441,
  ``InsertEdgeEdge`` :443-462, ``InsertFaceFace`` :464-513)
* face construction, edge sorting, orientation flip:
  ``src/Grids/Face.jl:48-167`` with ``OrientFaceSphere`` (dot(n, Mid))
* edge orientation sweep:  ``src/Grids/Orientation.jl`` (``Orientation!``)
* face rotation:  ``src/Grids/Renumbering.jl`` (``RenumberingFace4!``)
* single-rank sub-grid renumbering (first-appearance order):
  ``src/Grids/SubGridGhost.jl:1-162``
* CG numbering:  ``src/DyCore/NumberingFem.jl:21-131`` (``NumberingFemCGQuad``)

Indices are kept 1-based in the returned ``glob`` (the reference's Int64
convention); the device side converts once.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def _cube_point(i1, i2, i3, n, x, y, z, rad):
    """``CubePoint`` (CubedGrid.jl:377-391): gnomonic point projected to the sphere."""
    if i1 > 0:
        x = math.tan(i1 * math.pi / (2 * n) - 0.25 * math.pi)
    if i2 > 0:
        y = math.tan(i2 * math.pi / (2 * n) - 0.25 * math.pi)
    if i3 > 0:
        z = math.tan(i3 * math.pi / (2 * n) - 0.25 * math.pi)
    nrm = math.sqrt(x * x + y * y + z * z)
    return (x / nrm * rad, y / nrm * rad, z / nrm * rad)


@dataclass
class _Edge:
    N: list  # two node numbers (1-based)
    E: int = 0
    F: list = field(default_factory=lambda: [0, 0])
    NumF: int = 0


@dataclass
class _Face:
    N: list
    E: list
    Orientation: int = 1


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _make_face(edges_f, nodes, edges, pos):
    """``Face(EdgesF,Nodes,Edges,Pos,...)`` (Face.jl:48-167), quad case, ChangeOrient=3."""
    edges_f = list(edges_f)
    ne = len(edges_f)
    for ie in range(ne):
        e = edges[edges_f[ie]]
        e.NumF += 1
        e.F[e.NumF - 1] = pos
    # sort the edges into a closed cycle
    fe = [0] * ne
    fe[0] = edges_f[0]
    n2 = edges[fe[0]].N[1]
    for ie in range(1, ne):
        for ie1 in range(ie, ne):
            cand = edges[edges_f[ie1]]
            if n2 == cand.N[0]:
                fe[ie] = edges_f[ie1]
                n2 = cand.N[1]
                edges_f[ie1] = edges_f[ie]
                edges_f[ie] = fe[ie]
                break
            elif n2 == cand.N[1]:
                fe[ie] = edges_f[ie1]
                n2 = cand.N[0]
                edges_f[ie1] = edges_f[ie]
                edges_f[ie] = fe[ie]
                break
    fn = [0] * ne
    fn[0], fn[1] = edges[fe[0]].N[0], edges[fe[0]].N[1]
    for ie in range(1, ne - 1):
        if fn[ie] == edges[fe[ie]].N[0]:
            fn[ie + 1] = edges[fe[ie]].N[1]
        else:
            fn[ie + 1] = edges[fe[ie]].N[0]
    pts = [nodes[k] for k in fn]
    mid = [0.0, 0.0, 0.0]
    for p in pts:
        mid = [mid[0] + p[0], mid[1] + p[1], mid[2] + p[2]]
    mid = [c / float(ne) for c in mid]
    nrm = math.sqrt(sum(c * c for c in mid))
    mid = [c / nrm for c in mid]
    nv = _cross(pts[ne - 1], pts[0])
    for i in range(ne - 1):
        c = _cross(pts[i], pts[i + 1])
        nv = (nv[0] + c[0], nv[1] + c[1], nv[2] + c[2])
    face = _Face(N=fn, E=fe, Orientation=1)
    # OrientFaceSphere(n, Mid) = dot(n, Mid); reverse when negative (NumE=4 > ChangeOrient=3)
    if nv[0] * mid[0] + nv[1] * mid[1] + nv[2] * mid[2] < 0:
        nt, et = list(fn), list(fe)
        face.N = [nt[ne - i - 1] for i in range(ne)]
        face.E = [et[ne - i - 2] for i in range(ne - 1)] + [et[ne - 1]]
        face.Orientation = 1
    return face


def _edge_direction(n1, n2, en):
    return 1 if n1 == en[0] else -1


def _shift(i, s):
    i_s = i + s
    if i_s > 4:
        i_s -= 4
    return i_s


def _orientation(edges, faces):
    """``Orientation!`` (Orientation.jl).

    ``Edges[iE].F`` is a 1-D vector in the reference, so ``size(F,2)`` is 1:
    every sweep stops after one opposite-edge step and ``iFEnd`` stays 0.
    That behaviour is reproduced literally.
    """
    num_edges = len(edges) - 1
    visit = [0] * (num_edges + 1)
    for ie0 in range(1, num_edges + 1):
        if visit[ie0] == 1:
            continue
        ie = ie0
        visit[ie] = 1
        if1 = edges[ie].F[0]
        ie1 = 4
        for k in range(1, 5):
            ie1 = k
            if ie == faces[if1].E[k - 1]:
                break
        # single pass of the while-loop (iLoop == 0)
        nf = faces[if1].N + [faces[if1].N[0]]
        o = _edge_direction(nf[ie1 - 1], nf[ie1], edges[ie].N)
        iep1 = _shift(ie1, 2)
        iep = faces[if1].E[iep1 - 1]
        if _edge_direction(nf[iep1 - 1], nf[iep1], edges[iep].N) == o:
            if o == -1:
                edges[iep].N = [nf[iep1 - 1], nf[iep1]]
            else:
                edges[iep].N = [nf[iep1], nf[iep1 - 1]]
        visit[iep] = 1


def _renumber_face4(face, edges):
    """``RenumberingFace4!`` (Renumbering.jl): rotate so N[1] starts two edges."""
    nn = len(face.N)
    i_n = nn
    for k in range(1, nn + 1):
        i_n = k
        node = face.N[k - 1]
        num = 0
        for ie in range(len(face.E)):
            if node == edges[face.E[ie]].N[0]:
                num += 1
            if num == 2:
                break
        if num == 2:
            break
    if i_n > 1:
        nt = face.N + face.N
        et = face.E + face.E
        face.N = [nt[i + i_n - 1] for i in range(nn)]
        face.E = [et[i + i_n - 1] for i in range(nn)]


@dataclass
class CubedSphere:
    """Single-rank cubed sphere after the reference's sub-grid renumbering."""

    ne: int
    rad: float
    nodes: np.ndarray  # (NumNodes, 3) coordinates, sub-grid order
    face_nodes: np.ndarray  # (NF, 4) 1-based sub-grid node numbers
    face_edges: np.ndarray  # (NF, 4) 1-based sub-grid edge numbers
    edge_nodes: np.ndarray  # (NumEdges, 2) 1-based sub-grid node numbers

    @property
    def num_faces(self):
        return self.face_nodes.shape[0]

    @property
    def num_nodes(self):
        return self.nodes.shape[0]

    @property
    def num_edges(self):
        return self.edge_nodes.shape[0]

    def face_corners(self):
        """Corner coordinates ``F[4,3,NF]`` as used by ``DiscretizationCG.jl:13-28``,
        returned as a (NF, 4, 3) array (P[i] is the point of node N[i])."""
        return self.nodes[self.face_nodes - 1]


def cubed_sphere(ne: int, rad: float = 6.37122e6) -> CubedSphere:
    """Build the cubed sphere with ``ne`` elements per panel edge (``nPanel``)."""
    n = int(ne)
    if n < 2:
        raise ValueError("cubed sphere needs ne >= 2")
    num_nodes = 6 * (n - 1) * (n - 1) + 12 * (n - 1) + 8
    nodes = [None] * (num_nodes + 1)
    cnt = [1]

    def add(p):
        nodes[cnt[0]] = p
        cnt[0] += 1

    start = {}
    # face interiors (CubedGrid.jl:15-81)
    start["W"] = cnt[0]
    for k in range(1, n):
        for j in range(1, n):
            add(_cube_point(-1, j, k, n, -1.0, 0.0, 0.0, rad))
    start["E"] = cnt[0]
    for k in range(1, n):
        for j in range(1, n):
            add(_cube_point(-1, j, k, n, 1.0, 0.0, 0.0, rad))
    start["S"] = cnt[0]
    for k in range(1, n):
        for i in range(1, n):
            add(_cube_point(i, -1, k, n, 0.0, -1.0, 0.0, rad))
    start["N"] = cnt[0]
    for k in range(1, n):
        for i in range(1, n):
            add(_cube_point(i, -1, k, n, 0.0, 1.0, 0.0, rad))
    start["B"] = cnt[0]
    for j in range(1, n):
        for i in range(1, n):
            add(_cube_point(i, j, -1, n, 0.0, 0.0, -1.0, rad))
    start["T"] = cnt[0]
    for j in range(1, n):
        for i in range(1, n):
            add(_cube_point(i, j, -1, n, 0.0, 0.0, 1.0, rad))
    # cube edges (CubedGrid.jl:82-184)
    for key, (y, z) in (("WEmm", (-1.0, -1.0)), ("WEpm", (1.0, -1.0)),
                        ("WEmp", (-1.0, 1.0)), ("WEpp", (1.0, 1.0))):
        start[key] = cnt[0]
        for i in range(1, n):
            add(_cube_point(i, -1, -1, n, 0.0, y, z, rad))
    for key, (x, z) in (("SNmm", (-1.0, -1.0)), ("SNpm", (1.0, -1.0)),
                        ("SNmp", (-1.0, 1.0)), ("SNpp", (1.0, 1.0))):
        start[key] = cnt[0]
        for j in range(1, n):
            add(_cube_point(-1, j, -1, n, x, 0.0, z, rad))
    for key, (x, y) in (("BTmm", (-1.0, -1.0)), ("BTpm", (1.0, -1.0)),
                        ("BTmp", (-1.0, 1.0)), ("BTpp", (1.0, 1.0))):
        start[key] = cnt[0]
        for k in range(1, n):
            add(_cube_point(-1, -1, k, n, x, y, 0.0, rad))
    # cube corners (CubedGrid.jl:185-234)
    for key, (x, y, z) in (("mmm", (-1, -1, -1)), ("pmm", (1, -1, -1)), ("mpm", (-1, 1, -1)),
                           ("ppm", (1, 1, -1)), ("mmp", (-1, -1, 1)), ("pmp", (1, -1, 1)),
                           ("mpp", (-1, 1, 1)), ("ppp", (1, 1, 1))):
        start[key] = cnt[0]
        add(_cube_point(-1, -1, -1, n, float(x), float(y), float(z), rad))
    assert cnt[0] == num_nodes + 1

    num_edges = 12 * (n - 1) * n + 12 * n
    edges = [None] * (num_edges + 1)
    ecnt = [1]

    def add_edge(n1, n2):
        edges[ecnt[0]] = _Edge(N=[n1, n2], E=ecnt[0])
        ecnt[0] += 1

    def insert_face_edge(ns, e1s1, e2s1, e1s2, e2s2):
        # InsertFaceEdge (CubedGrid.jl:393-441)
        node = ns
        ne1, ne2 = e1s1, e2s1
        s1 = ecnt[0]
        for _j in range(1, n):
            for i in range(1, n + 1):
                n1, n2 = node - 1, node
                if i == 1:
                    n1 = ne1
                    ne1 += 1
                if i == n:
                    n2 = ne2
                    ne2 += 1
                    node -= 1
                add_edge(n1, n2)
                node += 1
        node = ns
        ne1, ne2 = e1s2, e2s2
        s2 = ecnt[0]
        for j in range(1, n + 1):
            for _i in range(1, n):
                n1, n2 = node - (n - 1), node
                if j == 1:
                    n1 = ne1
                    ne1 += 1
                if j == n:
                    n2 = ne2
                    ne2 += 1
                add_edge(n1, n2)
                node += 1
        return s1, s2

    def insert_edge_edge(ns, e1, e2):
        # InsertEdgeEdge (CubedGrid.jl:443-462)
        node = ns
        s = ecnt[0]
        for i in range(1, n + 1):
            n1, n2 = node - 1, node
            if i == 1:
                n1 = e1
            if i == n:
                n2 = e2
            add_edge(n1, n2)
            node += 1
        return s

    S = start
    fe = {}
    fe["W"] = insert_face_edge(S["W"], S["BTmm"], S["BTmp"], S["SNmm"], S["SNmp"])
    fe["E"] = insert_face_edge(S["E"], S["BTpm"], S["BTpp"], S["SNpm"], S["SNpp"])
    fe["S"] = insert_face_edge(S["S"], S["BTmm"], S["BTpm"], S["WEmm"], S["WEmp"])
    fe["N"] = insert_face_edge(S["N"], S["BTmp"], S["BTpp"], S["WEpm"], S["WEpp"])
    fe["B"] = insert_face_edge(S["B"], S["SNmm"], S["SNpm"], S["WEmm"], S["WEpm"])
    fe["T"] = insert_face_edge(S["T"], S["SNmp"], S["SNpp"], S["WEmp"], S["WEpp"])
    ee = {}
    ee["WEmm"] = insert_edge_edge(S["WEmm"], S["mmm"], S["pmm"])
    ee["WEpm"] = insert_edge_edge(S["WEpm"], S["mpm"], S["ppm"])
    ee["WEmp"] = insert_edge_edge(S["WEmp"], S["mmp"], S["pmp"])
    ee["WEpp"] = insert_edge_edge(S["WEpp"], S["mpp"], S["ppp"])
    ee["SNmm"] = insert_edge_edge(S["SNmm"], S["mmm"], S["mpm"])
    ee["SNpm"] = insert_edge_edge(S["SNpm"], S["pmm"], S["ppm"])
    ee["SNmp"] = insert_edge_edge(S["SNmp"], S["mmp"], S["mpp"])
    ee["SNpp"] = insert_edge_edge(S["SNpp"], S["pmp"], S["ppp"])
    ee["BTmm"] = insert_edge_edge(S["BTmm"], S["mmm"], S["mmp"])
    ee["BTpm"] = insert_edge_edge(S["BTpm"], S["pmm"], S["pmp"])
    ee["BTmp"] = insert_edge_edge(S["BTmp"], S["mpm"], S["mpp"])
    ee["BTpp"] = insert_edge_edge(S["BTpp"], S["ppm"], S["ppp"])
    assert ecnt[0] == num_edges + 1

    num_faces = 6 * n * n
    faces = [None] * (num_faces + 1)
    fcnt = [1]

    def insert_face_face(s1, s2, ew1, ew2, sn1, sn2):
        # InsertFaceFace (CubedGrid.jl:464-513)
        en1, en2 = s1, s2
        for j in range(1, n + 1):
            for i in range(1, n + 1):
                e1, e2, e3, e4 = en1 - n, en1, en2 - 1, en2
                if j == 1:
                    e1 = ew1
                    ew1 += 1
                if j == n:
                    e2 = ew2
                    ew2 += 1
                if i == 1:
                    e3 = sn1
                    sn1 += 1
                if i == n:
                    e4 = sn2
                    sn2 += 1
                    en2 -= 1
                en1 += 1
                en2 += 1
                faces[fcnt[0]] = _make_face([e1, e2, e3, e4], nodes, edges, fcnt[0])
                fcnt[0] += 1

    insert_face_face(*fe["W"], ee["SNmm"], ee["SNmp"], ee["BTmm"], ee["BTmp"])
    insert_face_face(*fe["E"], ee["SNpm"], ee["SNpp"], ee["BTpm"], ee["BTpp"])
    insert_face_face(*fe["S"], ee["WEmm"], ee["WEmp"], ee["BTmm"], ee["BTpm"])
    insert_face_face(*fe["N"], ee["WEpm"], ee["WEpp"], ee["BTmp"], ee["BTpp"])
    insert_face_face(*fe["B"], ee["WEmm"], ee["WEpm"], ee["SNmm"], ee["SNpm"])
    insert_face_face(*fe["T"], ee["WEmp"], ee["WEpp"], ee["SNmp"], ee["SNpp"])
    assert fcnt[0] == num_faces + 1

    # order=true: Orientation! then Renumbering! on the global grid (CubedGrid.jl:329-332)
    _orientation(edges, faces)
    for iF in range(1, num_faces + 1):
        _renumber_face4(faces[iF], edges)

    # single-rank sub grid: first-appearance renumbering (SubGridGhost.jl:19-142)
    node_new = {}
    edge_new = {}
    node_order = []
    edge_order = []
    for iF in range(1, num_faces + 1):
        f = faces[iF]
        for i in range(len(f.N)):
            if f.E[i] not in edge_new:
                edge_new[f.E[i]] = len(edge_order) + 1
                edge_order.append(f.E[i])
            if f.N[i] not in node_new:
                node_new[f.N[i]] = len(node_order) + 1
                node_order.append(f.N[i])
    sub_edges = [None]
    for old in edge_order:
        e = edges[old]
        sub_edges.append(_Edge(N=[node_new[e.N[0]], node_new[e.N[1]]], E=len(sub_edges)))
    sub_faces = [None]
    for iF in range(1, num_faces + 1):
        f = faces[iF]
        sub_faces.append(_Face(N=[node_new[k] for k in f.N], E=[edge_new[k] for k in f.E],
                               Orientation=f.Orientation))
    # Renumbering!(Edges,Faces) on the sub grid (SubGridGhost.jl:158-162)
    for iF in range(1, num_faces + 1):
        _renumber_face4(sub_faces[iF], sub_edges)

    coords = np.array([nodes[k] for k in node_order], dtype=np.float64)
    face_nodes = np.array([sub_faces[i].N for i in range(1, num_faces + 1)], dtype=np.int64)
    face_edges = np.array([sub_faces[i].E for i in range(1, num_faces + 1)], dtype=np.int64)
    edge_nodes = np.array([sub_edges[i].N for i in range(1, len(sub_edges))], dtype=np.int64)
    return CubedSphere(ne=n, rad=rad, nodes=coords, face_nodes=face_nodes,
                       face_edges=face_edges, edge_nodes=edge_nodes)


def _permutation(p: int) -> np.ndarray:
    """``Per`` of ``NumberingFemCGQuad`` (NumberingFem.jl:22-69), 0-based."""
    tables = {
        1: [1, 2, 4, 3],
        2: [1, 5, 2, 8, 9, 6, 4, 7, 3],
        3: [1, 5, 6, 2, 12, 13, 14, 7, 11, 15, 16, 8, 4, 10, 9, 3],
        4: [1, 5, 6, 7, 2, 16, 17, 18, 19, 8, 15, 20, 21, 22, 9, 14, 23, 24, 25, 10,
            4, 13, 12, 11, 3],
    }
    if p not in tables:
        # the reference's p > 4 branch is MATLAB syntax and never runs
        raise ValueError("OrdPoly must be 1..4")
    return np.asarray(tables[p], dtype=np.int64) - 1


def numbering_cg_quad(grid: CubedSphere, p: int):
    """``NumberingFemCGQuad`` (NumberingFem.jl:21-131).

    Returns ``(glob, num_g)`` with ``glob`` of shape (NF, (p+1)^2), 1-based,
    i.e. the transpose view of the reference's column-major ``Glob[ID, IF]``.
    """
    per = _permutation(p)
    num_n = grid.num_nodes
    num_e = grid.num_edges * (p - 1)
    nf = grid.num_faces
    glob = np.zeros((nf, (p + 1) * (p + 1)), dtype=np.int64)
    loc = np.zeros((p + 1) * (p + 1), dtype=np.int64)
    en = grid.edge_nodes
    for iF in range(nf):
        ii = 0
        fnodes = grid.face_nodes[iF]
        for i in range(4):
            loc[ii] = fnodes[i]
            ii += 1
        for i in range(4):
            iE = grid.face_edges[iF, i]
            base = (iE - 1) * (p - 1) + num_n
            if en[iE - 1, 0] == fnodes[i]:
                for j in range(1, p):
                    loc[ii] = base + j
                    ii += 1
            else:
                for j in range(p - 1, 0, -1):
                    loc[ii] = base + j
                    ii += 1
        for j in range(1, p):
            for i in range(1, p):
                loc[ii] = num_n + num_e + i + (j - 1) * (p - 1) + iF * (p - 1) * (p - 1)
                ii += 1
        glob[iF] = loc[per]
    num_g = num_n + num_e + nf * (p - 1) * (p - 1)
    return glob, num_g
