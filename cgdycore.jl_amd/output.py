"""Output and restart: the VTK sphere output of ``unstructured_vtkSphere``
(src/Outputs/vtkSphere.jl:121-222,337-507) with its interpolation on the
device (``cgd_interpolate``, src/Outputs/vtkOutputKernel.jl:49-158), and a
binary restart file (the reference has none).

VTK: every element and layer is split into OrdPrint x OrdPrint hexahedral
sub-cells whose corners are mapped by ``TransSphereX`` (src/Grids/Trans.jl:58-75)
and whose cell values are the fields interpolated to the sub-cell centres.
The file is legacy VTK, binary (big-endian), readable by ParaView/VisIt.

Restart: ``CGDRST1\\n`` (8 bytes), the byte length of a UTF-8 JSON header as
a little-endian uint64, the header, then the state as little-endian float64
in the reference's Julia column-major layout (Nz, NumG, 5+NumTr), i.e. the
bytes of the device array.  A run resumed from a restart file continues bit
for bit (the step reads nothing else).
"""
from __future__ import annotations

import ctypes as C
import json
import struct

import numpy as np
import torch

from .host import CGDError, _ptr, _stream, native

_vp = C.c_void_p


# --------------------------------------------------------------------------
# interpolation to the sub-cell centres

def lagrange(x: float, xp: np.ndarray, ip: int) -> float:
    """``Lagrange(x,xP,iP)`` (src/DG/Lagrange.jl), 0-based ``ip``."""
    f = 1.0
    for i in range(len(xp)):
        if i != ip:
            f = f * (x - xp[i]) / (xp[ip] - xp[i])
    return f


def vtk_inter(xw: np.ndarray, ord_print: int) -> np.ndarray:
    """``vtkInter[iRef,jRef,i,j]`` (vtkSphere.jl:193-210): the Lagrange basis
    at the sub-cell centres; numpy (N, N, P, P) = Julia (P, P, N, N)."""
    n = len(xw)
    inter = np.zeros((n, n, ord_print, ord_print))  # [j, i, jRef, iRef]
    dd = 2.0 / ord_print
    eta0 = -1.0
    for jr in range(ord_print):
        ksi0 = -1.0
        eta1 = eta0 + dd
        for ir in range(ord_print):
            ksi1 = ksi0 + dd
            for j in range(n):
                for i in range(n):
                    inter[j, i, jr, ir] = inter[j, i, jr, ir] + lagrange(0.5 * (ksi0 + ksi1), xw, i) * \
                        lagrange(0.5 * (eta0 + eta1), xw, j)
            ksi0 = ksi1
        eta0 = eta1
    return inter


def _out_check(rc, what):
    if rc != 0:
        raise CGDError(f"{what}: {native().cgd_output_error().decode()} (code {rc})")


def _bind():
    L = native()
    if not getattr(L, "_out_bound", False):
        L.cgd_output_error.restype = C.c_char_p
        L.cgd_interpolate.argtypes = [C.c_int] * 6 + [_vp] * 8 + [_vp]
        L._out_bound = True
    return L


def _interp(mode, cCell, c, Inter, Glob, rho=None, u=None, v=None, dXdxI=None):
    L = _bind()
    P = cCell.shape[-1]
    nf, nz = cCell.shape[0], cCell.shape[1]
    n = Inter.shape[0]
    num_g = c.shape[0]
    _out_check(L.cgd_interpolate(mode, n, nz, nf, num_g, P, _ptr(Glob), _ptr(Inter), _ptr(c),
                                 _ptr(rho) if rho is not None else None, _ptr(u) if u is not None else None,
                                 _ptr(v) if v is not None else None, _ptr(dXdxI) if dXdxI is not None else None,
                                 _ptr(cCell), _stream()), "cgd_interpolate")


def InterpolateGPU(cCell, c, Inter, Glob):
    """``InterpolateGPU!`` (vtkOutputKernel.jl:215-233): cCell (NF, Nz, P, P) from c (NumG, Nz)."""
    _interp(0, cCell, c, Inter, Glob)


def InterpolateRhoGPU(cCell, c, Rho, Inter, Glob):
    """``InterpolateRhoGPU!`` (vtkOutputKernel.jl:254-272): c / rho at the sub-cell centres."""
    _interp(1, cCell, c, Inter, Glob, rho=Rho)


def InterpolateWBGPU(cCell, u, v, w, Inter, dXdxI, Glob):
    """``InterpolateWBGPU!`` (vtkOutputKernel.jl:274-292): the cell-centre w."""
    _interp(2, cCell, w, Inter, Glob, u=u, v=v, dXdxI=dXdxI)


# --------------------------------------------------------------------------
# points of the sub-cells

def vtk_points(X: np.ndarray, xw: np.ndarray, ord_print: int, rad: float, rad_print: float | None = None):
    """Corners of every sub-cell by ``TransSphereX`` (Trans.jl:58-75), in the
    order of ``vtkStruct`` (vtkSphere.jl:135-181): elements, layers, sub-cell
    rows and columns, and the hexahedron's 8 corners (bottom then top).
    ``X``: numpy (NF, Nz, 3, 2, N^2).  Returns (NF*Nz*P*P*8, 3)."""
    rp = rad if rad_print is None else rad_print
    nf, nz = X.shape[0], X.shape[1]
    n = len(xw)
    P = ord_print
    dd = 2.0 / P
    edges = [-1.0 + dd * k for k in range(P + 1)]
    Lx = np.array([[lagrange(x, xw, i) for i in range(n)] for x in edges])  # (P+1, N)
    Xg = X.reshape(nf, nz, 3, 2, n, n)  # [..., c, K, J, I]
    # XP at (edge ksi a, edge eta b, K): sum_ij L_i(ksi_a) L_j(eta_b) X[I=i, J=j, K]
    XP = np.einsum("ai,bj,fzckji->fzkbac", Lx, Lx, Xg)  # (NF, Nz, K, b, a, 3)
    r = np.linalg.norm(XP, axis=-1, keepdims=True)
    z = np.maximum(r - rad, 0.0)
    XP = XP / r * (rp + z)
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    pts = np.empty((nf, nz, P, P, 8, 3))
    for jr in range(P):
        for ir in range(P):
            for q, (di, dj, K) in enumerate(corners):
                pts[:, :, jr, ir, q] = XP[:, :, K, jr + dj, ir + di]
    return pts.reshape(-1, 3)


def write_vtk(path: str, points: np.ndarray, cell_data: dict, title: str = "CGDycore output"):
    """Legacy binary VTK unstructured grid of hexahedra (8 consecutive points
    per cell) with double cell data."""
    npts = points.shape[0]
    ncell = npts // 8
    with open(path, "wb") as f:
        f.write(f"# vtk DataFile Version 3.0\n{title}\nBINARY\nDATASET UNSTRUCTURED_GRID\n".encode())
        f.write(f"POINTS {npts} double\n".encode())
        f.write(np.ascontiguousarray(points, dtype=">f8").tobytes())
        f.write(f"\nCELLS {ncell} {9 * ncell}\n".encode())
        conn = np.empty((ncell, 9), dtype=">i4")
        conn[:, 0] = 8
        conn[:, 1:] = np.arange(npts, dtype=np.int64).reshape(ncell, 8)
        f.write(conn.tobytes())
        f.write(f"\nCELL_TYPES {ncell}\n".encode())
        f.write(np.full(ncell, 12, dtype=">i4").tobytes())  # VTK_HEXAHEDRON
        f.write(f"\nCELL_DATA {ncell}\n".encode())
        for name, val in cell_data.items():
            val = np.asarray(val, dtype=">f8").reshape(-1)
            if val.size != ncell:
                raise CGDError(f"write_vtk: field {name} has {val.size} values for {ncell} cells")
            f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n".encode())
            f.write(val.tobytes())
            f.write(b"\n")


def read_vtk_header(path: str) -> dict:
    """Counts and field names of a file written by ``write_vtk`` (checks the layout)."""
    data = open(path, "rb").read()
    out = {"fields": []}
    i = 0

    def line():
        nonlocal i
        j = data.index(b"\n", i)
        s = data[i:j].decode()
        i = j + 1
        return s

    assert line().startswith("# vtk DataFile")
    line()
    assert line() == "BINARY"
    assert line() == "DATASET UNSTRUCTURED_GRID"
    _, npts, _ = line().split()
    out["points"] = int(npts)
    i += 24 * int(npts) + 1
    _, ncell, nint = line().split()
    out["cells"] = int(ncell)
    conn = np.frombuffer(data, dtype=">i4", count=int(nint), offset=i).reshape(-1, 9)
    out["connectivity_ok"] = bool((conn[:, 0] == 8).all() and conn[-1, -1] == int(npts) - 1)
    i += 4 * int(nint) + 1
    line()
    i += 4 * int(ncell) + 1
    assert line().split()[0] == "CELL_DATA"
    while i < len(data):
        name = line().split()[1]
        line()
        vals = np.frombuffer(data, dtype=">f8", count=int(ncell), offset=i)
        out["fields"].append(name)
        out[name] = vals.astype(np.float64)
        i += 8 * int(ncell) + 1
    return out


def unstructured_vtkSphere(path, U, FE, Metric, Global, xw, rad, ord_print=4, points=None):
    """``unstructured_vtkSphere`` (vtkSphere.jl:337-507) for the fields rho,
    u, v, w (cell centre, ``wB``), Th (rho*theta / rho) and the tracer mixing
    ratios, interpolated on the device.  ``U`` is the device state
    (5+NumTr, NumG, Nz); ``points`` may be passed from an earlier call."""
    nf, nz = Global.NF, Global.Nz
    dev = U.device
    Inter = torch.as_tensor(vtk_inter(np.asarray(xw), ord_print), dtype=torch.float64, device=dev)
    cCell = torch.empty((nf, nz, ord_print, ord_print), dtype=torch.float64, device=dev)
    fields = {}

    def take(name):
        torch.cuda.synchronize()
        fields[name] = cCell.cpu().numpy().copy()

    InterpolateGPU(cCell, U[0], Inter, FE.Glob)
    take("rho")
    InterpolateGPU(cCell, U[1], Inter, FE.Glob)
    take("u")
    InterpolateGPU(cCell, U[2], Inter, FE.Glob)
    take("v")
    InterpolateWBGPU(cCell, U[1], U[2], U[3], Inter, Metric.dXdxI, FE.Glob)
    take("w")
    InterpolateRhoGPU(cCell, U[4], U[0], Inter, FE.Glob)
    take("Th")
    for t in range(U.shape[0] - 5):
        InterpolateRhoGPU(cCell, U[5 + t], U[0], Inter, FE.Glob)
        take(f"Tr{t + 1}")
    if points is None:
        points = vtk_points(Metric.X.cpu().numpy(), np.asarray(xw), ord_print, rad)
    write_vtk(path, points, fields)
    return points, fields


# --------------------------------------------------------------------------
# restart

_MAGIC = b"CGDRST1\n"


def write_restart(path: str, V: torch.Tensor, Global, step: int = 0, time: float = 0.0, extra: dict | None = None):
    """Write the state ``V`` (device or host, (5+NumTr, NumG, Nz) float64)."""
    a = V.detach().cpu().numpy() if isinstance(V, torch.Tensor) else np.asarray(V)
    if a.dtype != np.float64 or a.ndim != 3 or a.shape[1:] != (Global.NumG, Global.Nz):
        raise CGDError(f"write_restart: expected float64 (nv, {Global.NumG}, {Global.Nz}), got {a.dtype} {a.shape}")
    head = {"format": "CGDRST1", "layout": "Julia column-major (Nz, NumG, NumV+NumTr), float64 little-endian",
            "Nz": Global.Nz, "NumG": Global.NumG, "NF": Global.NF, "N": Global.N, "NumV": 5,
            "NumTr": int(a.shape[0] - 5), "step": int(step), "time": float(time),
            "precision": Global.Precision}
    if extra:
        head["extra"] = extra
    hb = json.dumps(head).encode()
    with open(path, "wb") as f:
        f.write(_MAGIC)
        f.write(struct.pack("<Q", len(hb)))
        f.write(hb)
        f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def read_restart(path: str, Global=None, device=None):
    """Read a restart file: (state, header).  With ``Global`` the grid
    dimensions are checked; with ``device`` the state is returned there."""
    with open(path, "rb") as f:
        if f.read(8) != _MAGIC:
            raise CGDError(f"read_restart: {path} is not a restart file")
        (hl,) = struct.unpack("<Q", f.read(8))
        head = json.loads(f.read(hl).decode())
        nv = head["NumV"] + head["NumTr"]
        shape = (nv, head["NumG"], head["Nz"])
        a = np.fromfile(f, dtype="<f8", count=int(np.prod(shape)))
    if a.size != np.prod(shape):
        raise CGDError(f"read_restart: {path} is truncated")
    if Global is not None:
        want = (Global.Nz, Global.NumG, Global.NF, Global.N)
        got = (head["Nz"], head["NumG"], head["NF"], head["N"])
        if want != got:
            raise CGDError(f"read_restart: grid (Nz, NumG, NF, N) = {got} does not match {want}")
    a = a.reshape(shape).astype(np.float64)
    if device is not None:
        return torch.as_tensor(a, device=device), head
    return a, head
