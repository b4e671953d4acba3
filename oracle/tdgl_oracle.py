"""CPU FP64 oracle for tdgl3d's trapezoidal time step.

TEST INFRASTRUCTURE ONLY.  This module is imported by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
and only as the checker (or the timed CPU baseline).  The library
(``nanowire_tdgl_amd``) never imports it and never falls back to it.

It restates, in plain numpy/scipy, the algorithm of the reference package
``tdgl3d`` (omedeiro/nanowire_tdgl); every function cites the reference
``file:line`` it follows.  The right-hand side exists twice:

* :func:`eval_f_csr` keeps the reference's cost model (ten scipy CSR operators
  rebuilt per call, row-sliced, mat-vec'd); it is the CPU baseline;
* :func:`eval_f_stencil` evaluates the same operator as a vectorised 3-D
  stencil; the two cross-check each other.

Pinning (what checks each function; see ``tests/test_oracle.py``):

* indices / holes / material  -- golden fixtures written by tdgl3d itself
  (``tests/golden/indices_*.npz``), plus brute-force and hand-derived counts;
* apply_bc + eval_f           -- golden ``evalf_*.npz`` (applied field, holes,
  trilayer, 2-D, periodic), closed forms (FPSI values, zero-gauge Laplacian,
  LPHI diagonal, uniform periodic equilibrium, C4 antisymmetry), and the
  CSR-vs-stencil cross-check;
* jvp_stencil                 -- central differences of eval_f_csr on the ten
  golden eval_f inputs (second-order convergence to it), hole rows zero;
* tgcr / newton / trapezoidal -- golden ``tgcr_trap``/``newton_trap``/``trap_*``
  (values and iteration counts), dense linear solves, scalar Newton problems;
* forward_euler, observables   -- golden ``euler_small`` and ``obs_aniso``.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
import scipy.sparse as sp

TAU_RELAX = 0.1  # operators/sparse_operators.py:223


# ---------------------------------------------------------------------------
# Grid parameters -- core/parameters.py:9-94
# ---------------------------------------------------------------------------

@dataclass
class Grid:
    Nx: int
    Ny: int
    Nz: int
    hx: float = 1.0
    hy: float = 1.0
    hz: float = 1.0
    kappa: float = 5.0
    periodic_x: bool = False
    periodic_y: bool = False
    periodic_z: bool = False

    @property
    def is_3d(self) -> bool:  # parameters.py:54-55
        return self.Nz > 1

    @property
    def n_interior(self) -> int:  # parameters.py:58-60
        return (self.Nx - 1) * (self.Ny - 1) * max(self.Nz - 1, 1)

    @property
    def n_state(self) -> int:  # parameters.py:69-72
        return (4 if self.is_3d else 3) * self.n_interior

    @property
    def dim_x(self) -> int:  # parameters.py:75-79
        nzp = self.Nz + 1 if self.is_3d else 1
        return (self.Nx + 1) * (self.Ny + 1) * nzp

    @property
    def mj(self) -> int:  # parameters.py:82-84
        return self.Nx + 1

    @property
    def mk(self) -> int:  # parameters.py:86-89
        return (self.Nx + 1) * (self.Ny + 1)

    @property
    def full_shape(self) -> tuple[int, int, int]:
        """(k, j, i) shape of the full grid; C order gives m = i + mj*j + mk*k."""
        return ((self.Nz + 1) if self.is_3d else 1, self.Ny + 1, self.Nx + 1)


def grid_from(p) -> Grid:
    """Copy any object with SimulationParameters' fields into a :class:`Grid`."""
    return Grid(int(p.Nx), int(p.Ny), int(p.Nz), float(p.hx), float(p.hy), float(p.hz),
                float(p.kappa), bool(p.periodic_x), bool(p.periodic_y), bool(p.periodic_z))


# ---------------------------------------------------------------------------
# Index maps -- mesh/indices.py:239-450
# ---------------------------------------------------------------------------

def _lin(i, j, k, g: Grid):
    return i + g.mj * j + g.mk * k  # indices.py:46-54


def _plane(fixed_axis: str, value: int, r1, r2, g: Grid) -> np.ndarray:
    """Linear indices of the plane ``axis == value`` over the meshgrid(r1, r2, 'ij')."""
    a, b = np.meshgrid(np.asarray(r1, np.intp), np.asarray(r2, np.intp), indexing="ij")
    a, b = a.ravel(), b.ravel()
    v = np.full(a.size, value, dtype=np.intp)
    if fixed_axis == "x":      # others (j, k)   indices.py:319-328
        return _lin(v, a, b, g)
    if fixed_axis == "y":      # others (i, k)   indices.py:333-341
        return _lin(a, v, b, g)
    return _lin(a, b, v, g)    # others (i, j)   indices.py:349-357


def construct_indices(g: Grid) -> dict:
    """All index arrays of ``construct_indices`` (indices.py:239-450), same order."""
    Nx, Ny, Nz = g.Nx, g.Ny, g.Nz
    all_i = np.arange(Nx + 1, dtype=np.intp)
    all_j = np.arange(Ny + 1, dtype=np.intp)
    all_k = np.arange(Nz + 1, dtype=np.intp) if g.is_3d else np.array([0], np.intp)
    int_i = np.arange(1, Nx, dtype=np.intp)
    int_j = np.arange(1, Ny, dtype=np.intp)
    int_k = np.arange(1, Nz, dtype=np.intp) if g.is_3d else np.array([0], np.intp)

    # interior order: i slowest, k fastest  (indices.py:271-278)
    ii, jj, kk = np.meshgrid(int_i, int_j, int_k, indexing="ij")
    out = {"interior_to_full": _lin(ii.ravel(), jj.ravel(), kk.ravel(), g)}

    # bfield_interior assumes an i-fastest interior numbering (indices.py:283-300)
    nxi, nyi = Nx - 1, Ny - 1
    bi, bj = np.arange(Nx - 2, dtype=np.intp), np.arange(Ny - 2, dtype=np.intp)
    if g.is_3d:
        bk = np.arange(Nz - 2, dtype=np.intp)
        a, b, c = np.meshgrid(bi, bj, bk, indexing="ij")
        out["bfield_interior"] = a.ravel() + nxi * b.ravel() + nxi * nyi * c.ravel()
    else:
        a, b = np.meshgrid(bi, bj, indexing="ij")
        out["bfield_interior"] = a.ravel() + nxi * b.ravel()

    spec = {"x": (Nx, all_j, all_k, int_j, int_k),
            "y": (Ny, all_i, all_k, int_i, int_k),
            "z": (Nz, all_i, all_j, int_i, int_j)}
    empty = np.array([], dtype=np.intp)
    for ax, (N, o1, o2, n1, n2) in spec.items():  # indices.py:305-387
        if ax == "z" and not g.is_3d:
            for nm in ("face_lo", "first", "last", "face_hi"):
                out[f"z_{nm}"] = empty
                out[f"z_{nm}_inner"] = empty
            continue
        for nm, v in (("face_lo", 0), ("first", 1), ("last", N - 1), ("face_hi", N)):
            out[f"{ax}_{nm}"] = _plane(ax, v, o1, o2, g)
            out[f"{ax}_{nm}_inner"] = _plane(ax, v, n1, n2, g)
    # normal-component masks: planes {0, N-1} over all other indices (indices.py:392-430)
    for ax, (N, o1, o2, _, _) in spec.items():
        if ax == "z" and not g.is_3d:
            out["z_normal_bc_mask"] = empty
            continue
        out[f"{ax}_normal_bc_mask"] = np.concatenate(
            [_plane(ax, 0, o1, o2, g), _plane(ax, N - 1, o1, o2, g)])
    for ax in "xyz":
        out[f"hole_{ax}_bc_mask"] = empty.copy()
        out[f"hole_{ax}_bc_interior"] = empty.copy()
    return out


# ---------------------------------------------------------------------------
# Holes -- mesh/holes.py:43-257, mesh/indices.py:136-232
# ---------------------------------------------------------------------------

def point_in_polygon(x: float, y: float, verts) -> bool:
    """Even-odd ray cast toward +x with the reference's half-open tests (holes.py:43-100)."""
    inside = False
    n = len(verts)
    ax, ay = verts[0]
    xint = None
    for s in range(1, n + 1):
        bx, by = verts[s % n]
        if min(ay, by) < y <= max(ay, by) and x <= max(ax, bx):
            if ay != by:
                xint = (y - ay) * (bx - ax) / (by - ay) + ax
            # the reference reuses a stale crossing when ay == by (holes.py:92-96)
            if ax == bx or x <= xint:
                inside = not inside
        ax, ay = bx, by
    return inside


def identify_hole_nodes(verts, z_range, hx, hy, Nx, Ny, Nz) -> np.ndarray:
    """Boolean (Nx+1, Ny+1, Nz+1) node mask, extruded over z (holes.py:103-157)."""
    mask = np.zeros((Nx + 1, Ny + 1, Nz + 1), dtype=bool)
    k0, k1 = z_range
    ks = list(range(k0, min(k1 + 1, Nz + 1)))
    for i in range(Nx + 1):
        for j in range(Ny + 1):
            if point_in_polygon(i * hx, j * hy, verts):
                for k in ks:
                    mask[i, j, k] = True
    return mask


def identify_boundary_links(mask: np.ndarray, direction: str, is_3d: bool = True) -> np.ndarray:
    """Links whose two end nodes differ in the mask; loop orders fix the output
    order: x -> (k, j, i), y -> (k, i, j), z -> (i, j, k)  (holes.py:160-257)."""
    Nx, Ny, Nz = (s - 1 for s in mask.shape)
    mj, mk = Nx + 1, (Nx + 1) * (Ny + 1)

    def lin(i, j, k):
        return i + mj * j + (mk * k if is_3d else 0)

    links = []
    if direction == "x":
        for k in range(Nz + 1):
            for j in range(Ny + 1):
                for i in range(Nx):
                    if mask[i, j, k] != mask[i + 1, j, k]:
                        links.append(lin(i, j, k))
    elif direction == "y":
        for k in range(Nz + 1):
            for i in range(Nx + 1):
                for j in range(Ny):
                    if mask[i, j, k] != mask[i, j + 1, k]:
                        links.append(lin(i, j, k))
    elif direction == "z":
        for i in range(Nx + 1):
            for j in range(Ny + 1):
                for k in range(Nz):
                    if mask[i, j, k] != mask[i, j, k + 1]:
                        links.append(lin(i, j, k))
    else:
        raise ValueError(f"Invalid direction '{direction}'. Use 'x', 'y', or 'z'.")
    return np.array(links, dtype=np.int64)


def define_hole_polygon(idx: dict, verts, z_range, g: Grid) -> None:
    """Append the hole's boundary links (full and interior numbering) to idx
    (indices.py:136-232)."""
    mask = identify_hole_nodes(verts, z_range, g.hx, g.hy, g.Nx, g.Ny, g.Nz)
    f2i = np.full(g.dim_x, -1, dtype=np.intp)
    f2i[idx["interior_to_full"]] = np.arange(g.n_interior, dtype=np.intp)
    for ax in "xyz":
        new = identify_boundary_links(mask, ax, is_3d=g.is_3d)
        idx[f"hole_{ax}_bc_mask"] = np.concatenate([idx[f"hole_{ax}_bc_mask"], new])
        inter = [f2i[m] for m in new if f2i[m] >= 0]
        if inter:
            idx[f"hole_{ax}_bc_interior"] = np.concatenate(
                [idx[f"hole_{ax}_bc_interior"], np.array(inter, dtype=np.intp)])


# ---------------------------------------------------------------------------
# Material -- core/material.py:26-254, core/device.py:80-153
# ---------------------------------------------------------------------------

@dataclass
class Material:
    kappa: np.ndarray             # float64 [dim_x]
    sc_mask: np.ndarray           # float64 [dim_x]
    interior_sc_mask: np.ndarray  # float64 [n_interior]


def uniform_material(g: Grid) -> Material:
    """The map Device.add_hole creates when none exists (device.py:146-153)."""
    return Material(np.full(g.dim_x, g.kappa), np.ones(g.dim_x), np.ones(g.n_interior))


def carve_hole(mat: Material, verts, z_range, g: Grid, idx: dict) -> None:
    """sc_mask := 0 on hole nodes; interior copy refreshed (material.py:43-112)."""
    mask = identify_hole_nodes(verts, z_range, g.hx, g.hy, g.Nx, g.Ny, g.Nz)
    ii, jj, kk = np.where(mask)
    lin = ii + g.mj * jj + (g.mk * kk if g.is_3d else 0)
    mat.sc_mask[lin] = 0.0
    mat.interior_sc_mask = mat.sc_mask[idx["interior_to_full"]]


def trilayer_material(g: Grid, layers, idx: dict) -> Material:
    """Per-z-plane kappa/sc for a (bottom, insulator, top) stack; layers are
    (thickness, kappa, is_sc) triples (material.py:182-254)."""
    (tb, kb, sb), (ti, ki, si), (tt, kt, st) = layers
    ranges = [(0, tb, kb, sb), (tb, tb + ti, ki, si), (tb + ti, tb + ti + tt, kt, st)]
    kap = np.empty(g.dim_x)
    sc = np.empty(g.dim_x)
    for k in range(g.Nz + 1):
        sel = (kt, st)  # the top plane k == Nz falls through to "top"
        for lo, hi, kv, sv in ranges:
            if lo <= k < hi:
                sel = (kv, sv)
                break
        kap[k * g.mk:(k + 1) * g.mk] = sel[0]
        sc[k * g.mk:(k + 1) * g.mk] = 1.0 if sel[1] else 0.0
    return Material(kap, sc, sc[idx["interior_to_full"]])


class Film:
    """Grid + index maps + optional material, the parts of Device the RHS reads."""

    def __init__(self, g: Grid, layers=None):
        self.g = g
        self.idx = construct_indices(g)
        self.material: Optional[Material] = (
            trilayer_material(g, layers, self.idx) if layers is not None else None)

    def add_hole(self, verts, z_range=None) -> None:  # device.py:80-153
        if z_range is None:
            z_range = (0, self.g.Nz)
        define_hole_polygon(self.idx, verts, z_range, self.g)
        if self.material is None:
            self.material = uniform_material(self.g)
        carve_hole(self.material, verts, z_range, self.g, self.idx)


# ---------------------------------------------------------------------------
# Applied field -- physics/applied_field.py:42-91
# ---------------------------------------------------------------------------

def applied_field_at(t, t_stop, B, t_on_fraction=2.0 / 3.0, ramp=False, ramp_fraction=0.5):
    """(Bx, By, Bz) at time t: zero for t <= 0 and after t_on_fraction*t_stop
    (applied_field.py:42-56)."""
    if t <= 0 or t > t_stop * t_on_fraction:
        return 0.0, 0.0, 0.0
    scale = 1.0
    if ramp:
        tr = t_stop * ramp_fraction
        scale = min(t / tr, 1.0) if tr > 0 else 1.0
    return B[0] * scale, B[1] * scale, B[2] * scale


def boundary_vectors(bx, by, bz, g: Grid, idx: dict):
    """Full-grid Bx/By/Bz vectors, non-zero on the face-inner nodes
    (applied_field.py:59-91)."""
    out = []
    faces = {"x": ("y", "z"), "y": ("z", "x"), "z": ("x", "y")}
    for comp, val in (("x", bx), ("y", by), ("z", bz)):
        v = np.zeros(g.dim_x)
        if val != 0.0:
            a, b = faces[comp]
            sel = np.concatenate([idx[f"{a}_face_lo_inner"], idx[f"{b}_face_lo_inner"],
                                  idx[f"{a}_face_hi_inner"], idx[f"{b}_face_hi_inner"]])
            np.add.at(v, sel, val)
        out.append(v)
    return tuple(out)


# ---------------------------------------------------------------------------
# Boundary conditions -- physics/rhs.py:43-139
# ---------------------------------------------------------------------------

def expand(vals, g: Grid, idx: dict) -> np.ndarray:
    full = np.zeros(g.dim_x, dtype=np.complex128)  # rhs.py:43-51
    full[idx["interior_to_full"]] = vals
    return full


def apply_bc(x, y1, y2, y3, g: Grid, idx: dict, u) -> None:
    """In-place ghost fill, in the reference's order (rhs.py:54-139):
    normal-link zeroing, snapshot, per-axis ghost writes, hole-link zeroing."""
    Bx, By, Bz = u
    hx, hy, hz = g.hx, g.hy, g.hz
    y1[idx["x_normal_bc_mask"]] = 0.0
    y2[idx["y_normal_bc_mask"]] = 0.0
    if g.is_3d:
        y3[idx["z_normal_bc_mask"]] = 0.0
    x0, y10, y20, y30 = x.copy(), y1.copy(), y2.copy(), y3.copy()

    def I(name):
        return idx[name]

    if g.periodic_x:
        x[I("x_face_lo_inner")] += x0[I("x_last_inner")]
        x[I("x_face_hi_inner")] += x0[I("x_first_inner")]
        y1[I("x_face_lo_inner")] += y10[I("x_last_inner")]
        y1[I("x_face_hi_inner")] += y10[I("x_first_inner")]
    else:
        lo, hi, fi, la = I("x_face_lo_inner"), I("x_face_hi_inner"), I("x_first_inner"), I("x_last_inner")
        x[lo] += x0[fi] * np.exp(-1j * y10[lo])
        x[hi] += x0[la] * np.exp(1j * y10[la])
        y2[lo] += -Bz[lo] * hx * hy + y20[fi]
        y2[hi] += Bz[hi] * hx * hy + y20[la]
        y3[lo] += By[lo] * hz * hx + y30[fi]
        y3[hi] += -By[hi] * hz * hx + y30[la]
    if g.periodic_y:
        x[I("y_face_lo_inner")] += x0[I("y_last_inner")]
        x[I("y_face_hi_inner")] += x0[I("y_first_inner")]
        y2[I("y_face_lo_inner")] += y20[I("y_last_inner")]
        y2[I("y_face_hi_inner")] += y20[I("y_first_inner")]
    else:
        lo, hi, fi, la = I("y_face_lo_inner"), I("y_face_hi_inner"), I("y_first_inner"), I("y_last_inner")
        x[lo] += x0[fi] * np.exp(-1j * y20[lo])
        x[hi] += x0[la] * np.exp(1j * y20[la])
        y1[lo] += Bz[lo] * hx * hy + y10[fi]
        y1[hi] += -Bz[hi] * hx * hy + y10[la]
        y3[lo] += -Bx[lo] * hy * hz + y30[fi]
        y3[hi] += Bx[hi] * hy * hz + y30[la]
    if g.is_3d:
        if g.periodic_z:
            x[I("z_face_lo_inner")] += x0[I("z_last_inner")]
            x[I("z_face_hi_inner")] += x0[I("z_first_inner")]
            y3[I("z_face_lo_inner")] += y30[I("z_last_inner")]
            y3[I("z_face_hi_inner")] += y30[I("z_first_inner")]
        else:
            lo, hi, fi, la = I("z_face_lo_inner"), I("z_face_hi_inner"), I("z_first_inner"), I("z_last_inner")
            x[lo] += x0[fi] * np.exp(-1j * y30[lo])
            x[hi] += x0[la] * np.exp(1j * y30[la])
            y1[lo] += -By[lo] * hz * hx + y10[fi]
            y1[hi] += By[hi] * hz * hx + y10[la]
            y2[lo] += Bx[lo] * hy * hz + y20[fi]
            y2[hi] += -Bx[hi] * hy * hz + y20[la]
    if idx["hole_x_bc_mask"].size:
        y1[idx["hole_x_bc_mask"]] = 0.0
    if idx["hole_y_bc_mask"].size:
        y2[idx["hole_y_bc_mask"]] = 0.0
    if g.is_3d and idx["hole_z_bc_mask"].size:
        y3[idx["hole_z_bc_mask"]] = 0.0


def expand_with_bc(X, g: Grid, idx: dict, u):
    n = g.n_interior
    x = expand(X[:n], g, idx)
    y1 = expand(X[n:2 * n], g, idx)
    y2 = expand(X[2 * n:3 * n], g, idx)
    y3 = expand(X[3 * n:4 * n] if g.is_3d else np.zeros(n, np.complex128), g, idx)
    apply_bc(x, y1, y2, y3, g, idx, u)
    return x, y1, y2, y3


# ---------------------------------------------------------------------------
# RHS, CSR form -- physics/rhs.py:147-239, operators/sparse_operators.py:19-317
# ---------------------------------------------------------------------------

def _kappa_rows(g: Grid, idx: dict, mat: Optional[Material]) -> np.ndarray:
    m = idx["interior_to_full"]  # sparse_operators.py:19-28
    return mat.kappa[m] if mat is not None else np.full(m.size, g.kappa)


def _csr(vals, rows, cols, N):
    return sp.csr_matrix((np.asarray(vals, np.complex128), (rows, cols)), shape=(N, N),
                         dtype=np.complex128)


def _lpsi(y, axis: str, g: Grid, idx: dict):
    """Gauge-covariant 1-D Laplacian rows (sparse_operators.py:35-101)."""
    N, m = g.dim_x, idx["interior_to_full"]
    if axis == "z" and not g.is_3d:
        return sp.csr_matrix((N, N), dtype=np.complex128)
    s = {"x": 1, "y": g.mj, "z": g.mk}[axis]
    L = _csr(np.full(m.size, -2.0), m, m, N)
    L = L + _csr(np.exp(-1j * y[m - s]), m, m - s, N)
    L = L + _csr(np.exp(1j * y[m]), m, m + s, N)
    return L


def _lphi(axis: str, g: Grid, idx: dict, mat):
    """Link Laplacian over the two directions transverse to `axis`
    (sparse_operators.py:108-195)."""
    N, m = g.dim_x, idx["interior_to_full"]
    k2 = _kappa_rows(g, idx, mat) ** 2
    c = {"x": k2 / g.hx ** 2, "y": k2 / g.hy ** 2,
         "z": k2 / g.hz ** 2 if g.is_3d else np.zeros_like(k2)}
    if axis == "z":
        c["z"] = np.zeros_like(k2)  # LPHI_z has no z part
    trans = [a for a in "xyz" if a != axis]
    if axis in ("x", "y") and not g.is_3d:
        trans = [a for a in trans if a != "z"]
    diag = -2.0 * sum(c[a] for a in [t for t in "xyz" if t != axis])
    L = _csr(diag, m, m, N)
    for a in trans:
        s = {"x": 1, "y": g.mj, "z": g.mk}[a]
        L = L + _csr(c[a], m, m + s, N)
        L = L + _csr(c[a], m, m - s, N)
    return L


def _fpsi(x, g: Grid, idx: dict, mat):
    pm = x[idx["interior_to_full"]]  # sparse_operators.py:202-226
    gl = (1.0 - np.conj(pm) * pm) * pm
    if mat is None:
        return gl
    sc = mat.interior_sc_mask
    return sc * gl - (1.0 - sc) * pm / TAU_RELAX


def _fphi(axis: str, x, ys, g: Grid, idx: dict, mat):
    """Curl cross terms + supercurrent for one link component
    (sparse_operators.py:229-317)."""
    m = idx["interior_to_full"]
    k2 = _kappa_rows(g, idx, mat) ** 2
    s = {"x": 1, "y": g.mj, "z": g.mk}
    h = {"x": g.hx, "y": g.hy, "z": g.hz}
    yv = {"x": ys[0], "y": ys[1], "z": ys[2]}
    if axis == "z" and not g.is_3d:
        return np.zeros(g.n_interior, dtype=np.complex128)
    a = s[axis]
    sup = np.imag(np.exp(-1j * yv[axis][m]) * np.conj(x[m]) * x[m + a])
    others = [b for b in "xyz" if b != axis and (g.is_3d or b != "z")]
    curl = 0.0
    for b in others:
        yb, sb = yv[b], s[b]
        curl = curl + (k2 / h[b] ** 2) * (-yb[m + a] + yb[m] + yb[m + a - sb] - yb[m - sb])
    return (curl + sup).astype(np.complex128)


def eval_f_csr(X, g: Grid, idx: dict, u, mat: Optional[Material] = None) -> np.ndarray:
    """tdgl3d's eval_f with its cost model: operators rebuilt as CSR per call
    (rhs.py:147-239)."""
    x, y1, y2, y3 = expand_with_bc(np.asarray(X, np.complex128), g, idx, u)
    m = idx["interior_to_full"]
    LX, LY, LZ = (_lpsi(y, a, g, idx) for y, a in ((y1, "x"), (y2, "y"), (y3, "z")))
    PX, PY, PZ = (_lphi(a, g, idx, mat) for a in "xyz")
    LX, LY, LZ, PX, PY, PZ = (M[m, :] for M in (LX, LY, LZ, PX, PY, PZ))
    dpsi = (LX / g.hx ** 2 + LY / g.hy ** 2 + LZ / g.hz ** 2) @ x + _fpsi(x, g, idx, mat)
    ys = (y1, y2, y3)
    dpx = (PY + PZ) @ y1 + _fphi("x", x, ys, g, idx, mat)
    dpy = (PX + PZ) @ y2 + _fphi("y", x, ys, g, idx, mat)
    dpz = (PX + PY) @ y3 + _fphi("z", x, ys, g, idx, mat)
    for arr, ax in ((dpx, "x"), (dpy, "y"), (dpz, "z")):
        hm = idx[f"hole_{ax}_bc_interior"]
        if hm.size and (ax != "z" or g.is_3d):
            arr[hm] = 0.0
    blocks = [dpsi, dpx, dpy] + ([dpz] if g.is_3d else [])
    return np.concatenate(blocks)


# ---------------------------------------------------------------------------
# RHS, stencil form (same operator, independent formulation)
# ---------------------------------------------------------------------------

def eval_f_stencil(X, g: Grid, idx: dict, u, mat: Optional[Material] = None) -> np.ndarray:
    """The operator of rhs.py:219-234 written as array shifts on the (k, j, i) grid.

    dpsi = sum_a (U_a^- psi(-a) - 2 psi + U_a^+ psi(+a)) / h_a^2 + FPSI
    dphi_a = k^2 [2 D_aa/h_a^2 + sum_{b!=a} D_bb/h_b^2] phi_a + curl cross terms
             + Im(e^{-i phi_a} conj(psi) psi(+a))
    """
    x, y1, y2, y3 = (v.reshape(g.full_shape) for v in expand_with_bc(
        np.asarray(X, np.complex128), g, idx, u))
    nz3 = g.is_3d
    Ki = slice(1, g.Nz) if nz3 else slice(0, 1)
    Ji, Ii = slice(1, g.Ny), slice(1, g.Nx)

    def sh(A, di=0, dj=0, dk=0):
        ks = slice(Ki.start + dk, Ki.stop + dk)
        return A[ks, Ji.start + dj:Ji.stop + dj, Ii.start + di:Ii.stop + di]

    c = sh(x)
    lap = 0.0
    axes = [("x", (1, 0, 0), y1, g.hx), ("y", (0, 1, 0), y2, g.hy)]
    if nz3:
        axes.append(("z", (0, 0, 1), y3, g.hz))
    for _, (di, dj, dk), y, h in axes:
        lap = lap + (np.exp(-1j * sh(y, -di, -dj, -dk)) * sh(x, -di, -dj, -dk) - 2.0 * c
                     + np.exp(1j * sh(y)) * sh(x, di, dj, dk)) / h ** 2
    m = idx["interior_to_full"]
    if mat is not None:
        k2 = np.zeros(g.dim_x)
        k2[m] = mat.kappa[m] ** 2
        sc = np.zeros(g.dim_x)
        sc[m] = mat.interior_sc_mask
        k2 = sh(k2.reshape(g.full_shape))
        sc = sh(sc.reshape(g.full_shape))
        gl = (1.0 - np.conj(c) * c) * c
        fpsi = sc * gl - (1.0 - sc) * c / TAU_RELAX
    else:
        k2 = g.kappa ** 2
        fpsi = (1.0 - np.conj(c) * c) * c
    dpsi = lap + fpsi

    ph = {"x": (y1, (1, 0, 0), g.hx), "y": (y2, (0, 1, 0), g.hy)}
    if nz3:
        ph["z"] = (y3, (0, 0, 1), g.hz)
    outs = []
    for a, (ya, da, ha) in ph.items():
        acc = 0.0
        for b, (yb, db, hb) in ph.items():
            d2 = sh(ya, *[-t for t in db]) - 2.0 * sh(ya) + sh(ya, *db)
            acc = acc + (2.0 if b == a else 1.0) * k2 / hb ** 2 * d2
            if b != a:
                # -phi_b(+a) + phi_b + phi_b(+a-b) - phi_b(-b)
                apb = tuple(p - q for p, q in zip(da, db))
                mb = tuple(-q for q in db)
                acc = acc + k2 / hb ** 2 * (-sh(yb, *da) + sh(yb) + sh(yb, *apb) - sh(yb, *mb))
        acc = acc + np.imag(np.exp(-1j * sh(ya)) * np.conj(c) * sh(x, *da))
        outs.append(acc.astype(np.complex128))

    def to_ref(A):  # (k, j, i) interior block -> reference order (i slowest, k fastest)
        return np.ascontiguousarray(A.transpose(2, 1, 0)).ravel()

    res = [to_ref(dpsi)] + [to_ref(o) for o in outs]
    n = g.n_interior
    for q, ax in enumerate("xyz"):
        hm = idx[f"hole_{ax}_bc_interior"]
        if hm.size and q + 1 < len(res):
            res[q + 1][hm] = 0.0
    assert all(r.size == n for r in res)
    return np.concatenate(res)


def jvp_stencil(X, P, g: Grid, idx: dict, u, mat: Optional[Material] = None) -> np.ndarray:
    """Exact directional derivative J(X) P of eval_f (the operator of rhs.py:147-239).

    The boundary rules are affine in the state (the applied field enters as
    constants), so the perturbation is expanded with the same rules and a zero
    field.  Per term: d[e^{i t} psi] = e^{i t}(d psi + i dt psi);
    d[(1-|psi|^2)psi] = (1 - 2|psi|^2) d psi - psi^2 conj(d psi);
    d Im[e^{-i t} conj(psi) psi1] = Im[e^{-i t}(-i dt conj(psi) psi1 + conj(d psi) psi1
    + conj(psi) d psi1)]; the link operators are linear.  Hole-boundary rows are 0.
    """
    X = np.asarray(X, np.complex128)
    P = np.asarray(P, np.complex128)
    zero = tuple(np.zeros(g.dim_x) for _ in range(3))
    x, y1, y2, y3 = (v.reshape(g.full_shape) for v in expand_with_bc(X, g, idx, u))
    dx, d1, d2, d3 = (v.reshape(g.full_shape) for v in expand_with_bc(P, g, idx, zero))
    nz3 = g.is_3d
    Ki = slice(1, g.Nz) if nz3 else slice(0, 1)
    Ji, Ii = slice(1, g.Ny), slice(1, g.Nx)

    def sh(A, di=0, dj=0, dk=0):
        return A[Ki.start + dk:Ki.stop + dk, Ji.start + dj:Ji.stop + dj, Ii.start + di:Ii.stop + di]

    c, dc = sh(x), sh(dx)
    lap = 0.0
    axes = [((1, 0, 0), y1, d1, g.hx), ((0, 1, 0), y2, d2, g.hy)]
    if nz3:
        axes.append(((0, 0, 1), y3, d3, g.hz))
    for (di, dj, dk), y, dy, h in axes:
        m_ = (-di, -dj, -dk)
        lap = lap + (np.exp(-1j * sh(y, *m_)) * (sh(dx, *m_) - 1j * sh(dy, *m_) * sh(x, *m_))
                     - 2.0 * dc
                     + np.exp(1j * sh(y)) * (sh(dx, di, dj, dk) + 1j * sh(dy) * sh(x, di, dj, dk))) / h ** 2
    m = idx["interior_to_full"]
    dgl = (1.0 - 2.0 * np.abs(c) ** 2) * dc - c * c * np.conj(dc)
    if mat is not None:
        k2 = np.zeros(g.dim_x)
        k2[m] = mat.kappa[m] ** 2
        sc = np.zeros(g.dim_x)
        sc[m] = mat.interior_sc_mask
        k2 = sh(k2.reshape(g.full_shape))
        sc = sh(sc.reshape(g.full_shape))
        dfpsi = sc * dgl - (1.0 - sc) * dc / TAU_RELAX
    else:
        k2 = g.kappa ** 2
        dfpsi = dgl
    out = [lap + dfpsi]
    ph = {"x": (y1, d1, (1, 0, 0), g.hx), "y": (y2, d2, (0, 1, 0), g.hy)}
    if nz3:
        ph["z"] = (y3, d3, (0, 0, 1), g.hz)
    for a, (ya, da_, va, ha) in ph.items():
        acc = 0.0
        for b, (yb, db_, vb, hb) in ph.items():
            acc = acc + (2.0 if b == a else 1.0) * k2 / hb ** 2 * (
                sh(da_, *[-t for t in vb]) - 2.0 * sh(da_) + sh(da_, *vb))
            if b != a:
                apb = tuple(p - q for p, q in zip(va, vb))
                mb = tuple(-q for q in vb)
                acc = acc + k2 / hb ** 2 * (-sh(db_, *va) + sh(db_) + sh(db_, *apb) - sh(db_, *mb))
        x1, dx1 = sh(x, *va), sh(dx, *va)
        acc = acc + np.imag(np.exp(-1j * sh(ya)) * (-1j * sh(da_) * np.conj(c) * x1
                                                   + np.conj(dc) * x1 + np.conj(c) * dx1))
        out.append(acc.astype(np.complex128))

    def to_ref(A):
        return np.ascontiguousarray(A.transpose(2, 1, 0)).ravel()

    res = [to_ref(o) for o in out]
    for q, ax in enumerate("xyz"):
        hm = idx[f"hole_{ax}_bc_interior"]
        if hm.size and q + 1 < len(res):
            res[q + 1][hm] = 0.0
    return np.concatenate(res)


eval_f = eval_f_csr


# ---------------------------------------------------------------------------
# Solvers -- solvers/tgcr.py:96-156, newton.py:76-120, integrators.py:35-206
# ---------------------------------------------------------------------------

def tgcr_trap(f: Callable, x_lin, b, dt, tol=1e-4, max_iter=None, eps_mf=1e-4, stats=None):
    """Truncated GCR on (I - dt/2 J) dx = b with the finite-difference product
    J p ~ (f(x_lin + eps p) - f(x_lin)) / eps and modified Gram-Schmidt
    (tgcr.py:96-156).  Returns an empty array when it does not converge."""
    N = len(b)
    if max_iter is None:
        max_iter = max(N, int(round(0.2 * N)))
    x = np.zeros_like(b)
    r = b.copy()
    r0 = np.linalg.norm(r, 2)
    rk = r0
    if stats is not None:
        stats["tgcr_iters"] = 0
    if r0 == 0.0:
        return x
    P, AP = [], []
    xn = np.linalg.norm(x_lin)
    k = 0
    while rk / r0 > tol and k < max_iter:
        k += 1
        p = r.copy()
        eps = eps_mf * (1.0 + xn) / np.linalg.norm(p)
        Ap = p - (dt / 2.0) * ((f(x_lin + eps * p) - f(x_lin)) / eps)
        for pj, apj in zip(P, AP):
            beta = np.vdot(Ap, apj).real
            p = p - beta * pj
            Ap = Ap - beta * apj
        nrm = np.linalg.norm(Ap, 2)
        if nrm < 1e-14:
            break
        Ap /= nrm
        p /= nrm
        P.append(p)
        AP.append(Ap)
        alpha = np.vdot(r, Ap).real
        x = x + alpha * p
        r = r - alpha * Ap
        rk = np.linalg.norm(r, 2)
    if stats is not None:
        stats["tgcr_iters"] = k
    if rk > tol * r0:
        return np.array([], dtype=np.complex128)
    return x


def newton_trap(f: Callable, x0, gamma, dt, tol_f=1e-3, tol_dx=1e-3, max_iter=20,
                tol_gcr=1e-4, eps_mf=1e-4, tgcr_max_iter=None, stats=None):
    """Newton on g(x) = x - dt/2 f(x) - gamma with TGCR inner solves
    (newton.py:76-120).  Returns (x, converged, iterations)."""
    x = np.array(x0, dtype=np.complex128)
    tg = []
    for k in range(1, max_iter + 1):
        g = x - (dt / 2.0) * f(x) - gamma
        st = {}
        dx = tgcr_trap(f, x, -g, dt, tol=tol_gcr, max_iter=tgcr_max_iter, eps_mf=eps_mf,
                       stats=st)
        tg.append(st["tgcr_iters"])
        if dx.size == 0:
            if stats is not None:
                stats["tgcr"] = tg
            return x, False, k
        x = x + dx
        g = x - (dt / 2.0) * f(x) - gamma
        if np.linalg.norm(g, np.inf) <= tol_f and np.linalg.norm(dx, np.inf) <= tol_dx:
            if stats is not None:
                stats["tgcr"] = tg
            return x, True, k
    if stats is not None:
        stats["tgcr"] = tg
    return x, False, max_iter


def trapezoidal(x0, film: Film, field_at: Callable, t_start, t_stop, dt, *,
                newton_tol_f=1e-3, newton_tol_dx=1e-3, newton_max_iter=20, tol_gcr=1e-4,
                eps_mf=1e-4, save_every=1, adaptive=True, dt_min=1e-8, tgcr_max_iter=None,
                max_accepted=None, log=None, predictor="euler"):
    """Adaptive trapezoidal stepping (integrators.py:92-206).  ``field_at(t)``
    returns (Bx, By, Bz); the field is frozen at t_n for the implicit solve.
    ``max_accepted`` stops after that many accepted steps (a bounded sample).
    ``predictor="extrapolate"`` (an extension with no counterpart in the
    reference): after the first accepted step Newton starts from
    X + (dt / dt_prev) (X - X_prev), X_prev the state before the last accepted
    step and dt_prev its size, instead of X + dt f(X) (integrators.py:156)."""
    g, idx, mat = film.g, film.idx, film.material
    X = np.array(x0, dtype=np.complex128)
    t = t_start
    cur = dt
    times, hist = [t], [X.copy()]
    step = 0
    x_prev, dt_prev = None, 0.0
    while t < t_stop - 1e-14:
        dta = min(cur, t_stop - t)
        u = boundary_vectors(*field_at(t), g, idx)

        def f(Y, u=u):
            return eval_f_csr(Y, g, idx, u, mat)

        fn = f(X)
        xp = X + dta * fn
        if predictor == "extrapolate" and x_prev is not None:
            xp = X + (dta / dt_prev) * (X - x_prev)
        gam = X + (dta / 2.0) * fn
        st = {}
        xn, conv, it = newton_trap(f, xp, gam, dta, newton_tol_f, newton_tol_dx,
                                   newton_max_iter, tol_gcr, eps_mf, tgcr_max_iter, stats=st)
        if log is not None:
            log.append((bool(conv), int(it), float(dta), list(st.get("tgcr", []))))
        if conv:
            x_prev, dt_prev = X, dta
            X = xn
            t += dta
            step += 1
            if step % save_every == 0:
                times.append(t)
                hist.append(X.copy())
            cur = dt
            if max_accepted is not None and step >= max_accepted:
                break
        else:
            if not adaptive:
                raise RuntimeError(f"Newton did not converge at t={t:.6e} with dt={dta:.3e}.")
            cur /= 10.0
            if cur < dt_min:
                raise RuntimeError(f"dt reduced below dt_min={dt_min:.1e} at t={t:.6e}; aborting.")
    if times[-1] < t:
        times.append(t)
        hist.append(X.copy())
    return np.array(times), np.column_stack(hist)


def forward_euler(x0, film: Film, field_at: Callable, t_start, t_stop, dt, save_every=1):
    """X += dt f(X) with the reference's step count and save rule (integrators.py:35-89)."""
    g, idx, mat = film.g, film.idx, film.material
    n_steps = int(np.ceil((t_stop - t_start) / dt))
    X = np.array(x0, dtype=np.complex128)
    times, hist = [t_start], [X.copy()]
    t = t_start
    for s in range(n_steps):
        dta = min(dt, t_stop - t)
        u = boundary_vectors(*field_at(t), g, idx)
        X = X + dta * eval_f_csr(X, g, idx, u, mat)
        t += dta
        if (s + 1) % save_every == 0 or s == n_steps - 1:
            times.append(t)
            hist.append(X.copy())
    return np.array(times), np.column_stack(hist)


# ---------------------------------------------------------------------------
# Observables -- physics/bfield.py:12-78, current_density.py:14-72, solution.py:94-217
# ---------------------------------------------------------------------------

def bfield_full(X, g: Grid, idx: dict, u=None):
    """B = curl(phi) at every interior node on the expanded state; with ``u``
    the BCs are applied first (solution.py:124-158, bfield.py:12-78)."""
    n = g.n_interior
    X = np.asarray(X, np.complex128)
    if u is None:
        x = expand(X[:n], g, idx)
        y1 = expand(X[n:2 * n], g, idx)
        y2 = expand(X[2 * n:3 * n], g, idx)
        y3 = expand(X[3 * n:4 * n] if g.is_3d else np.zeros(n), g, idx)
    else:
        x, y1, y2, y3 = expand_with_bc(X, g, idx, u)
    m, mj, mk = idx["interior_to_full"], g.mj, g.mk
    bz = (1.0 / (g.hx * g.hy)) * (y1[m] - y1[m + mj] - y2[m] + y2[m + 1])
    if g.is_3d:
        bx = (1.0 / (g.hy * g.hz)) * (y2[m] - y2[m + mk] - y3[m] + y3[m + mj])
        by = (1.0 / (g.hz * g.hx)) * (y3[m] - y3[m + 1] - y1[m] + y1[m + mk])
    else:
        bx = np.zeros(n)
        by = np.zeros(n)
    return np.real(bx), np.real(by), np.real(bz)


def supercurrent(X, g: Grid, idx: dict):
    """J_s,a = Im[conj(psi) e^{-i phi_a} psi(+a)] on the expanded state without BCs
    (solution.py:180-217, current_density.py:14-72)."""
    n = g.n_interior
    X = np.asarray(X, np.complex128)
    x = expand(X[:n], g, idx)
    ys = [expand(X[(q + 1) * n:(q + 2) * n], g, idx) for q in range(3 if g.is_3d else 2)]
    m = idx["interior_to_full"]
    st = [1, g.mj, g.mk]
    J = [np.imag(np.conj(x[m]) * np.exp(-1j * ys[q][m]) * x[m + st[q]]) for q in range(len(ys))]
    if not g.is_3d:
        J.append(np.zeros(n))
    return tuple(J)
