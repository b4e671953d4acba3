"""Grid index maps (mirror of tdgl3d's GridIndices / construct_indices,
src/tdgl3d/mesh/indices.py:61-450).

The arrays are built by broadcasting; their values and order equal the
reference's.  The GPU kernels do not read them: they use index arithmetic on
the internal layout.  These arrays are the public interface (Solution, user
code) and the source of the hole masks uploaded to the device.
"""

from __future__ import annotations

from dataclasses import dataclass, field, fields

import numpy as np

from ..core.parameters import SimulationParameters


def _empty():
    return field(default_factory=lambda: np.array([], dtype=np.intp))


@dataclass
class GridIndices:
    interior_to_full: np.ndarray = _empty()
    bfield_interior: np.ndarray = _empty()
    x_face_lo: np.ndarray = _empty()
    x_first: np.ndarray = _empty()
    x_last: np.ndarray = _empty()
    x_face_hi: np.ndarray = _empty()
    x_face_lo_inner: np.ndarray = _empty()
    x_first_inner: np.ndarray = _empty()
    x_last_inner: np.ndarray = _empty()
    x_face_hi_inner: np.ndarray = _empty()
    y_face_lo: np.ndarray = _empty()
    y_first: np.ndarray = _empty()
    y_last: np.ndarray = _empty()
    y_face_hi: np.ndarray = _empty()
    y_face_lo_inner: np.ndarray = _empty()
    y_first_inner: np.ndarray = _empty()
    y_last_inner: np.ndarray = _empty()
    y_face_hi_inner: np.ndarray = _empty()
    z_face_lo: np.ndarray = _empty()
    z_first: np.ndarray = _empty()
    z_last: np.ndarray = _empty()
    z_face_hi: np.ndarray = _empty()
    z_face_lo_inner: np.ndarray = _empty()
    z_first_inner: np.ndarray = _empty()
    z_last_inner: np.ndarray = _empty()
    z_face_hi_inner: np.ndarray = _empty()
    x_normal_bc_mask: np.ndarray = _empty()
    y_normal_bc_mask: np.ndarray = _empty()
    z_normal_bc_mask: np.ndarray = _empty()
    hole_x_bc_mask: np.ndarray = _empty()
    hole_y_bc_mask: np.ndarray = _empty()
    hole_z_bc_mask: np.ndarray = _empty()
    hole_x_bc_interior: np.ndarray = _empty()
    hole_y_bc_interior: np.ndarray = _empty()
    hole_z_bc_interior: np.ndarray = _empty()

    def define_hole_polygon(self, vertices, z_range, params: SimulationParameters) -> None:
        """Append the polygon's boundary links, in full-grid and interior
        numbering (indices.py:136-232)."""
        from .holes import identify_boundary_links, identify_hole_nodes

        mask = identify_hole_nodes(vertices, z_range, params.hx, params.hy,
                                   params.Nx, params.Ny, params.Nz)
        full_to_int = np.full(params.dim_x, -1, dtype=np.intp)
        full_to_int[self.interior_to_full] = np.arange(params.n_interior, dtype=np.intp)
        for ax in "xyz":
            links = identify_boundary_links(mask, ax, is_3d=params.is_3d)
            name = f"hole_{ax}_bc_mask"
            setattr(self, name, np.concatenate([getattr(self, name), links]))
            inter = full_to_int[links]
            inter = inter[inter >= 0].astype(np.intp)
            if inter.size:
                iname = f"hole_{ax}_bc_interior"
                setattr(self, iname, np.concatenate([getattr(self, iname), inter]))

    def hole_link_bits(self, params: SimulationParameters) -> np.ndarray | None:
        """uint8 per full-grid node: bit 0/1/2 = x/y/z hole-boundary link."""
        if not (self.hole_x_bc_mask.size or self.hole_y_bc_mask.size or self.hole_z_bc_mask.size):
            return None
        bits = np.zeros(params.dim_x, dtype=np.uint8)
        for b, arr in ((1, self.hole_x_bc_mask), (2, self.hole_y_bc_mask), (4, self.hole_z_bc_mask)):
            if arr.size and (b != 4 or params.is_3d):
                bits[np.asarray(arr, dtype=np.intp)] |= b
        return bits

    def array_names(self) -> list[str]:
        return [f.name for f in fields(self)]


def construct_indices(params: SimulationParameters) -> GridIndices:
    """All index arrays for `params` (indices.py:239-450)."""
    Nx, Ny, Nz = params.Nx, params.Ny, params.Nz
    mj, mk = params.mj, params.mk
    three_d = params.is_3d

    def rng(a, b):
        return np.arange(a, b, dtype=np.intp)

    ax_all = {"x": rng(0, Nx + 1), "y": rng(0, Ny + 1),
              "z": rng(0, Nz + 1) if three_d else np.zeros(1, np.intp)}
    ax_int = {"x": rng(1, Nx), "y": rng(1, Ny),
              "z": rng(1, Nz) if three_d else np.zeros(1, np.intp)}
    stride = {"x": 1, "y": mj, "z": mk}

    def lin(**coords) -> np.ndarray:
        """Linear ids over the outer product of the given axes, in the order given
        (first axis slowest)."""
        axes = list(coords)
        out = np.zeros([len(coords[a]) for a in axes], dtype=np.intp)
        for d, a in enumerate(axes):
            shape = [1] * len(axes)
            shape[d] = -1
            out = out + stride[a] * np.asarray(coords[a], np.intp).reshape(shape)
        return out.ravel()

    idx = GridIndices()
    # interior order: i slowest, k fastest (indices.py:271-278)
    idx.interior_to_full = lin(x=ax_int["x"], y=ax_int["y"], z=ax_int["z"])

    # bfield_interior: an i-fastest numbering one layer in (indices.py:283-300)
    nxi, nyi = Nx - 1, Ny - 1
    bi, bj = rng(0, Nx - 2), rng(0, Ny - 2)
    if three_d:
        bk = rng(0, Nz - 2)
        idx.bfield_interior = (bi[:, None, None] + nxi * bj[None, :, None]
                               + nxi * nyi * bk[None, None, :]).ravel()
    else:
        idx.bfield_interior = (bi[:, None] + nxi * bj[None, :]).ravel()

    others = {"x": ("y", "z"), "y": ("x", "z"), "z": ("x", "y")}
    last = {"x": Nx, "y": Ny, "z": Nz}
    for a in "xyz":
        o1, o2 = others[a]
        if a == "z" and not three_d:
            continue
        N = last[a]
        for nm, v in (("face_lo", 0), ("first", 1), ("last", N - 1), ("face_hi", N)):
            val = np.array([v], np.intp)
            for suffix, src in (("", ax_all), ("_inner", ax_int)):
                # the two transverse axes vary (first slowest); `a` is fixed at v
                coords = {o1: src[o1], o2: src[o2], a: val}
                setattr(idx, f"{a}_{nm}{suffix}", lin(**coords))
        # both faces perpendicular to `a`, over every transverse node (indices.py:392-430)
        parts = [lin(**{o1: ax_all[o1], o2: ax_all[o2], a: np.array([v], np.intp)})
                 for v in (0, N - 1)]
        setattr(idx, f"{a}_normal_bc_mask", np.concatenate(parts))
    return idx
