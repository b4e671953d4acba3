"""Host-side logic of the library (no GPU): the tdgl3d call surface, index
maps, hole masks, material maps, applied-field vectors and the data uploaded
to the device, checked bit-exact against tdgl3d's fixtures and the oracle."""

from __future__ import annotations

import numpy as np
import pytest

import nanowire_tdgl_amd as T
import specs as S
from conftest import golden
from helpers import device, oracle_film
from nanowire_tdgl_amd.mesh import holes as H
from nanowire_tdgl_amd.mesh.indices import construct_indices
from nanowire_tdgl_amd.physics.applied_field import build_boundary_field_vectors
from oracle import tdgl_oracle as O


@pytest.mark.parametrize("name", list(S.INDEX_CASES))
def test_indices_bit_exact(name):
    g = golden(f"indices_{name}")
    dev = device(S.INDEX_CASES[name])
    for k in g.files:
        got = getattr(dev.material, k) if k in ("sc_mask", "interior_sc_mask", "kappa") else getattr(dev.idx, k)
        assert np.array_equal(got, g[k]), k
        assert got.dtype.kind == g[k].dtype.kind, k


def test_film128_hole_bit_exact():
    import hashlib
    g = golden("indices_film128_hole")
    dev = device(S.FILM128)
    for k in g.files:
        if k.startswith("hole_"):
            assert np.array_equal(getattr(dev.idx, k), g[k]), k
    for k, dig in zip(g["digest_keys"], g["digest_vals"]):
        a = np.ascontiguousarray(np.asarray(getattr(dev.idx, str(k)), np.int64))
        assert hashlib.sha256(a.tobytes()).hexdigest() == str(dig), k
    assert int(np.sum(dev.material.sc_mask == 0.0)) == int(g["n_sc_zero"])


def test_random_polygons_match_oracle():
    rng = np.random.default_rng(11)
    for trial in range(12):
        Nx, Ny, Nz = (int(v) for v in rng.integers(4, 16, 3))
        hx, hy = rng.choice([0.5, 0.7, 1.0], 2)
        nv = int(rng.integers(3, 7))
        ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
        cx, cy, r = Nx * hx / 2, Ny * hy / 2, min(Nx * hx, Ny * hy) / 3
        verts = [(float(cx + r * np.cos(a) * rng.uniform(0.5, 1)),
                  float(cy + r * np.sin(a) * rng.uniform(0.5, 1))) for a in ang]
        # put some vertices exactly on grid lines to exercise the half-open tests
        if trial % 3 == 0:
            verts = [(round(x / hx) * hx, round(y / hy) * hy) for x, y in verts]
        zr = (int(rng.integers(0, Nz)), int(rng.integers(0, Nz + 1))) if trial % 2 else None
        case = dict(p=dict(Nx=Nx, Ny=Ny, Nz=Nz, hx=float(hx), hy=float(hy)), holes=[(verts, zr)])
        dev = device(case)
        f = oracle_film(case)
        for k in ("hole_x_bc_mask", "hole_y_bc_mask", "hole_z_bc_mask", "hole_x_bc_interior",
                  "hole_y_bc_interior", "hole_z_bc_interior"):
            assert np.array_equal(getattr(dev.idx, k), f.idx[k]), (trial, k)
        assert np.array_equal(dev.material.sc_mask, f.material.sc_mask)


def test_point_in_polygon_scalar_api():
    tri = [(0.0, 0.0), (10.0, 0.0), (5.0, 10.0)]
    assert H.point_in_polygon((5.0, 5.0), tri)
    assert not H.point_in_polygon((0.0, 5.0), tri)
    assert not H.point_in_polygon((15.0, 5.0), tri)
    with pytest.raises(ValueError, match="Invalid direction"):
        H.identify_boundary_links(np.zeros((3, 3, 3), bool), "w")


@pytest.mark.parametrize("field", [(0.0, 0.0, 0.3), (0.11, -0.23, 0.37), (0.5, 0.0, 0.0)])
def test_boundary_vectors_match_oracle(field):
    case = S.EVALF_CASES["aniso_bxyz"]
    dev = device(case)
    f = oracle_film(case)
    got = build_boundary_field_vectors(*field, dev.params, dev.idx)
    want = O.boundary_vectors(*field, f.g, f.idx)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert got.uniform == tuple(float(v) for v in field)


def test_parameters_and_state_surface():
    with pytest.raises(ValueError):
        T.SimulationParameters(Nx=1)
    with pytest.raises(ValueError):
        T.SimulationParameters(Nz=0)
    p = T.SimulationParameters(Nx=5, Ny=4, Nz=3)
    assert (p.n_interior, p.n_state, p.dim_x, p.mj, p.mk) == (24, 96, 120, 6, 30)
    p2 = T.SimulationParameters(Nx=5, Ny=4, Nz=1)
    assert (p2.n_interior, p2.n_state, p2.dim_x, p2.is_3d) == (12, 36, 30, False)
    sv = T.StateVector.uniform_superconducting(p)
    assert np.all(sv.psi == 1) and np.all(sv.phi_z == 0)
    with pytest.raises(ValueError):
        T.StateVector(np.zeros(5, complex), p)
    with pytest.raises(AttributeError):
        _ = T.StateVector.uniform_superconducting(p2).phi_z


def test_initial_state_zero_in_hole():
    dev = device(S.INDEX_CASES["hole3d"])
    st = dev.initial_state()
    assert np.sum(st.psi == 0) == np.sum(dev.material.interior_sc_mask == 0) > 0


def test_trilayer_material_matches_oracle():
    case = S.EVALF_CASES["trilayer_hole"]
    dev = device(case)
    f = oracle_film(case)
    for k in ("kappa", "sc_mask", "interior_sc_mask"):
        assert np.array_equal(getattr(dev.material, k), getattr(f.material, k)), k


def test_hole_link_bits():
    dev = device(S.INDEX_CASES["twoholes_zr"])
    bits = dev.idx.hole_link_bits(dev.params)
    for b, k in ((1, "hole_x_bc_mask"), (2, "hole_y_bc_mask"), (4, "hole_z_bc_mask")):
        assert set(np.nonzero(bits & b)[0].tolist()) == set(getattr(dev.idx, k).tolist())
    assert construct_indices(dev.params).hole_link_bits(dev.params) is None


def test_applied_field_evaluate():
    af = T.AppliedField(Bz=1.0, t_on_fraction=0.5)
    assert af.evaluate(0.3, 1.0)[2] == 1.0
    assert af.evaluate(0.6, 1.0)[2] == 0.0
    assert af.evaluate(0.0, 1.0)[2] == 0.0
    ramp = T.AppliedField(Bz=2.0, ramp=True, t_on_fraction=1.0)
    assert ramp.evaluate(0.25, 1.0)[2] == pytest.approx(1.0)
    custom = T.AppliedField(field_func=lambda t, ts: (0.0, 0.0, t / ts))
    assert custom.evaluate(0.5, 1.0)[2] == pytest.approx(0.5)


def test_bench_reads_the_committed_traffic_lines():
    """bench.py attaches the HBM traffic of each section's dominant kernel from
    the committed rocprofv3 summaries: every line it reads parses, per kernel."""
    import bench
    sweeps = bench.pmc_traffic("cheb_sweeps_mixed_sections.md")
    for section in ("batched", "mixed", "large"):
        assert sweeps[section]["cheb"] > 1e6, section
    fp64 = bench.pmc_traffic("pmc_traffic_film128_hole.md")
    assert fp64["mgs"] > 1e6
    roof = {"kernel": "cheb", "traffic": None}
    assert bench.with_traffic(dict(roof), "cheb_sweeps_mixed_sections.md", "batched")["traffic"] > 1e6


def test_partition_films_covers_every_film_once_and_balances():
    """The concurrent-group split: every film in exactly one group, groups
    ascending, none empty, and the heaviest group within one film of the mean
    (largest-first greedy)."""
    from nanowire_tdgl_amd.solvers.batch import partition_films
    rng = np.random.default_rng(5)
    for F, G in [(1, 1), (1, 3), (5, 2), (64, 3), (64, 8), (7, 7)]:
        sizes = [int(x) for x in rng.integers(100, 100000, F)]
        parts = partition_films(sizes, G)
        assert len(parts) == min(G, F)
        assert sorted(i for p in parts for i in p) == list(range(F))
        assert all(p == sorted(p) and p for p in parts)
        loads = [sum(sizes[i] for i in p) for p in parts]
        assert max(loads) <= sum(sizes) / len(parts) + max(sizes)
    assert partition_films([5, 9, 1, 3, 3, 7], 3) == [[1], [4, 5], [0, 2, 3]]
    with pytest.raises(ValueError):
        partition_films([1, 2], 0)


def test_partition_films_by_key_is_contiguous_in_key_order():
    """With keys, each group is a contiguous run of the films sorted by (key,
    index), every film in exactly one group, none empty, and the heaviest group
    within one film of the mean; a hand-worked case pins the cut rule (each film
    to the run its weight's midpoint falls in)."""
    from nanowire_tdgl_amd.solvers.batch import partition_films
    rng = np.random.default_rng(6)
    for _ in range(500):
        F, G = int(rng.integers(1, 70)), int(rng.integers(1, 10))
        sizes = [int(x) for x in rng.integers(100, 100000, F)]
        keys = [float(x) for x in rng.integers(0, 20, F)]
        parts = partition_films(sizes, G, keys)
        assert len(parts) == min(G, F)
        assert sorted(i for p in parts for i in p) == list(range(F))
        assert all(p == sorted(p) and p for p in parts)
        for a, b in zip(parts[:-1], parts[1:]):
            assert max((keys[i], i) for i in a) < min((keys[i], i) for i in b)
        loads = [sum(sizes[i] for i in p) for p in parts]
        assert max(loads) <= sum(sizes) / len(parts) + max(sizes)
    # keys order 2, 0, 3, 1 (weights 1, 1, 6, 2 of 10): midpoints 0.5, 1.5, 5, 9
    # of 10 fall in halves 0, 0, 1, 1
    assert partition_films([1, 2, 1, 6], 2, keys=[1.0, 3.0, 0.0, 2.0]) == [[0, 2], [1, 3]]
    # the heavy film's midpoint (52 of 104) falls in run 2 of 4, but it takes
    # run 1, which no film would hold otherwise
    assert partition_films([1, 1, 100, 1, 1], 4, keys=[0, 1, 2, 3, 4]) == [[0, 1], [2], [3], [4]]
    with pytest.raises(ValueError, match="one key per film"):
        partition_films([1, 2], 2, keys=[1.0])


def test_trapezoidal_batch_argument_errors():
    """Host-side checks of trapezoidal_batch, raised before any device work:
    an unknown predictor, a plain runner handed to a grouped run, and groups
    built for another number of films."""
    from nanowire_tdgl_amd.solvers.batch import BatchGroups, partition_films, trapezoidal_batch
    with pytest.raises(ValueError, match="predictor"):
        trapezoidal_batch([], 0.0, 1.0, 0.1, predictor="bdf2")
    with pytest.raises(ValueError, match="BatchGroups"):
        trapezoidal_batch([], 0.0, 1.0, 0.1, groups=2, runner=object())
    grp = BatchGroups.__new__(BatchGroups)  # the partition only: no batches on the device
    grp.parts = partition_films([5, 4, 3], 2)
    with pytest.raises(ValueError, match="holds 3 films; 2 were given"):
        trapezoidal_batch([object(), object()], 0.0, 1.0, 0.1, groups=2, runner=grp)
