"""GPU: batches of uneven films in lockstep (mixed precision).

* A film's arithmetic does not depend on the other films of its batch: each
  film owns its work items, streaming blocks and reduction order, so a batch
  reproduces every film's own batch-of-one step bit for bit (with or without
  a Chebyshev preconditioner, of a fixed degree or of each film's own
  automatic degree).
* Films switched off for an attempt are left untouched.
* Against FP64 steps of the same films: |psi|^2, B_z and J_s agree to 1e-6
  relative (of the largest entry) after several steps.
* The stencil tiling (tile shape and z depth per work item) changes no node's
  arithmetic in the JVP stencil: its output is bitwise identical for every
  shape and depth.
"""

from __future__ import annotations

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _films():
    import nanowire_tdgl_amd as T
    from nanowire_tdgl_amd.solvers.batch import BatchFilm
    specs = [
        dict(p=dict(Nx=20, Ny=14, Nz=6, hx=0.5, hy=0.5, hz=0.5, kappa=2.0), bz=0.3,
             hole=[(2.0, 1.5), (4.5, 1.5), (4.5, 4.0), (2.0, 4.0)]),
        dict(p=dict(Nx=70, Ny=9, Nz=4, hx=0.5, hy=0.6, hz=0.7, kappa=3.5), bz=0.6, hole=None),
        dict(p=dict(Nx=12, Ny=33, Nz=9, hx=0.5, hy=0.5, hz=0.5, kappa=1.5), bz=0.1, hole=None,
             spheres=([(6.0, 16.0, 4.0)], [2.5])),
    ]
    films = []
    for f, s in enumerate(specs):
        p = T.SimulationParameters(**s["p"])
        dev = T.Device(p, applied_field=T.AppliedField(Bz=s["bz"], t_on_fraction=1.0))
        if s["hole"]:
            dev.add_hole(s["hole"])
        if s.get("spheres"):
            dev.add_spheres(*s["spheres"])
        n = p.n_interior
        rng = np.random.default_rng(40 + f)
        x0 = np.concatenate([1.0 + 0.02 * (rng.standard_normal(n) + 1j * rng.standard_normal(n)),
                             np.zeros(3 * n, complex)])
        films.append(BatchFilm(dev, x0))
    return films


def _single(film, steps, dt, precision, extrapolate=False, **kw):
    from nanowire_tdgl_amd import _native
    from nanowire_tdgl_amd.engine import engine_for
    d = film.device
    eng = engine_for(d.params, d.idx, d.material)
    X = eng.ctx.to_internal(film.x0)
    opts = _native.solver_opts(precision=precision, **kw)
    stats = []
    x_prev = None
    for s in range(steps):
        eng.set_field(*d.applied_field.evaluate(s * dt, 100.0))
        x_in = X.clone() if extrapolate else None
        st = eng.ctx.trap_step(X, dt, opts, x_prev=x_prev, dt_prev=dt)
        assert st.converged == 1
        stats.append(st.as_dict())
        x_prev = x_in
    return eng, X, stats


def _observables(eng, X, n):
    b, j = eng.ctx.observables(X)
    psi = eng.ctx.to_reference(X).cpu().numpy()[:n]
    return np.abs(psi) ** 2, b.cpu().numpy()[2 * n:], j.cpu().numpy()


@pytest.mark.parametrize("precond", ["none", 6, "chebyshev"])
def test_batch_reproduces_each_film_bitwise(precond):
    from nanowire_tdgl_amd import _native
    from nanowire_tdgl_amd.solvers.batch import BatchRunner
    films = _films()
    steps, dt = 3, 0.2
    singles = [_single(f, steps, dt, "mixed", precond=precond) for f in films]
    got = [s[1].clone() for s in singles]
    runner = BatchRunner([f.device for f in films])
    X = runner.pack([f.x0 for f in films])
    opts = _native.solver_opts(precision="mixed", precond=precond)
    for s in range(steps):
        runner.set_fields([s * dt] * len(films), 100.0)
        st = runner.batch.trap_step(X, dt, opts)
        for f, sf in enumerate(st):
            assert sf.converged == 1
            ref = singles[f][2][s]
            assert sf.newton_iters == ref["newton_iters"]
            assert sf.tgcr_iters_total == ref["tgcr_iters_total"]
    for f in range(len(films)):
        assert runner.batch.view(X, f).equal(got[f]), f


@pytest.mark.parametrize("tile", [0, 3])
def test_batch_with_general_field_film_reproduces_each_film_bitwise(tile):
    """One film of the batch carries full-grid field vectors (a non-uniform
    B_z on its faces): the batch then runs the general-field sweeps, per-cell
    ghost rules for that film and the uniform-field masks for the others, and
    still reproduces every film's own run bit for bit."""
    from nanowire_tdgl_amd import _native
    from nanowire_tdgl_amd.engine import engine_for
    from nanowire_tdgl_amd.physics.applied_field import build_boundary_field_vectors
    from nanowire_tdgl_amd.solvers.batch import BatchRunner
    films = _films()
    d1 = films[1].device
    Bx, By, Bz = (np.array(v, dtype=np.float64) for v in build_boundary_field_vectors(0.0, 0.0, 0.6, d1.params, d1.idx))
    Bz = Bz * (1.0 + 0.3 * np.random.default_rng(11).uniform(size=Bz.shape))  # non-uniform on the faces
    steps, dt = 2, 0.2
    opts = _native.solver_opts(precision="mixed")

    def fields(engines):
        for f, (film, eng) in enumerate(zip(films, engines)):
            if f == 1:
                eng.ctx.set_field_vectors(Bx, By, Bz)
                eng._field = None
            else:
                eng.set_field(*film.device.applied_field.evaluate(0.0, 100.0))

    engines = [engine_for(f.device.params, f.device.idx, f.device.material) for f in films]
    want = []
    for f, (film, eng) in enumerate(zip(films, engines)):
        eng.ctx.set_sweep_tile(tile)
        fields(engines)
        X = eng.ctx.to_internal(film.x0)
        for _ in range(steps):
            st = eng.ctx.trap_step(X, dt, opts)
            assert st.converged == 1
        want.append(X)
    runner = BatchRunner([f.device for f in films])
    runner.batch.set_tile(tile)
    X = runner.pack([f.x0 for f in films])
    for _ in range(steps):
        fields(runner.engines)
        st = runner.batch.trap_step(X, dt, opts)
        assert all(q.converged for q in st)
    for f in range(len(films)):
        assert runner.batch.view(X, f).equal(want[f]), f
    for eng in engines:
        eng.ctx.set_sweep_tile(0)


def test_batch_extrapolating_predictor_reproduces_each_film_bitwise():
    """Newton started from each film's extrapolated last step: the batch still
    reproduces every film's batch-of-one steps bit for bit (the first step has
    no history and takes the reference predictor), with fewer Newton
    iterations than the reference predictor."""
    from nanowire_tdgl_amd import _native
    from nanowire_tdgl_amd.solvers.batch import BatchRunner
    films = _films()
    steps, dt = 4, 0.2
    singles = [_single(f, steps, dt, "mixed", extrapolate=True, precond=6) for f in films]
    plain = [_single(f, steps, dt, "mixed", precond=6) for f in films]
    runner = BatchRunner([f.device for f in films])
    X = runner.pack([f.x0 for f in films])
    opts = _native.solver_opts(precision="mixed", precond=6)
    x_prev = None
    for s in range(steps):
        runner.set_fields([s * dt] * len(films), 100.0)
        x_in = X.clone()
        st = runner.batch.trap_step(X, dt, opts, x_prev=x_prev, dt_prev=[dt] * len(films))
        x_prev = x_in
        for f, sf in enumerate(st):
            assert sf.converged == 1
            assert sf.newton_iters == singles[f][2][s]["newton_iters"]
            assert sf.tgcr_iters_total == singles[f][2][s]["tgcr_iters_total"]
    for f in range(len(films)):
        assert runner.batch.view(X, f).equal(singles[f][1]), f
    ne = sum(r["newton_iters"] for s in singles for r in s[2])
    nr = sum(r["newton_iters"] for s in plain for r in s[2])
    print(f"Newton iterations over the films: extrapolating {ne}, reference predictor {nr}")
    assert ne < nr
    with pytest.raises(ValueError):
        runner.batch.trap_step(X, dt, opts, x_prev=X[:10], dt_prev=dt)


def test_trapezoidal_batch_extrapolating_matches_single_films():
    """trapezoidal_batch's per-film history bookkeeping (predictor="extrapolate")
    against each film's own trapezoidal run with the same predictor: the same
    final states bit for bit."""
    from nanowire_tdgl_amd.solvers import integrators
    from nanowire_tdgl_amd.solvers.batch import trapezoidal_batch
    from nanowire_tdgl_amd.solvers.runner import _make_eval_u
    films = _films()
    t_end, dt = 1.2, 0.2
    res = trapezoidal_batch(films, 0.0, t_end, dt, predictor="extrapolate")
    for f, film in enumerate(films):
        d = film.device
        eval_u = _make_eval_u(d.applied_field, d.params, d.idx, t_end)
        log = []
        times, hist = integrators.trapezoidal(film.x0, d.params, d.idx, eval_u, 0.0, t_end, dt,
                                              precision="mixed", predictor="extrapolate",
                                              material=d.material, progress=False, step_log=log)
        assert res.steps[f] == sum(1 for r in log if r["converged"])
        assert np.array_equal(hist[:, -1], res.states[f]), f


def test_inactive_films_are_untouched():
    from nanowire_tdgl_amd import _native
    from nanowire_tdgl_amd.solvers.batch import BatchRunner
    films = _films()
    runner = BatchRunner([f.device for f in films])
    X = runner.pack([f.x0 for f in films])
    before = runner.batch.view(X, 1).clone()
    runner.set_fields([0.2] * 3, 100.0)
    st = runner.batch.trap_step(X, [0.2, 0.2, 0.2], _native.solver_opts(precision="mixed"),
                                active=[1, 0, 1])
    assert st[0].converged == 1 and st[2].converged == 1
    assert st[1].converged == 0 and st[1].newton_iters == 0 and st[1].rhs_evals == 0
    assert runner.batch.view(X, 1).equal(before)


def test_batch_rejects_fp64_and_bad_dt():
    from nanowire_tdgl_amd import _native
    from nanowire_tdgl_amd.solvers.batch import BatchRunner
    films = _films()
    runner = BatchRunner([f.device for f in films])
    X = runner.pack([f.x0 for f in films])
    with pytest.raises(ValueError, match="mixed"):
        runner.batch.trap_step(X, 0.2, _native.solver_opts())
    with pytest.raises(ValueError, match="dt"):
        runner.batch.trap_step(X, [0.2, -1.0, 0.2], _native.solver_opts(precision="mixed"))


def test_batch_error_against_fp64():
    from nanowire_tdgl_amd.solvers.batch import trapezoidal_batch
    films = _films()
    steps, dt = 6, 0.2
    res = trapezoidal_batch(films, 0.0, steps * dt, dt, tol_gcr=1e-8)
    assert res.steps == [steps] * len(films)
    for f, film in enumerate(films):
        n = film.device.params.n_interior
        eng, X, _ = _single(film, steps, dt, "f64", tol_gcr=1e-8)
        a = _observables(eng, X, n)
        Xm = eng.ctx.to_internal(res.states[f])
        b = _observables(eng, Xm, n)
        for k, name in enumerate(("|psi|^2", "B_z", "J_s")):
            err = rel_err(b[k], a[k])
            print(f"film {f}: {name} relative difference {err:.2e}")
            assert err < 1e-6, (f, name)


def test_stencil_depth_does_not_change_the_jvp(tmp_path):
    import torch
    from nanowire_tdgl_amd.solvers.batch import BatchRunner
    films = _films()
    runner = BatchRunner([f.device for f in films])
    ctx = runner.engines[2].ctx
    g = torch.Generator(device="cuda").manual_seed(5)
    x = 0.5 + 0.1 * torch.rand(ctx.len, device="cuda", generator=g, dtype=torch.float32)
    p = torch.rand(ctx.len, device="cuda", generator=g, dtype=torch.float32) - 0.5
    outs = []
    for shape in (0, 1, 2, 3):
        ctx.set_sweep_tile(shape)
        assert ctx.get_sweep_tile() == shape
        for kz in (1, 3, 8):
            out = torch.empty_like(x)
            ctx.set_stencil_kz(kz)
            assert ctx.get_stencil_kz() == kz
            ctx.jvp_trap32(x, p, 0.2, out)
            outs.append(out)
    assert all(o.equal(outs[0]) for o in outs[1:])
    with pytest.raises(ValueError):
        ctx.set_sweep_tile(4)
    # the work split of the batch's sweeps changes no node's arithmetic either:
    # a whole mixed step gives the same bits for every split
    from nanowire_tdgl_amd import _native
    steps = []
    for shape, split in ((0, (8, 2)), (0, (1, 1)), (0, (3, 3)), (3, (2, 2)), (3, (64, 1)), (1, (8, 2))):
        runner.batch.set_tile(shape)
        runner.batch.set_split(*split)
        assert runner.batch.get_split() == split
        Xb = runner.pack([f.x0 for f in films])
        st = runner.batch.trap_step(Xb, 0.2, _native.solver_opts(precision="mixed"))
        assert all(q.converged for q in st)
        steps.append(Xb)
    assert all(o.equal(steps[0]) for o in steps[1:])
    with pytest.raises(ValueError):
        runner.batch.set_split(2, 4)
    cache = tmp_path / "autotune.json"
    info = runner.tune(reps=2, cache=str(cache))
    assert len(info["kz"]) == 3 and info["ms"] > 0 and info["shape"] in (0, 1, 2, 3)
    assert runner.batch.get_tile() == info["shape"] and not info["cached"]
    assert list(runner.batch.get_split()) == info["split"]
    assert cache.exists()
    # a second tuning of the same films reads the pick from the cache
    other = BatchRunner([f.device for f in films])
    again = other.tune(reps=2, cache=str(cache))
    assert again["cached"] and again["split"] == info["split"] and again["shape"] == info["shape"]
    assert other.batch.get_tile() == info["shape"]
    assert list(other.batch.get_split()) == info["split"]
    # different films: a different key, tuned afresh
    fewer = BatchRunner([f.device for f in films[:2]])
    assert not fewer.tune(reps=1, cache=str(cache))["cached"]


def test_graph_replay_matches_eager_launches():
    """TGCR iterations replayed from the captured hipGraph give the same bits
    as the same kernels launched one by one, over several steps with eager
    work (field changes, observables) between the replays."""
    from nanowire_tdgl_amd import _native
    from nanowire_tdgl_amd.solvers.batch import BatchRunner
    films = _films()
    opts = _native.solver_opts(precision="mixed")
    outs = []
    for graphs in (False, True):
        runner = BatchRunner([f.device for f in films])
        runner.batch.set_graphs(graphs)
        X = runner.pack([f.x0 for f in films])
        for s in range(4):
            runner.set_fields([0.1 * s] * len(films), 100.0)
            st = runner.batch.trap_step(X, 0.2, opts)
            assert all(x.converged for x in st)
            runner.engines[0].ctx.observables(runner.batch.view(X, 0))
        outs.append(X.clone())
    assert outs[0].equal(outs[1])


@pytest.mark.parametrize("groups,keyed", [(2, False), (3, False), (8, False), (2, True)])
def test_concurrent_groups_reproduce_one_batch_bitwise(groups, keyed):
    """trapezoidal_batch(groups=G): the films run as G sub-batches on their own
    streams and host threads at once (graph capture on), and every film ends
    in the same state, time and step count as in one batch, bit for bit; the
    per-film logs agree too, attempt by attempt.  More groups than films give
    one film per group; films grouped by a key (here the plane count) too."""
    from nanowire_tdgl_amd.solvers.batch import BatchGroups, trapezoidal_batch
    films = _films() + _films()[:2]
    one = trapezoidal_batch(films, 0.0, 0.8, 0.2, keep_log=True)
    keys = [f.device.params.Nz for f in films] if keyed else None
    grp = BatchGroups([f.device for f in films], groups, keys=keys)
    assert len(grp) == min(groups, len(films))
    assert sorted(i for p in grp.parts for i in p) == list(range(len(films)))
    # the groups' autotuner (timed together) leaves every group on its pick
    pick = grp.tune(reps=2, repeats=1)
    assert tuple(pick["split"]) in BatchGroups.GROUP_SPLITS
    assert all(r.batch.get_tile() == pick["shape"] and list(r.batch.get_split()) == pick["split"]
               for r in grp.runners)
    for pred in ("euler", "extrapolate"):
        ref = one if pred == "euler" else trapezoidal_batch(films, 0.0, 0.8, 0.2, keep_log=True,
                                                            predictor=pred)
        res = trapezoidal_batch(films, 0.0, 0.8, 0.2, keep_log=True, predictor=pred, runner=grp,
                                groups=groups)
        assert res.steps == ref.steps and res.times == ref.times
        for f in range(len(films)):
            assert np.array_equal(res.states[f], ref.states[f]), (pred, f)
            mine = [row[f] for row in res.log if row[f] is not None]
            theirs = [row[f] for row in ref.log][:len(mine)]
            assert mine == theirs, (pred, f)
