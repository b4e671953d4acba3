"""Case definitions shared by the fixture generator and the tests.

Each case is plain data: SimulationParameters fields, holes as (vertices,
z_range) pairs, an optional trilayer as three (thickness, kappa, is_sc)
triples, the applied field and the seed of the random state.
"""

INDEX_CASES = {
    "plain3d": dict(p=dict(Nx=7, Ny=6, Nz=5), holes=[]),
    "plain2d": dict(p=dict(Nx=6, Ny=5, Nz=1), holes=[]),
    "hole3d": dict(p=dict(Nx=20, Ny=20, Nz=5, hx=1.0, hy=1.0, hz=1.0),
                   holes=[([(5.0, 5.0), (15.0, 5.0), (15.0, 15.0), (5.0, 15.0)], None)]),
    "lshape3d": dict(p=dict(Nx=30, Ny=30, Nz=5, hx=1.0, hy=1.0, hz=1.0),
                     holes=[([(5.0, 5.0), (15.0, 5.0), (15.0, 10.0), (10.0, 10.0),
                              (10.0, 15.0), (5.0, 15.0)], (0, 5))]),
    "twoholes_zr": dict(p=dict(Nx=24, Ny=20, Nz=6, hx=0.5, hy=0.5, hz=0.5),
                        holes=[([(2.0, 2.0), (4.0, 2.0), (4.0, 4.5), (2.0, 4.5)], (1, 3)),
                               ([(6.0, 5.0), (9.0, 5.5), (7.0, 8.0)], None)]),
    "hole2d": dict(p=dict(Nx=20, Ny=20, Nz=1, hx=1.0, hy=1.0, hz=1.0),
                   holes=[([(7.0, 7.0), (13.0, 7.0), (13.0, 13.0), (7.0, 13.0)], None)]),
}

# 128x128x16 film, h = 0.5, centred 32x32-cell square hole through all z.
FILM128 = dict(p=dict(Nx=128, Ny=128, Nz=16, hx=0.5, hy=0.5, hz=0.5, kappa=4.0),
               holes=[([(24.0, 24.0), (40.0, 24.0), (40.0, 40.0), (24.0, 40.0)], None)])

EVALF_CASES = {
    "c1_bz": dict(p=dict(Nx=16, Ny=16, Nz=8, hx=0.5, hy=0.5, hz=0.5, kappa=2.0),
                  field=(0.0, 0.0, 0.3), holes=[], seed=10),
    "aniso_bxyz": dict(p=dict(Nx=9, Ny=7, Nz=6, hx=0.5, hy=0.7, hz=0.9, kappa=1.5),
                       field=(0.11, -0.23, 0.37), holes=[], seed=11),
    "hole_bz": dict(p=dict(Nx=20, Ny=18, Nz=5, hx=1.0, hy=1.0, hz=1.0, kappa=2.0),
                    field=(0.05, 0.1, 0.4),
                    holes=[([(5.0, 5.0), (13.0, 5.0), (13.0, 12.0), (5.0, 12.0)], None)], seed=12),
    "hole_zr": dict(p=dict(Nx=16, Ny=14, Nz=6, hx=0.5, hy=0.5, hz=0.5, kappa=3.0),
                    field=(0.2, 0.0, 0.5),
                    holes=[([(2.0, 2.0), (5.0, 2.0), (5.0, 4.5), (2.0, 4.5)], (1, 3))], seed=13),
    "hole2d": dict(p=dict(Nx=14, Ny=12, Nz=1, hx=1.0, hy=1.0, hz=1.0, kappa=2.0),
                   field=(0.0, 0.0, 0.3),
                   holes=[([(4.0, 4.0), (9.0, 4.0), (9.0, 8.0), (4.0, 8.0)], None)], seed=14),
    "plain2d": dict(p=dict(Nx=10, Ny=9, Nz=1, hx=0.8, hy=0.6, kappa=2.5),
                    field=(0.0, 0.0, 0.45), holes=[], seed=15),
    "periodic_x": dict(p=dict(Nx=8, Ny=7, Nz=5, hx=0.5, hy=0.5, hz=0.5, kappa=2.0,
                              periodic_x=True),
                       field=(0.1, 0.2, 0.3), holes=[], seed=16),
    "periodic_xyz": dict(p=dict(Nx=6, Ny=7, Nz=5, hx=0.5, hy=0.6, hz=0.7, kappa=2.0,
                                periodic_x=True, periodic_y=True, periodic_z=True),
                         field=(0.1, 0.2, 0.3), holes=[], seed=17),
    "trilayer": dict(p=dict(Nx=10, Ny=9, Nz=7, hx=0.5, hy=0.5, hz=0.5, kappa=2.0),
                     field=(0.0, 0.1, 0.35), holes=[], seed=18,
                     trilayer=((2, 2.0, True), (3, 0.7, False), (2, 3.0, True))),
    "trilayer_hole": dict(p=dict(Nx=12, Ny=12, Nz=7, hx=0.5, hy=0.5, hz=0.5, kappa=2.0),
                          field=(0.0, 0.0, 0.35),
                          holes=[([(2.0, 2.0), (4.0, 2.0), (4.0, 4.0), (2.0, 4.0)], (0, 3))],
                          seed=19, trilayer=((3, 2.0, True), (2, 0.0, False), (2, 2.5, True))),
}

# trapezoidal trajectories
TRAP_C1 = dict(p=dict(Nx=16, Ny=16, Nz=8, hx=0.5, hy=0.5, hz=0.5, kappa=2.0), holes=[],
               field=dict(Bz=0.3, t_on_fraction=1.0), seed=0, noise=0.01, dt=0.1,
               tol_gcr=1e-8, tgcr_max_iter=30, t_stop=7.571, t_stop_abort=(10.0, 20.0))
TRAP_HOLE = dict(p=dict(Nx=12, Ny=10, Nz=4, hx=0.5, hy=0.5, hz=0.5, kappa=4.0),
                 holes=[([(2.0, 1.5), (3.6, 1.5), (3.6, 3.4), (2.0, 3.4)], None)],
                 field=dict(Bz=0.4), seed=1, noise=1e-3, dt=0.2, tol_gcr=1e-8,
                 tgcr_max_iter=None, t_stop=6.0)
TRAP_DTCUT = dict(p=dict(Nx=6, Ny=6, Nz=3, hx=0.5, hy=0.5, hz=0.5, kappa=2.0), holes=[],
                  field=dict(Bz=0.5, t_on_fraction=1.0), seed=5, noise=0.05, dt=0.5,
                  tol_gcr=1e-6, tgcr_max_iter=4, t_stop=2.0)
EULER_SMALL = dict(p=dict(Nx=10, Ny=9, Nz=4, hx=0.5, hy=0.5, hz=0.5, kappa=2.0),
                   holes=[([(1.5, 1.5), (3.0, 1.5), (3.0, 2.6), (1.5, 2.6)], None)],
                   field=dict(Bz=0.6), seed=6, noise=0.02, dt=0.02, t_stop=1.0, save_every=10)
OBS_ANISO = dict(p=dict(Nx=9, Ny=8, Nz=5, hx=0.5, hy=0.6, hz=0.7, kappa=2.0), holes=[],
                 field=dict(Bx=0.1, By=-0.2, Bz=0.3), seed=30, times=(0.0, 0.5, 1.0))
