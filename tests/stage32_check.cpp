// CPU check of the FP32 sweeps' cell staging (tdgl_stage32.hpp): cell_pair32
// must give bit for bit what eff_cell_pair gives for every cell of the
// extended grid (ghost faces, edges, corners and cells outside the grid
// included), on random 2-D and 3-D grids with periodic axes, uniform and
// per-node applied fields and hole-boundary link bits; and the mask form of
// the uniform-field staging (stage32u) must give exactly the bits of the
// per-cell rules (stage32), signs of zeros included.  The FP64 mask staging
// (cell64, plain and x + eps p) must give exactly the bits of eff_cell.
// Host code only: hipcc compiles the __host__ __device__ functions for the CPU.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <vector>

#include "tdgl_stage32.hpp"

using namespace tdgl;

static bool same(float a, float b) { return memcmp(&a, &b, sizeof(float)) == 0 || (a == 0.f && b == 0.f); }
static bool bits_equal(float a, float b) { return memcmp(&a, &b, sizeof(float)) == 0; }

template <bool IS3D>
static long long check(std::mt19937& rng, int trial) {
  std::uniform_int_distribution<int> dn(2, 9), dz(2, 7), bit(0, 1), b8(0, 15);
  std::uniform_real_distribution<double> ur(-1.0, 1.0);
  Geo g{};
  g.Nx = dn(rng);
  g.Ny = dn(rng);
  g.Nz = IS3D ? dz(rng) : 1;
  g.nx = g.Nx - 1;
  g.ny = g.Ny - 1;
  g.nz = IS3D ? g.Nz - 1 : 1;
  g.is3d = IS3D;
  g.per_x = bit(rng);
  g.per_y = bit(rng);
  g.per_z = IS3D ? bit(rng) : 0;
  g.mj = g.Nx + 1;
  g.mk = (long long)(g.Nx + 1) * (g.Ny + 1);
  g.n = (long long)g.nx * g.ny * g.nz;
  g.nphi = IS3D ? 3 : 2;
  const long long dimx = g.mk * (IS3D ? g.Nz + 1 : 1);
  const long long len = (2 + g.nphi) * g.n;
  std::vector<float> x(len), p(len);
  for (auto& v : x) v = (float)ur(rng);
  for (auto& v : p) v = (float)ur(rng);
  std::vector<double> bx(dimx), by(dimx), bz(dimx);
  for (long long m = 0; m < dimx; ++m) {
    bx[m] = ur(rng);
    by[m] = ur(rng);
    bz[m] = ur(rng);
  }
  Bnd b{};
  b.hx = 0.5 + 0.5 * bit(rng);
  b.hy = 0.7;
  b.hz = 0.9;
  const bool general = trial % 3 == 0;
  if (general) {
    b.bx = bx.data();
    b.by = by.data();
    b.bz = bz.data();
  } else {
    b.cx = ur(rng);
    b.cy = ur(rng);
    b.cz = ur(rng);
  }
  std::vector<uint8_t> flags(dimx);
  for (auto& f : flags) f = (uint8_t)(b8(rng) & 7);
  const uint8_t* fl = trial % 2 ? flags.data() : nullptr;
  const Src<float, false> sx{x.data(), nullptr, 0.f, g.n};
  const Src<float, false> sp{p.data(), nullptr, 0.f, g.n};
  // the FP32 sweeps read the mixed method's layout (each node's links together)
  std::vector<float> xl(len), pl(len);
  for (long long e = 0; e < len; ++e) {
    const long long i = internal_of_l32(e, g.n, g.nphi);
    xl[e] = x[i];
    pl[e] = p[i];
    if (l32_of_internal(i, g.n, g.nphi) != e) {
      printf("LAYOUT MAP MISMATCH n=%lld nphi=%d e=%lld\n", g.n, g.nphi, e);
      return -1;
    }
  }
  const V32 X = planes32(xl.data(), (unsigned)g.n, IS3D), P = planes32(pl.data(), (unsigned)g.n, IS3D);
  std::vector<double> xd(len), pd(len);
  for (long long e = 0; e < len; ++e) {
    xd[e] = (double)x[e] + 1e-9 * ur(rng);
    pd[e] = (double)p[e];
  }
  const double eps = 1e-4 * (1.0 + ur(rng));
  const Src<double, false> s64{xd.data(), nullptr, 0.0, g.n};
  const Src<double, true> s64p{xd.data(), pd.data(), eps, g.n};
  const V64 X64 = planes64(xd.data(), (unsigned)g.n), P64 = planes64(pd.data(), (unsigned)g.n);
  long long cells = 0, bad = 0;
  for (int k = IS3D ? -1 : 0; k <= (IS3D ? g.Nz + 1 : 0); ++k)
    for (int j = -1; j <= g.Ny + 1; ++j)
      for (int i = -1; i <= g.Nx + 1; ++i) {
        Cell<float> ax, ap, cx, cp, mx, mp;
        eff_cell_pair<float, IS3D>(g, b, fl, sx, sp, i, j, k, ax, ap);
        cell_pair32<IS3D>(g, b, fl, X, P, i, j, k, cx, cp);
        const float u[10] = {ax.pr, ax.pi, ax.x, ax.y, ax.z, ap.pr, ap.pi, ap.x, ap.y, ap.z};
        const float w[10] = {cx.pr, cx.pi, cx.x, cx.y, cx.z, cp.pr, cp.pi, cp.x, cp.y, cp.z};
        ++cells;
        for (int c = 0; c < 10; ++c)
          if (!same(u[c], w[c])) {
            if (bad < 10)
              printf("MISMATCH 3d=%d trial=%d N=(%d,%d,%d) per=(%d,%d,%d) cell=(%d,%d,%d) comp=%d: %g vs %g\n",
                     (int)IS3D, trial, g.Nx, g.Ny, g.Nz, g.per_x, g.per_y, g.per_z, i, j, k, c, u[c], w[c]);
            ++bad;
          }
        if (!general && k >= 0 && k <= (IS3D ? g.Nz : 0)) {  // FP64 mask staging vs eff_cell
          for (int pert = 0; pert < 2; ++pert) {
            const Cell<double> ref = pert ? eff_cell<double, IS3D, true>(g, b, fl, s64p, i, j, k)
                                          : eff_cell<double, IS3D, false>(g, b, fl, s64, i, j, k);
            const Cell<double> got = pert ? cell64<IS3D, true>(g, b, fl, X64, P64, eps, i, j, k)
                                          : cell64<IS3D, false>(g, b, fl, X64, P64, eps, i, j, k);
            const double u64[5] = {ref.pr, ref.pi, ref.x, ref.y, ref.z};
            const double w64[5] = {got.pr, got.pi, got.x, got.y, got.z};
            for (int c = 0; c < 5; ++c)
              if (memcmp(&u64[c], &w64[c], sizeof(double)) != 0) {
                if (bad < 10)
                  printf("FP64 MISMATCH 3d=%d pert=%d trial=%d N=(%d,%d,%d) per=(%d,%d,%d) cell=(%d,%d,%d) comp=%d: %.17g vs %.17g\n",
                         (int)IS3D, pert, trial, g.Nx, g.Ny, g.Nz, g.per_x, g.per_y, g.per_z, i, j, k, c, u64[c], w64[c]);
                ++bad;
              }
          }
        }
        if (!general && k >= 0 && k <= (IS3D ? g.Nz : 0)) {  // the planes a sweep stages
          cell_pair32<IS3D>(g, b, fl, X, P, i, j, k, mx, mp, true);
          const float v[10] = {mx.pr, mx.pi, mx.x, mx.y, mx.z, mp.pr, mp.pi, mp.x, mp.y, mp.z};
          for (int c = 0; c < 10; ++c)
            if (!bits_equal(w[c], v[c])) {
              if (bad < 10)
                printf("MASK MISMATCH 3d=%d trial=%d N=(%d,%d,%d) per=(%d,%d,%d) cell=(%d,%d,%d) comp=%d: %g vs %g\n",
                       (int)IS3D, trial, g.Nx, g.Ny, g.Nz, g.per_x, g.per_y, g.per_z, i, j, k, c, w[c], v[c]);
              ++bad;
            }
        }
      }
  return bad ? -bad : cells;
}

int main() {
  std::mt19937 rng(12345);
  long long cells = 0;
  for (int t = 0; t < 300; ++t) {
    const long long r = (t % 4 == 0) ? check<false>(rng, t) : check<true>(rng, t);
    if (r < 0) {
      printf("FAILED\n");
      return 1;
    }
    cells += r;
  }
  printf("OK %lld cells\n", cells);
  return 0;
}
