// Adjoint (vector-Jacobian) kernels of eval_f and of the B_z observable, FP64.
//
// The Jacobian-vector product of eval_f at x is a sum, over every output
// component of every interior node m, of terms coef * (perturbation of one
// component of one expanded-grid cell) -- the cells being the node, its six
// neighbours and the six diagonal cells of the curl cross terms, all read
// through the boundary rules (the ghost cells are affine in the interior
// values).  The transpose is applied in two steps:
//
//   k_vjp_cells   each node m accumulates w_m * coef per (target cell,
//                 component) -- its own cell, the six neighbours, the six
//                 diagonal cells: 41 planes, each (plane, m) written by node
//                 m's thread alone (no atomics, deterministic);
//   k_cells_gather  every full-grid cell sums what its (up to 13) source nodes
//                 left for it, in a fixed order: the cell adjoint A[cell][component];
//   k_bc_gather   each interior node q gathers A over the cells whose values
//                 come from q under the boundary rules of rhs.py:54-139 (its
//                 own cell, the face ghosts copied from it, periodic images),
//                 giving (J^T w)_q;
//   k_field_grad  the face ghosts' transverse links also carry the applied
//                 field flux (cx, cy, cz) of the ghost rules, so the same A
//                 gives d<w, F>/d(cx, cy, cz).
//
// k_vjp_cells also emits the parameter derivatives of <w, F>: kappa (the
// link operators scale with kappa^2, the supercurrent does not) and the
// superconducting mask sc of every interior node (FPSI's sc * (1 - |psi|^2)
// psi - (1 - sc) psi / tau, sparse_operators.py:202-226).

#include <hip/hip_runtime.h>
#include <math.h>

#include "tdgl_device.hpp"
#include "tdgl_launch.hpp"

namespace tdgl {

namespace {

enum { CR = 0, CI = 1, LX = 2, LY = 3, LZ = 4 };
// Target cells of a node's scatter, relative to the node: itself, -x, +x, -y,
// +y, -z, +z (all five components each: plane slot * 5 + component), then the
// diagonal cells of the curl cross terms (one component each: planes 35-40).
enum { S0 = 0, SXM = 1, SXP = 2, SYM = 3, SYP = 4, SZM = 5, SZP = 6 };
enum { D_XP_YM = 35, D_XP_ZM = 36, D_XM_YP = 37, D_YP_ZM = 38, D_XM_ZP = 39, D_YM_ZP = 40 };
constexpr int kVjpPlanes = 41;
__device__ __forceinline__ int plane(int slot, int comp) { return slot * 5 + comp; }

template <bool IS3D>
__device__ __forceinline__ void node_of(const Geo& g, long long q, int& i, int& j, int& k) {
  i = 1 + (int)(q % g.nx);
  const long long t = q / g.nx;
  j = 1 + (int)(t % g.ny);
  k = IS3D ? 1 + (int)(t / g.ny) : 0;
}

}  // namespace

// w: adjoint weights (internal layout).  A: full grid x 5, accumulated.
// gsc: per interior node (internal order) d<w,F>/d(sc_q), or nullptr.
// pk: per-block partials of d<w,F>/d(kappa) (uniform kappa `kappa`), or nullptr.
// w: adjoint weights (internal layout).  B: kVjpPlanes x n, every (plane, node)
// entry written: a node's 41 sums are accumulated in registers, in the order of
// the terms below, and stored once.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_vjp_cells(Geo g, Bnd b, Mat mt, const double* x,
                                                           const double* w, double* B, double* gsc,
                                                           double kappa, double* pk, const double* stop) {
  if (stop && *stop != 0.0) return;  // a queued TGCR iteration after the loop ended
  const long long n = g.n;
  const Src<double, false> s{x, nullptr, 0.0, n};
  double kap[1] = {0.0};
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < n;
       q += (long long)gridDim.x * kRedThreads) {
    int i, j, k;
    node_of<IS3D>(g, q, i, j, k);
    const double2 w0 = reinterpret_cast<const double2*>(w)[q];
    const double mr = w0.x, mi = w0.y;
    const uint8_t fl = mt.flags ? mt.flags[mfull(g, i, j, k)] : (uint8_t)0;
    const double mx = (fl & kHoleX) ? 0.0 : w[2 * n + q];
    const double my = (fl & kHoleY) ? 0.0 : w[3 * n + q];
    const double mz = (IS3D && !(fl & kHoleZ)) ? w[4 * n + q] : 0.0;

    auto cell = [&](int ii, int jj, int kk) { return eff_cell<double, IS3D, false>(g, b, mt.flags, s, ii, jj, kk); };
    // the node's entries of the planes (B[plane * n + q] belongs to this thread)
    double bp[kVjpPlanes];
#pragma unroll
    for (int pl = 0; pl < kVjpPlanes; ++pl) bp[pl] = 0.0;
    auto acc = [&](int pl, double v) { bp[pl] += v; };
    const Cell<double> c0 = cell(i, j, k), xm = cell(i - 1, j, k), xp = cell(i + 1, j, k);
    const Cell<double> ym = cell(i, j - 1, k), yp = cell(i, j + 1, k);
    Cell<double> zm{}, zp{};
    if (IS3D) {
      zm = cell(i, j, k - 1);
      zp = cell(i, j, k + 1);
    }

    const double pr = c0.pr, pi = c0.pi;
    double sxp, cxp, sxm, cxm, syp, cyp, sym, cym, szp = 0, czp = 1, szm = 0, czm = 1;
    sincos(c0.x, &sxp, &cxp);
    sincos(xm.x, &sxm, &cxm);
    sincos(c0.y, &syp, &cyp);
    sincos(ym.y, &sym, &cym);
    if (IS3D) {
      sincos(c0.z, &szp, &czp);
      sincos(zm.z, &szm, &czm);
    }

    // ---- psi equation: d psi_m = dgpsi dpsi_0 + sum of the six hopping terms + FPSI'
    acc(plane(S0, CR), mt.dgpsi * mr);
    acc(plane(S0, CI), mt.dgpsi * mi);
    // minus neighbour nb = m - a, its link t = phi_a(m - a):
    //   w_r = dq_r + dt q_i, w_i = dq_i - dt q_r;  d(ar, ai) = h (c w_r + s w_i, c w_i - s w_r)
    auto minus = [&](double c, double sn, double h, const Cell<double>& nb, int snb, int lc) {
      const double gr = h * (c * mr - sn * mi), gi = h * (sn * mr + c * mi);
      acc(plane(snb, CR), gr);
      acc(plane(snb, CI), gi);
      acc(plane(snb, lc), gr * nb.pi - gi * nb.pr);
    };
    // plus neighbour nb = m + a, the node's own link t = phi_a(m):
    //   w_r = dq_r - dt q_i, w_i = dq_i + dt q_r;  d(ar, ai) = h (c w_r - s w_i, c w_i + s w_r)
    auto plus = [&](double c, double sn, double h, const Cell<double>& nb, int snb, int lc) {
      const double gr = h * (c * mr + sn * mi), gi = h * (-sn * mr + c * mi);
      acc(plane(snb, CR), gr);
      acc(plane(snb, CI), gi);
      acc(plane(S0, lc), -gr * nb.pi + gi * nb.pr);
    };
    minus(cxm, sxm, mt.ihx2, xm, SXM, LX);
    plus(cxp, sxp, mt.ihx2, xp, SXP, LX);
    minus(cym, sym, mt.ihy2, ym, SYM, LY);
    plus(cyp, syp, mt.ihy2, yp, SYP, LY);
    if (IS3D) {
      minus(czm, szm, mt.ihz2, zm, SZM, LZ);
      plus(czp, szp, mt.ihz2, zp, SZP, LZ);
    }
    // FPSI: [fr, fi] = M [dr, di], M symmetric
    const double m2 = pr * pr + pi * pi, a2 = pr * pr - pi * pi, b2 = 2.0 * pr * pi;
    double sc = 1.0;
    if (mt.has_mat) sc = mt.sc_bits ? ((fl & kInsulator) ? 0.0 : 1.0) : mt.sc[q];
    const double tau = mt.has_mat ? (1.0 - sc) * mt.inv_tau : 0.0;
    const double M00 = sc * ((1.0 - 2.0 * m2) - a2) - tau, M11 = sc * ((1.0 - 2.0 * m2) + a2) - tau;
    const double M01 = -sc * b2;
    acc(plane(S0, CR), M00 * mr + M01 * mi);
    acc(plane(S0, CI), M01 * mr + M11 * mi);
    if (gsc) {  // d FPSI / d sc = (1 - |psi|^2) psi + psi / tau
      const double t1 = 1.0 - m2;
      gsc[q] = mr * (t1 * pr + pr * mt.inv_tau) + mi * (t1 * pi + pi * mt.inv_tau);
    }

    // ---- link equations
    double cx, cy, cz;
    if (mt.k2) {
      const double k2 = mt.k2[q];
      cx = k2 / mt.hx2;
      cy = k2 / mt.hy2;
      cz = IS3D ? k2 / mt.hz2 : 0.0;
    } else {
      cx = mt.cx;
      cy = mt.cy;
      cz = mt.cz;
    }
    // supercurrent derivative along link a with own phase (c, sn), neighbour nb = m + a:
    //   (c U_i - s U_r) - dt (c T_r + s T_i)
    auto dsup_t = [&](double c, double sn, double wa, const Cell<double>& nb, int snb, int lc) {
      if (wa == 0.0) return;
      const double qr = nb.pr, qi = nb.pi;
      const double Tr = pr * qr + pi * qi, Ti = pr * qi - pi * qr;
      const double gUi = c * wa, gUr = -sn * wa;
      acc(plane(S0, CR), gUr * qr + gUi * qi);
      acc(plane(S0, CI), gUr * qi - gUi * qr);
      acc(plane(snb, CR), gUr * pr - gUi * pi);
      acc(plane(snb, CI), gUr * pi + gUi * pr);
      acc(plane(S0, lc), -(c * Tr + sn * Ti) * wa);
    };
    // the kappa^2 part of each link equation, for d/d kappa
    double kx = 0.0, ky = 0.0, kzv = 0.0;
    if (mx != 0.0) {
      // jx = 2cx (x[xm] - 2 x0 + x[xp]) + cy (x[ym] - 2 x0 + x[yp]) + cy ((-y[xp] + y0) + y[xp-y] - y[ym])
      //    + cz (x[zm] - 2 x0 + x[zp]) + cz ((-z[xp] + z0) + z[xp-z] - z[zm]) + dsup_x
      const double c2x = 2.0 * cx;
      acc(plane(SXM, LX), c2x * mx);
      acc(plane(SXP, LX), c2x * mx);
      acc(plane(SYM, LX), cy * mx);
      acc(plane(SYP, LX), cy * mx);
      acc(plane(S0, LX), (-2.0 * c2x - 2.0 * cy - (IS3D ? 2.0 * cz : 0.0)) * mx);
      acc(plane(SXP, LY), -cy * mx);
      acc(plane(S0, LY), cy * mx);
      acc(D_XP_YM, cy * mx);
      acc(plane(SYM, LY), -cy * mx);
      const Cell<double> dxy = cell(i + 1, j - 1, k);
      kx = c2x * (xm.x - 2.0 * c0.x + xp.x) + cy * (ym.x - 2.0 * c0.x + yp.x) +
           cy * (((-xp.y + c0.y) + dxy.y) - ym.y);
      if (IS3D) {
        acc(plane(SZM, LX), cz * mx);
        acc(plane(SZP, LX), cz * mx);
        acc(plane(SXP, LZ), -cz * mx);
        acc(plane(S0, LZ), cz * mx);
        acc(D_XP_ZM, cz * mx);
        acc(plane(SZM, LZ), -cz * mx);
        const Cell<double> dxz = cell(i + 1, j, k - 1);
        kx += cz * (zm.x - 2.0 * c0.x + zp.x) + cz * (((-xp.z + c0.z) + dxz.z) - zm.z);
      }
      dsup_t(cxp, sxp, mx, xp, SXP, LX);
    }
    if (my != 0.0) {
      // jy = 2cy (y[ym] - 2 y0 + y[yp]) + cx (y[xm] - 2 y0 + y[xp]) + cx ((-x[yp] + x0) + x[yp-x] - x[xm])
      //    + cz (y[zm] - 2 y0 + y[zp]) + cz ((-z[yp] + z0) + z[yp-z] - z[zm]) + dsup_y
      const double c2y = 2.0 * cy;
      acc(plane(SYM, LY), c2y * my);
      acc(plane(SYP, LY), c2y * my);
      acc(plane(SXM, LY), cx * my);
      acc(plane(SXP, LY), cx * my);
      acc(plane(S0, LY), (-2.0 * c2y - 2.0 * cx - (IS3D ? 2.0 * cz : 0.0)) * my);
      acc(plane(SYP, LX), -cx * my);
      acc(plane(S0, LX), cx * my);
      acc(D_XM_YP, cx * my);
      acc(plane(SXM, LX), -cx * my);
      const Cell<double> dyx = cell(i - 1, j + 1, k);
      ky = c2y * (ym.y - 2.0 * c0.y + yp.y) + cx * (xm.y - 2.0 * c0.y + xp.y) +
           cx * (((-yp.x + c0.x) + dyx.x) - xm.x);
      if (IS3D) {
        acc(plane(SZM, LY), cz * my);
        acc(plane(SZP, LY), cz * my);
        acc(plane(SYP, LZ), -cz * my);
        acc(plane(S0, LZ), cz * my);
        acc(D_YP_ZM, cz * my);
        acc(plane(SZM, LZ), -cz * my);
        const Cell<double> dyz = cell(i, j + 1, k - 1);
        ky += cz * (zm.y - 2.0 * c0.y + zp.y) + cz * (((-yp.z + c0.z) + dyz.z) - zm.z);
      }
      dsup_t(cyp, syp, my, yp, SYP, LY);
    }
    if (IS3D && mz != 0.0) {
      // jz = 2cz (z[zm] - 2 z0 + z[zp]) + cx (z[xm] - 2 z0 + z[xp]) + cy (z[ym] - 2 z0 + z[yp])
      //    + cx ((-x[zp] + x0) + x[zp-x] - x[xm]) + cy ((-y[zp] + y0) + y[zp-y] - y[ym]) + dsup_z
      const double c2z = 2.0 * cz;
      acc(plane(SZM, LZ), c2z * mz);
      acc(plane(SZP, LZ), c2z * mz);
      acc(plane(SXM, LZ), cx * mz);
      acc(plane(SXP, LZ), cx * mz);
      acc(plane(SYM, LZ), cy * mz);
      acc(plane(SYP, LZ), cy * mz);
      acc(plane(S0, LZ), (-2.0 * c2z - 2.0 * cx - 2.0 * cy) * mz);
      acc(plane(SZP, LX), -cx * mz);
      acc(plane(S0, LX), cx * mz);
      acc(D_XM_ZP, cx * mz);
      acc(plane(SXM, LX), -cx * mz);
      acc(plane(SZP, LY), -cy * mz);
      acc(plane(S0, LY), cy * mz);
      acc(D_YM_ZP, cy * mz);
      acc(plane(SYM, LY), -cy * mz);
      const Cell<double> dzx = cell(i - 1, j, k + 1), dzy = cell(i, j - 1, k + 1);
      kzv = c2z * (zm.z - 2.0 * c0.z + zp.z) + cx * (xm.z - 2.0 * c0.z + xp.z) +
            cy * (ym.z - 2.0 * c0.z + yp.z) + cx * (((-zp.x + c0.x) + dzx.x) - xm.x) +
            cy * (((-zp.y + c0.y) + dzy.y) - ym.y);
      dsup_t(czp, szp, mz, zp, SZP, LZ);
    }
    if (pk) {  // the link operators are proportional to kappa^2
      const double kq = mt.k2 ? sqrt(mt.k2[q]) : kappa;
      kap[0] += (2.0 / kq) * ((mx * kx + my * ky) + mz * kzv);
    }
#pragma unroll
    for (int pl = 0; pl < kVjpPlanes; ++pl) B[(long long)pl * n + q] = bp[pl];
  }
  if (pk) publish<1, 0>(kap, pk);
}

// Cell adjoint A[cell][component] of the full grid: each cell sums what its
// source nodes (cell - offset, interior only) left in the planes of k_vjp_cells,
// in a fixed order (the six neighbours' and the node's own slots, then the
// diagonals).  Every (plane, node) entry is read by exactly one cell
// (k_vjp_cells writes them all on every call).
template <bool IS3D>
__device__ __forceinline__ void gather_cell(const Geo& g, const double* B, int ci, int cj, int ck,
                                            double (&a)[5]) {
  const long long n = g.n;
#pragma unroll
  for (int comp = 0; comp < 5; ++comp) a[comp] = 0.0;
  // the node that scattered into this cell with offset (di, dj, dk), if interior
  auto src = [&](int di, int dj, int dk, long long& q) {
    const int si = ci - di, sj = cj - dj, sk = ck - dk;
    if (si < 1 || si > g.Nx - 1 || sj < 1 || sj > g.Ny - 1) return false;
    if (IS3D && (sk < 1 || sk > g.Nz - 1)) return false;
    q = qix(g, si, sj, sk);
    return true;
  };
  constexpr int off[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  auto take = [&](int pl, long long q) { return B[(long long)pl * n + q]; };
#pragma unroll
  for (int s = 0; s < (IS3D ? 7 : 5); ++s) {
    long long q;
    if (src(off[s][0], off[s][1], off[s][2], q))
#pragma unroll
      for (int comp = 0; comp < 5; ++comp) a[comp] += take(plane(s, comp), q);
  }
  long long q;
  if (src(1, -1, 0, q)) a[LY] += take(D_XP_YM, q);
  if (src(-1, 1, 0, q)) a[LX] += take(D_XM_YP, q);
  if (IS3D) {
    if (src(1, 0, -1, q)) a[LZ] += take(D_XP_ZM, q);
    if (src(0, 1, -1, q)) a[LZ] += take(D_YP_ZM, q);
    if (src(-1, 0, 1, q)) a[LX] += take(D_XM_ZP, q);
    if (src(0, -1, 1, q)) a[LY] += take(D_YM_ZP, q);
  }
}

template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_cells_gather(Geo g, const double* B, double* A,
                                                              const double* stop) {
  if (stop && *stop != 0.0) return;
  const long long dimx = g.mk * (IS3D ? (long long)(g.Nz + 1) : 1LL);
  for (long long c = (long long)blockIdx.x * kRedThreads + threadIdx.x; c < dimx;
       c += (long long)gridDim.x * kRedThreads) {
    const int ci = (int)(c % g.mj);
    const long long t = c / g.mj;
    const int cj = (int)(t % (g.Ny + 1));
    const int ck = IS3D ? (int)(t / (g.Ny + 1)) : 0;
    double a[5];
    gather_cell<IS3D>(g, B, ci, cj, ck, a);
#pragma unroll
    for (int comp = 0; comp < 5; ++comp) A[c * 5 + comp] = a[comp];
  }
}

// (J^T w)_q: the cell adjoint gathered over the cells whose values come from q
// (its own cell and the face ghosts copied from it); cell(ii, jj, kk, a) gives
// a cell's five components.  Then J^T w is stored, or (ch.z) the recurrence
// step formed from it.
template <bool IS3D, class CellA>
__device__ __forceinline__ void bc_node(const Geo& g, const uint8_t* flags, long long q, double* out,
                                        const AdjCheb& ch, CellA cell) {
  const long long n = g.n;
  int i, j, k;
  node_of<IS3D>(g, q, i, j, k);
  auto F = [&](int ii, int jj, int kk) { return flags ? flags[mfull(g, ii, jj, kk)] : (uint8_t)0; };
  const uint8_t f0 = F(i, j, k);
  double a0[5];
  cell(i, j, k, a0);
  double gr = a0[CR], gi = a0[CI];
  double gx = (i != g.Nx - 1 && !(f0 & kHoleX)) ? a0[LX] : 0.0;
  double gy = (j != g.Ny - 1 && !(f0 & kHoleY)) ? a0[LY] : 0.0;
  double gz = (IS3D && k != g.Nz - 1 && !(f0 & kHoleZ)) ? a0[LZ] : 0.0;
  double ag[5];
  // x faces
  for (int side = 0; side < 2; ++side) {
    const int gi_ = side == 0 ? 0 : g.Nx;
    const int src = g.per_x ? (side == 0 ? g.Nx - 1 : 1) : (side == 0 ? 1 : g.Nx - 1);
    if (i != src) continue;
    const uint8_t fg = F(gi_, j, k);
    cell(gi_, j, k, ag);
    gr += ag[CR];
    gi += ag[CI];
    if (g.per_x) {
      if (side == 1 && g.Nx - 1 != 1 && !(fg & kHoleX)) gx += ag[LX];
    } else {
      if (j != g.Ny - 1 && !(fg & kHoleY)) gy += ag[LY];
      if (IS3D && k != g.Nz - 1 && !(fg & kHoleZ)) gz += ag[LZ];
    }
  }
  // y faces
  for (int side = 0; side < 2; ++side) {
    const int gj = side == 0 ? 0 : g.Ny;
    const int src = g.per_y ? (side == 0 ? g.Ny - 1 : 1) : (side == 0 ? 1 : g.Ny - 1);
    if (j != src) continue;
    const uint8_t fg = F(i, gj, k);
    cell(i, gj, k, ag);
    gr += ag[CR];
    gi += ag[CI];
    if (g.per_y) {
      if (side == 1 && g.Ny - 1 != 1 && !(fg & kHoleY)) gy += ag[LY];
    } else {
      if (i != g.Nx - 1 && !(fg & kHoleX)) gx += ag[LX];
      if (IS3D && k != g.Nz - 1 && !(fg & kHoleZ)) gz += ag[LZ];
    }
  }
  // z faces
  if (IS3D) {
    for (int side = 0; side < 2; ++side) {
      const int gk = side == 0 ? 0 : g.Nz;
      const int src = g.per_z ? (side == 0 ? g.Nz - 1 : 1) : (side == 0 ? 1 : g.Nz - 1);
      if (k != src) continue;
      const uint8_t fg = F(i, j, gk);
      cell(i, j, gk, ag);
      gr += ag[CR];
      gi += ag[CI];
      if (g.per_z) {
        if (side == 1 && g.Nz - 1 != 1 && !(fg & kHoleZ)) gz += ag[LZ];
      } else {
        if (i != g.Nx - 1 && !(fg & kHoleX)) gx += ag[LX];
        if (j != g.Ny - 1 && !(fg & kHoleY)) gy += ag[LY];
      }
    }
  }
  if (ch.z) {  // the recurrence step, J^T zp = (gr, gi, gx, gy, gz)
    auto step = [&](long long e, double gv) {
      const double a = ch.zp[e] - ch.hdt * gv;
      double v = ch.om * (ch.zp[e] + (ch.r[e] - a) * ch.thinv);
      if (ch.zpp) v = v - (ch.om - 1.0) * ch.zpp[e];
      ch.z[e] = v;
    };
    step(2 * q, gr);
    step(2 * q + 1, gi);
    step(2 * n + q, gx);
    step(3 * n + q, gy);
    if (IS3D) step(4 * n + q, gz);
  } else {
    reinterpret_cast<double2*>(out)[q] = make_double2(gr, gi);
    out[2 * n + q] = gx;
    out[3 * n + q] = gy;
    if (IS3D) out[4 * n + q] = gz;
  }
}

// J^T w from the materialised cell adjoint A (k_cells_gather).
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_bc_gather(Geo g, const uint8_t* flags,
                                                           const double* A, double* out,
                                                           const double* stop, AdjCheb ch) {
  if (stop && *stop != 0.0) return;
  auto cell = [&](int ii, int jj, int kk, double (&a)[5]) {
    const long long c = mfull(g, ii, jj, kk);
#pragma unroll
    for (int comp = 0; comp < 5; ++comp) a[comp] = A[c * 5 + comp];
  };
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < g.n;
       q += (long long)gridDim.x * kRedThreads)
    bc_node<IS3D>(g, flags, q, out, ch, cell);
}

// J^T w straight from the planes: each node gathers the cells it needs (its
// own, and at a face the ghosts copied from it) itself, so the cell adjoint is
// never stored.  Each of those cells is needed by this node only, and each is
// summed as k_cells_gather sums it: the same bits as the two kernels.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_planes_gather(Geo g, const uint8_t* flags,
                                                               const double* B, double* out,
                                                               const double* stop, AdjCheb ch) {
  if (stop && *stop != 0.0) return;
  auto cell = [&](int ii, int jj, int kk, double (&a)[5]) { gather_cell<IS3D>(g, B, ii, jj, kk, a); };
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < g.n;
       q += (long long)gridDim.x * kRedThreads)
    bc_node<IS3D>(g, flags, q, out, ch, cell);
}

// d/d(cx, cy, cz) of the face ghosts' transverse links (rhs.py:91-98, 107-112,
// 122-127), weighted by the cell adjoint; per-block partials.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_field_grad(Geo g, const uint8_t* flags,
                                                            const double* A, double* pout) {
  double v[3] = {0.0, 0.0, 0.0};  // d/dcx, d/dcy, d/dcz
  const long long dimx = g.mk * (IS3D ? (long long)(g.Nz + 1) : 1LL);
  for (long long m = (long long)blockIdx.x * kRedThreads + threadIdx.x; m < dimx;
       m += (long long)gridDim.x * kRedThreads) {
    const int i = (int)(m % g.mj);
    const long long t = m / g.mj;
    const int j = (int)(t % (g.Ny + 1));
    const int k = IS3D ? (int)(t / (g.Ny + 1)) : 0;
    const bool bi = (i == 0) || (i == g.Nx);
    const bool bj = (j == 0) || (j == g.Ny);
    const bool bk = IS3D && ((k == 0) || (k == g.Nz));
    if ((int)bi + (int)bj + (int)bk != 1) continue;
    const uint8_t fg = flags ? flags[m] : (uint8_t)0;
    const double ax = (fg & kHoleX) ? 0.0 : A[m * 5 + LX];
    const double ay = (fg & kHoleY) ? 0.0 : A[m * 5 + LY];
    const double az = (IS3D && !(fg & kHoleZ)) ? A[m * 5 + LZ] : 0.0;
    if (bi && !g.per_x) {
      v[2] += (i == 0 ? -1.0 : 1.0) * ay;
      if (IS3D) v[1] += (i == 0 ? 1.0 : -1.0) * az;
    } else if (bj && !g.per_y) {
      v[2] += (j == 0 ? 1.0 : -1.0) * ax;
      if (IS3D) v[0] += (j == 0 ? -1.0 : 1.0) * az;
    } else if (bk && !g.per_z) {
      v[1] += (k == 0 ? -1.0 : 1.0) * ax;
      v[0] += (k == 0 ? 1.0 : -1.0) * ay;
    }
  }
  publish<3, 0>(v, pout);
}

// Cell adjoint of B_z = (((x - x[+y]) - y) + y[+x]) / (hx hy) at every
// interior node (physics/bfield.py:58-68), weights in the reference order.
// Gathered per full-grid cell (no atomics, fixed order): the x link of a cell
// takes +v from its own node and -v from the node below it in y; the y link
// takes -v from its own node and +v from the node before it in x.  Every cell
// and component is written.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_bz_cells(Geo g, Bnd b, const double* gbz, double* A) {
  const long long dimx = g.mk * (IS3D ? (long long)(g.Nz + 1) : 1LL);
  const double s = 1.0 / (b.hx * b.hy);
  // s * gbz at interior node (i, j, k), reference order r = ((i-1) ny + (j-1)) nz + (k-1)
  auto v = [&](int i, int j, int k, double& out) {
    if (i < 1 || i > g.Nx - 1 || j < 1 || j > g.Ny - 1) return false;
    if (IS3D && (k < 1 || k > g.Nz - 1)) return false;
    out = s * gbz[((long long)(i - 1) * g.ny + (j - 1)) * g.nz + (IS3D ? k - 1 : 0)];
    return true;
  };
  for (long long c = (long long)blockIdx.x * kRedThreads + threadIdx.x; c < dimx;
       c += (long long)gridDim.x * kRedThreads) {
    const int ci = (int)(c % g.mj);
    const long long t = c / g.mj;
    const int cj = (int)(t % (g.Ny + 1));
    const int ck = IS3D ? (int)(t / (g.Ny + 1)) : 0;
    double ax = 0.0, ay = 0.0, w;
    if (v(ci, cj, ck, w)) {
      ax += w;
      ay -= w;
    }
    if (v(ci, cj - 1, ck, w)) ax -= w;
    if (v(ci - 1, cj, ck, w)) ay += w;
    A[c * 5 + CR] = 0.0;
    A[c * 5 + CI] = 0.0;
    A[c * 5 + LX] = ax;
    A[c * 5 + LY] = ay;
    A[c * 5 + LZ] = 0.0;
  }
}

// The adjoint step operator inside TGCR: Ap = r - hdt (J^T r) in place of
// Ap = J^T r, and the partials of <Ap, Ap_0> (or ||Ap||^2 without Ap_0) that
// the first Gram-Schmidt pass (or k_final_p) consumes -- what the FD-JVP
// stencil's epilogue gives the forward solve.
__global__ __launch_bounds__(kRedThreads) void k_adj_combine(long long len, const double* r, double* ap,
                                                             double hdt, const double* ap0, Red rd) {
  if (rd.res[RS_STOP] != 0.0) return;  // a queued iteration after the loop ended
  double v[1] = {0.0};
  for (long long e = (long long)blockIdx.x * kRedThreads + threadIdx.x; e < len;
       e += (long long)gridDim.x * kRedThreads) {
    const double a = r[e] - hdt * ap[e];
    ap[e] = a;
    v[0] += a * (ap0 ? ap0[e] : a);
  }
  publish<1, 0>(v, rd.pout);
}

// y = s x (len doubles; y may alias x).
__global__ __launch_bounds__(kRedThreads) void k_scale(long long len, double* y, double s, const double* x) {
  for (long long e = (long long)blockIdx.x * kRedThreads + threadIdx.x; e < len;
       e += (long long)gridDim.x * kRedThreads)
    y[e] = s * x[e];
}

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------

int launch_adj_combine(const Geo& g, const double* r, double* ap, double hdt, const double* ap0,
                       const Red& rd, hipStream_t st) {
  const long long len = (long long)(2 + g.nphi) * g.n;
  launch(k_adj_combine, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, len, r, ap, hdt, ap0, rd);
  return stream_blocks(g);
}
void launch_scale(const Geo& g, double* y, double s, const double* x, hipStream_t st) {
  const long long len = (long long)(2 + g.nphi) * g.n;
  launch(k_scale, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, len, y, s, x);
}

int vjp_planes() { return kVjpPlanes; }
void launch_vjp_cells(const Geo& g, const Bnd& b, const Mat& mt, const double* x, const double* w,
                      double* B, double* gsc, double kappa, double* pk, hipStream_t st, const double* stop) {
  const dim3 gr(stream_blocks(g)), bl(kRedThreads);
  if (g.is3d) launch(k_vjp_cells<true>, gr, bl, 0, st, g, b, mt, x, w, B, gsc, kappa, pk, stop);
  else launch(k_vjp_cells<false>, gr, bl, 0, st, g, b, mt, x, w, B, gsc, kappa, pk, stop);
}
void launch_cells_gather(const Geo& g, const double* B, double* A, hipStream_t st, const double* stop) {
  const long long dimx = g.mk * (g.is3d ? (long long)(g.Nz + 1) : 1LL);
  long long nb = (dimx + kRedThreads - 1) / kRedThreads;
  if (nb > 4096) nb = 4096;
  if (g.is3d) launch(k_cells_gather<true>, dim3((unsigned)nb), dim3(kRedThreads), 0, st, g, B, A, stop);
  else launch(k_cells_gather<false>, dim3((unsigned)nb), dim3(kRedThreads), 0, st, g, B, A, stop);
}
void launch_bc_gather(const Geo& g, const uint8_t* flags, const double* A, double* out, hipStream_t st,
                      const double* stop, const AdjCheb& ch) {
  const dim3 gr(stream_blocks(g)), bl(kRedThreads);
  if (g.is3d) launch(k_bc_gather<true>, gr, bl, 0, st, g, flags, A, out, stop, ch);
  else launch(k_bc_gather<false>, gr, bl, 0, st, g, flags, A, out, stop, ch);
}
void launch_planes_gather(const Geo& g, const uint8_t* flags, const double* B, double* out, hipStream_t st,
                          const double* stop, const AdjCheb& ch) {
  const dim3 gr(stream_blocks(g)), bl(kRedThreads);
  if (g.is3d) launch(k_planes_gather<true>, gr, bl, 0, st, g, flags, B, out, stop, ch);
  else launch(k_planes_gather<false>, gr, bl, 0, st, g, flags, B, out, stop, ch);
}
int launch_field_grad(const Geo& g, const uint8_t* flags, const double* A, double* pout, hipStream_t st) {
  const dim3 gr(stream_blocks(g)), bl(kRedThreads);
  if (g.is3d) launch(k_field_grad<true>, gr, bl, 0, st, g, flags, A, pout);
  else launch(k_field_grad<false>, gr, bl, 0, st, g, flags, A, pout);
  return stream_blocks(g);
}
void launch_bz_cells(const Geo& g, const Bnd& b, const double* gbz, double* A, hipStream_t st) {
  const long long dimx = g.mk * (g.is3d ? (long long)(g.Nz + 1) : 1LL);
  long long nb = (dimx + kRedThreads - 1) / kRedThreads;
  if (nb > 4096) nb = 4096;
  if (g.is3d) launch(k_bz_cells<true>, dim3((unsigned)nb), dim3(kRedThreads), 0, st, g, b, gbz, A);
  else launch(k_bz_cells<false>, dim3((unsigned)nb), dim3(kRedThreads), 0, st, g, b, gbz, A);
}

}  // namespace tdgl
