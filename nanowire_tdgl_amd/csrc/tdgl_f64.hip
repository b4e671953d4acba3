// FP64 kernels of the trapezoidal TDGL step (plain method).
//
// Compiled with -ffp-contract=off: every product and sum is rounded on its
// own, in the order tdgl3d's numpy/scipy code performs it, so the GPU result
// differs from the reference only by the ulp-level differences of sin/cos and
// of reduction order.
//
// Kernels
//   k_stencil  : the fused RHS (psi Laplacian through the Peierls phases,
//                link curl-curl, supercurrent, relaxation term, ghost rules,
//                hole masking) with five epilogues: plain RHS, trapezoidal
//                predictor, Newton residual, finite-difference JVP, Euler.
//                xy tile of 64x4 nodes + 1-node halo staged in LDS, marching
//                through z with a 3-plane ring.
//   k_mgs_*, k_final_p, k_update_r, k_accum : TGCR with modified Gram-Schmidt
//                (tgcr.py:135-152), vectors streamed node by node.
//   k_permute_* : reference complex128 state <-> internal layout.
//   k_obs      : B = curl(phi) and J_s (check-path observables).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>

#include "tdgl_device.hpp"
#include "tdgl_launch.hpp"
#include "tdgl_stage32.hpp"

namespace tdgl {

constexpr int TX = 64, TY = 4;
constexpr int CX = TX + 2, CY = TY + 2, CELLS = CX * CY;
constexpr int BS = TX * TY;
static_assert(BS == kRedThreads, "stencil block must match the reduction block");

// complex multiply-accumulate with numpy's formula (ar*br - ai*bi, ar*bi + ai*br)
__device__ __forceinline__ void cmac(double& sr, double& si, double cr, double ci, double vr,
                                     double vi) {
  const double tr = cr * vr - ci * vi;
  const double ti = cr * vi + ci * vr;
  sr = sr + tr;
  si = si + ti;
}

// Im[e^{-i phi} * conj(psi) * psi1] evaluated left to right as numpy does
// (sparse_operators.py:245, :275, :308).
__device__ __forceinline__ double supercurrent(double cs, double sn, double pr, double pi,
                                               double qr, double qi) {
  const double er = cs, ei = -sn;        // exp(-1j*phi)
  const double cr = pr, ci = -pi;        // conj(psi)
  const double tr = er * cr - ei * ci;
  const double ti = er * ci + ei * cr;
  return tr * qi + ti * qr;              // imag(t * psi1)
}

// Raw interior values a stencil epilogue reads at its own node, loaded before
// the next plane's halo so that they do not wait behind it.
struct Ctr {
  double xr, xi, xx, xy, xz;  // PRED/EULER/NEWTON: x;  JVP: p
  double gr, gi, gx, gy, gz;  // NEWTON: gamma;  JVP: f0
  double ar, ai, ax, ay, az;  // JVP with ap0: Ap_0
  double sc, k2;
  uint8_t fl;
};

template <int MODE, bool PERT>
struct RedN {
  static constexpr bool NEWTON = (MODE == M_NEWTON) || (MODE == M_NEWTON32);
  static constexpr int NS = NEWTON ? 2 : (MODE == M_JVP ? 1 : 0);
  static constexpr int NM = NEWTON ? 1 : 0;
  static constexpr int NV = NS + NM > 0 ? NS + NM : 1;
};

// One FP64 output vector on 32-bit byte offsets (as V64).
struct W64 {
  char* base;
  unsigned int pl;  // 8 n
  __device__ __forceinline__ void psi(unsigned int q, double2 v) const { stb<double2>(base, 16u * q, v); }
  __device__ __forceinline__ void ph(int a, unsigned int q, double v) const {
    stb<double>(base, (unsigned int)a * pl + 8u * q, v);
  }
};
__device__ __forceinline__ W64 wplanes64(double* v, unsigned int n) {
  return W64{reinterpret_cast<char*>(v), 8u * n};
}

// One work item of the fused stencil: an xy tile of TX x TY nodes, planes
// [kb, ke) of a z chunk.  The tile and its 1-node halo are staged in LDS with
// a 3-plane ring; the halo of plane k+2 is loaded into registers while plane
// k is computed (software pipelining), and committed to the ring after.
// Cells are staged by the column masks of stage64u (uniform field: the same
// bits as eff_cell, held in vector registers across the z march); GEN (some
// film has general field vectors) stages them with eff_cell, its column
// booleans recomputed on every plane.  Own-node loads and stores use 32-bit
// byte offsets from one base per vector (the host keeps 5 (8 n) < 2^32).
template <bool IS3D, bool PERT, int MODE, bool GEN>
__device__ __forceinline__ void stencil_tile(const Geo& g, const Bnd& b, const Mat& mt,
                                             const Src<double, PERT>& s, const Epi& e, int tile_x,
                                             int tile_y, int zc, int kz,
                                             double (&red)[RedN<MODE, PERT>::NV]) {
  constexpr int NSLOT = IS3D ? 3 : 1;
  __shared__ double l_pr[NSLOT][CELLS], l_pi[NSLOT][CELLS], l_x[NSLOT][CELLS], l_y[NSLOT][CELLS],
      l_z[IS3D ? NSLOT : 1][IS3D ? CELLS : 1];
  constexpr bool NEWTON = RedN<MODE, PERT>::NEWTON;
  constexpr int NS = RedN<MODE, PERT>::NS;

  const int i0 = 1 + tile_x * TX, j0 = 1 + tile_y * TY;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = (i <= g.Nx - 1) && (j <= g.Ny - 1);
  int kb = 0, ke = 1;
  if (IS3D) {
    kb = 1 + zc * kz;
    ke = min(kb + kz, g.Nz);
  }
  const unsigned int n = (unsigned int)g.n;
  const V64 X = planes64(s.x, n), P = planes64(PERT ? s.p : s.x, n);

  // halo cells handled by this thread: c0 = tid, c1 = tid + BS (if < CELLS)
  const int hc0 = threadIdx.x, hc1 = threadIdx.x + BS;
  const bool has1 = hc1 < CELLS;
  Plan64 pl0{}, pl1{};
  if constexpr (!GEN) {
    pl0 = plan64(g, b, i0 - 1 + hc0 % CX, j0 - 1 + hc0 / CX);
    pl1 = plan64(g, b, i0 - 1 + hc1 % CX, j0 - 1 + hc1 / CX);
  }
  Cell<double> v0, v1;
  // GEN: each cell staged whole (eff_cell); otherwise a plane's loads
  // (load64u) and rules (finish64u) in one stage, except in the prologue,
  // whose three planes' loads all go out before the first rules
  const bool hf = mt.flags != nullptr;
  auto ld = [&](int kp, Raw64& x0, Raw64& x1) __attribute__((always_inline)) {
    const Pl64 pp = plane64<IS3D>(g, b, kp);
    load64u<IS3D, PERT>(mt.flags, X, P, pl0, pp, x0);
    if (has1) load64u<IS3D, PERT>(mt.flags, X, P, pl1, pp, x1);
  };
  auto fin = [&](int kp, const Raw64& x0, const Raw64& x1) __attribute__((always_inline)) {
    const Pl64 pp = plane64<IS3D>(g, b, kp);
    if (IS3D && pp.fk) {
      finish64u<IS3D, true, PERT>(x0, hf, s.eps, pl0, pp, v0);
      if (has1) finish64u<IS3D, true, PERT>(x1, hf, s.eps, pl1, pp, v1);
    } else {
      finish64u<IS3D, false, PERT>(x0, hf, s.eps, pl0, pp, v0);
      if (has1) finish64u<IS3D, false, PERT>(x1, hf, s.eps, pl1, pp, v1);
    }
  };
  auto stage = [&](int kp) __attribute__((always_inline)) {
    if constexpr (GEN) {
      unsigned int h0 = hc0, h1 = hc1;
      opaque(h0);
      opaque(h1);
      v0 = eff_cell<double, IS3D, PERT>(g, b, mt.flags, s, i0 - 1 + h0 % CX, j0 - 1 + h0 / CX, kp);
      if (has1)
        v1 = eff_cell<double, IS3D, PERT>(g, b, mt.flags, s, i0 - 1 + h1 % CX, j0 - 1 + h1 / CX, kp);
    } else {
      Raw64 x0, x1;
      ld(kp, x0, x1);
      fin(kp, x0, x1);
    }
  };
  const unsigned int m_ij = (unsigned int)i + (unsigned int)g.mj * (unsigned int)j;  // full-grid column
  auto write = [&](int slot) __attribute__((always_inline)) {
    l_pr[slot][hc0] = v0.pr;
    l_pi[slot][hc0] = v0.pi;
    l_x[slot][hc0] = v0.x;
    l_y[slot][hc0] = v0.y;
    if (IS3D) l_z[slot][hc0] = v0.z;
    if (has1) {
      l_pr[slot][hc1] = v1.pr;
      l_pi[slot][hc1] = v1.pi;
      l_x[slot][hc1] = v1.x;
      l_y[slot][hc1] = v1.y;
      if (IS3D) l_z[slot][hc1] = v1.z;
    }
  };
  auto commit = [&](int slot) __attribute__((always_inline)) { write(slot); };

  if (IS3D) {
    if constexpr (GEN) {
      stage(kb - 1);
      commit((kb - 1) % 3);
      stage(kb);
      commit(kb % 3);
      stage(kb + 1);
    } else {  // the three planes' loads together: one memory round trip before the first step
      Raw64 a0{}, a1{}, b0{}, b1{}, c0{}, c1{};
      ld(kb - 1, a0, a1);
      ld(kb, b0, b1);
      ld(kb + 1, c0, c1);
      fin(kb - 1, a0, a1);
      write((kb - 1) % 3);
      fin(kb, b0, b1);
      write(kb % 3);
      fin(kb + 1, c0, c1);  // committed at the first step
    }
  } else {
    stage(0);
  }

  for (int k = kb; k < ke; ++k) {
    commit(IS3D ? (k + 1) % 3 : 0);
    __syncthreads();

    // own-node inputs first, then the halo two planes ahead
    Ctr c{};
    unsigned int q = 0;
    auto ld5 = [&](const V64& v, double& a, double& bb, double& xx, double& yy, double& zz)
                   __attribute__((always_inline)) {
      const double2 t = v.psi(q);
      a = t.x;
      bb = t.y;
      xx = v.ph(2, q);
      yy = v.ph(3, q);
      zz = IS3D ? v.ph(4, q) : 0.0;
    };
    if (act) {
      q = qix32(g, i, j, k, IS3D);
      c.fl = mt.flags ? ldu(mt.flags, m_ij + (IS3D ? (unsigned int)g.mk * (unsigned int)k : 0u)) : (uint8_t)0;
      if (mt.has_mat && !mt.sc_bits) c.sc = ldu(mt.sc, q);
      if (mt.k2) c.k2 = ldu(mt.k2, q);
      if (MODE == M_PRED || MODE == M_EULER || NEWTON) ld5(X, c.xr, c.xi, c.xx, c.xy, c.xz);
      if (NEWTON) ld5(planes64(e.gam, n), c.gr, c.gi, c.gx, c.gy, c.gz);
      if (MODE == M_JVP) {
        ld5(P, c.xr, c.xi, c.xx, c.xy, c.xz);
        ld5(planes64(e.f0, n), c.gr, c.gi, c.gx, c.gy, c.gz);
        if (e.ap0) ld5(planes64(e.ap0, n), c.ar, c.ai, c.ax, c.ay, c.az);
      }
    }
    if (IS3D && k + 2 <= ke) stage(k + 2);

    if (act) {
      const int s0 = IS3D ? k % 3 : 0;
      const int sm = IS3D ? (k + 2) % 3 : 0;  // k - 1
      const int sp = IS3D ? (k + 1) % 3 : 0;  // k + 1
      const int c0 = (ty + 1) * CX + (tx + 1);
      const int cxm = c0 - 1, cxp = c0 + 1, cym = c0 - CX, cyp = c0 + CX;

      // psi and links at the node and its neighbours (post-BC values)
      const double pr = l_pr[s0][c0], pi = l_pi[s0][c0];
      const double ex = l_x[s0][c0], ey = l_y[s0][c0];
      const double ez = IS3D ? l_z[s0][c0] : 0.0;

      // Peierls factors e^{i phi_a(m)} and e^{-i phi_a(m - a)}
      double sxp, cxp_, sxm, cxm_, syp, cyp_, sym, cym_, szp = 0, czp = 1, szm = 0, czm = 1;
      sincos(ex, &sxp, &cxp_);
      sincos(l_x[s0][cxm], &sxm, &cxm_);
      sincos(ey, &syp, &cyp_);
      sincos(l_y[s0][cym], &sym, &cym_);
      if (IS3D) {
        sincos(ez, &szp, &czp);
        sincos(l_z[sm][c0], &szm, &czm);
      }

      // ---- dpsi: row m of (LPSIX/hx^2 + LPSIY/hy^2 + LPSIZ/hz^2) @ x, columns
      // ascending (m-mk, m-mj, m-1, m, m+1, m+mj, m+mk)   rhs.py:220
      double ar = 0.0, ai = 0.0;
      if (IS3D) cmac(ar, ai, czm * mt.ihz2, -szm * mt.ihz2, l_pr[sm][c0], l_pi[sm][c0]);
      cmac(ar, ai, cym_ * mt.ihy2, -sym * mt.ihy2, l_pr[s0][cym], l_pi[s0][cym]);
      cmac(ar, ai, cxm_ * mt.ihx2, -sxm * mt.ihx2, l_pr[s0][cxm], l_pi[s0][cxm]);
      cmac(ar, ai, mt.dgpsi, 0.0, pr, pi);
      cmac(ar, ai, cxp_ * mt.ihx2, sxp * mt.ihx2, l_pr[s0][cxp], l_pi[s0][cxp]);
      cmac(ar, ai, cyp_ * mt.ihy2, syp * mt.ihy2, l_pr[s0][cyp], l_pi[s0][cyp]);
      if (IS3D) cmac(ar, ai, czp * mt.ihz2, szp * mt.ihz2, l_pr[sp][c0], l_pi[sp][c0]);

      // FPSI (sparse_operators.py:202-226)
      const double ss = pr * pr - (-pi) * pi;  // conj(psi)*psi
      const double t1 = 1.0 - ss;
      double fr = t1 * pr - 0.0 * pi, fi = t1 * pi + 0.0 * pr;
      const uint8_t fl = c.fl;
      if (mt.has_mat) {
        double sc;
        if (mt.sc_bits) sc = (fl & kInsulator) ? 0.0 : 1.0;
        else sc = c.sc;
        const double w = 1.0 - sc;
        fr = sc * fr - (w * pr) * mt.inv_tau;
        fi = sc * fi - (w * pi) * mt.inv_tau;
      }
      const double dpr = ar + fr, dpi = ai + fi;

      // ---- link equations (rhs.py:223-225 + sparse_operators.py:229-317)
      double cx, cy, cz, dgx, dgy, dgz;
      if (mt.k2) {
        const double k2 = c.k2;
        cx = k2 / mt.hx2;
        cy = k2 / mt.hy2;
        cz = IS3D ? k2 / mt.hz2 : 0.0;
        dgx = (-2.0 * (cx + cz)) + (-2.0 * (cx + cy));
        dgy = (-2.0 * (cy + cz)) + (-2.0 * (cx + cy));
        dgz = (-2.0 * (cy + cz)) + (-2.0 * (cx + cz));
      } else {
        cx = mt.cx; cy = mt.cy; cz = mt.cz;
        dgx = mt.dgx; dgy = mt.dgy; dgz = mt.dgz;
      }
      const double c2x = cx + cx, c2y = cy + cy, c2z = cz + cz;

      // dphi_x = (LPHIY + LPHIZ) @ y1 + FPHIX
      double dx_ = 0.0;
      if (IS3D) dx_ = dx_ + cz * l_x[sm][c0];
      dx_ = dx_ + cy * l_x[s0][cym];
      dx_ = dx_ + c2x * l_x[s0][cxm];
      dx_ = dx_ + dgx * ex;
      dx_ = dx_ + c2x * l_x[s0][cxp];
      dx_ = dx_ + cy * l_x[s0][cyp];
      if (IS3D) dx_ = dx_ + cz * l_x[sp][c0];
      double crl = cy * (((-l_y[s0][cxp] + ey) + l_y[s0][cxp - CX]) - l_y[s0][cym]);
      if (IS3D) crl = crl + cz * (((-l_z[s0][cxp] + ez) + l_z[sm][cxp]) - l_z[sm][c0]);
      dx_ = dx_ + (crl + supercurrent(cxp_, sxp, pr, pi, l_pr[s0][cxp], l_pi[s0][cxp]));

      // dphi_y = (LPHIX + LPHIZ) @ y2 + FPHIY
      double dy_ = 0.0;
      if (IS3D) dy_ = dy_ + cz * l_y[sm][c0];
      dy_ = dy_ + c2y * l_y[s0][cym];
      dy_ = dy_ + cx * l_y[s0][cxm];
      dy_ = dy_ + dgy * ey;
      dy_ = dy_ + cx * l_y[s0][cxp];
      dy_ = dy_ + c2y * l_y[s0][cyp];
      if (IS3D) dy_ = dy_ + cz * l_y[sp][c0];
      crl = cx * (((-l_x[s0][cyp] + ex) + l_x[s0][cyp - 1]) - l_x[s0][cxm]);
      if (IS3D) crl = crl + cz * (((-l_z[s0][cyp] + ez) + l_z[sm][cyp]) - l_z[sm][c0]);
      dy_ = dy_ + (crl + supercurrent(cyp_, syp, pr, pi, l_pr[s0][cyp], l_pi[s0][cyp]));

      // dphi_z = (LPHIX + LPHIY) @ y3 + FPHIZ
      double dz_ = 0.0;
      if (IS3D) {
        dz_ = dz_ + c2z * l_z[sm][c0];
        dz_ = dz_ + cy * l_z[s0][cym];
        dz_ = dz_ + cx * l_z[s0][cxm];
        dz_ = dz_ + dgz * ez;
        dz_ = dz_ + cx * l_z[s0][cxp];
        dz_ = dz_ + cy * l_z[s0][cyp];
        dz_ = dz_ + c2z * l_z[sp][c0];
        crl = cx * (((-l_x[sp][c0] + ex) + l_x[sp][cxm]) - l_x[s0][cxm]);
        crl = crl + cy * (((-l_y[sp][c0] + ey) + l_y[sp][cym]) - l_y[s0][cym]);
        dz_ = dz_ + (crl + supercurrent(czp, szp, pr, pi, l_pr[sp][c0], l_pi[sp][c0]));
      }
      // hole-boundary links keep dphi = 0 (rhs.py:229-234)
      if (fl & kHoleX) dx_ = 0.0;
      if (fl & kHoleY) dy_ = 0.0;
      if (IS3D && (fl & kHoleZ)) dz_ = 0.0;

      // ---- epilogues
      const W64 o0 = wplanes64(e.out0, n);
      auto st5 = [&](const W64& w, double a, double bb, double xx, double yy, double zz)
                     __attribute__((always_inline)) {
        w.psi(q, make_double2(a, bb));
        w.ph(2, q, xx);
        w.ph(3, q, yy);
        if (IS3D) w.ph(4, q, zz);
      };
      if (MODE == M_RHS) {
        st5(o0, dpr, dpi, dx_, dy_, dz_);
      } else if (MODE == M_PRED || MODE == M_EULER) {
        // X_pred = X + dt f ; gamma = X + dt/2 f   (integrators.py:156-158, :82)
        st5(o0, c.xr + e.dt * dpr, c.xi + e.dt * dpi, c.xx + e.dt * dx_, c.xy + e.dt * dy_,
            c.xz + e.dt * dz_);
        if (MODE == M_PRED)
          st5(wplanes64(e.out1, n), c.xr + e.hdt * dpr, c.xi + e.hdt * dpi, c.xx + e.hdt * dx_,
              c.xy + e.hdt * dy_, c.xz + e.hdt * dz_);
      } else if (NEWTON) {
        // f, r = -g with g = x - dt/2 f - gamma (newton.py:98-100, :109-111).
        // NEWTON32 (mixed precision) stores x and r in FP32 instead of f and r.
        if (MODE == M_NEWTON) st5(o0, dpr, dpi, dx_, dy_, dz_);
        const double gr = (c.xr - e.hdt * dpr) - c.gr;
        const double gi = (c.xi - e.hdt * dpi) - c.gi;
        const double gx = (c.xx - e.hdt * dx_) - c.gx;
        const double gy = (c.xy - e.hdt * dy_) - c.gy;
        const double gz = IS3D ? (c.xz - e.hdt * dz_) - c.gz : 0.0;
        if (MODE == M_NEWTON) {
          st5(wplanes64(e.out1, n), -gr, -gi, -gx, -gy, -gz);
        } else {
          // the mixed method's FP32 layout (each node's links together)
          const W32 a32 = wplanes32(e.o32a, n, IS3D), b32 = wplanes32(e.o32b, n, IS3D);
          a32.psi(q, make_float2((float)c.xr, (float)c.xi));
          a32.links<IS3D>(q, (float)c.xx, (float)c.xy, (float)c.xz);
          b32.psi(q, make_float2((float)-gr, (float)-gi));
          b32.links<IS3D>(q, (float)-gx, (float)-gy, (float)-gz);
        }
        red[0] += ((c.xr * c.xr + c.xi * c.xi) + c.xx * c.xx) + (c.xy * c.xy + c.xz * c.xz);
        red[1] += ((gr * gr + gi * gi) + gx * gx) + (gy * gy + gz * gz);
        red[NS] = fmax(red[NS], fmax(hypot(gr, gi), fmax(fabs(gx), fmax(fabs(gy), fabs(gz)))));
      } else if (MODE == M_JVP) {
        // Ap = p - dt/2 * ((f(x+eps p) - f0) * (1/eps))   (tgcr.py:129-133)
        const double jr = (dpr - c.gr) * e.inv_eps, ji = (dpi - c.gi) * e.inv_eps;
        const double jx = (dx_ - c.gx) * e.inv_eps;
        const double jy = (dy_ - c.gy) * e.inv_eps;
        const double apr = c.xr - e.hdt * jr, api = c.xi - e.hdt * ji;
        const double apx = c.xx - e.hdt * jx;
        const double apy = c.xy - e.hdt * jy;
        double apz = 0.0;
        if (IS3D) {
          const double jz = (dz_ - c.gz) * e.inv_eps;
          apz = c.xz - e.hdt * jz;
        }
        if (e.cz) {  // the recurrence step: z = om (p + (r - A p) thinv) - (om - 1) zpp
          const V64 R = planes64(e.cr, n);
          const double2 rp = R.psi(q);
          const double rx = R.ph(2, q), ry = R.ph(3, q), rz = IS3D ? R.ph(4, q) : 0.0;
          double2 zq = make_double2(0.0, 0.0);
          double zx = 0.0, zy = 0.0, zz = 0.0;
          if (e.czpp) {
            const V64 Z = planes64(e.czpp, n);
            zq = Z.psi(q);
            zx = Z.ph(2, q);
            zy = Z.ph(3, q);
            if (IS3D) zz = Z.ph(4, q);
          }
          auto step = [&](double p, double r, double a, double zpp) {
            double v = e.com * (p + (r - a) * e.cthinv);
            if (e.czpp) v = v - (e.com - 1.0) * zpp;
            return v;
          };
          st5(wplanes64(e.cz, n), step(c.xr, rp.x, apr, zq.x), step(c.xi, rp.y, api, zq.y),
              step(c.xx, rx, apx, zx), step(c.xy, ry, apy, zy), IS3D ? step(c.xz, rz, apz, zz) : 0.0);
        } else {
          st5(o0, apr, api, apx, apy, apz);
        }
        if (e.ap0) {
          red[0] += ((apr * c.ar + api * c.ai) + apx * c.ax) + (apy * c.ay + apz * c.az);
        } else {
          red[0] += ((apr * apr + api * api) + apx * apx) + (apy * apy + apz * apz);
        }
      }
    }
    __syncthreads();
  }
}

template <bool IS3D, bool PERT, int MODE, bool GEN>
__global__ __launch_bounds__(BS) void k_stencil(Geo g, Bnd b, Mat mt, Src<double, PERT> s, Epi e,
                                                int tiles_x, int tiles_y, int kz, Red r) {
  const unsigned int L = xcd_remap(blockIdx.x, gridDim.x);
  const int tile_x = (int)(L % (unsigned)tiles_x);
  const unsigned int rest = L / (unsigned)tiles_x;
  const int tile_y = (int)(rest % (unsigned)tiles_y);
  const int zc = (int)(rest / (unsigned)tiles_y);
  using RN = RedN<MODE, PERT>;
  if constexpr (MODE == M_JVP) {
    if (e.tgcr) {  // tgcr.py:129 with ||p_k|| = ||r_k||, from the device's ||r||^2
      if (r.res[RS_STOP] != 0.0) return;
      s.eps = e.eps_num / sqrt(r.res[RS_RR]);
      e.inv_eps = 1.0 / s.eps;
    }
  }
  double red[RN::NV];
#pragma unroll
  for (int a = 0; a < RN::NV; ++a) red[a] = 0.0;
  stencil_tile<IS3D, PERT, MODE, GEN>(g, b, mt, s, e, tile_x, tile_y, zc, kz, red);
  if constexpr (RN::NS + RN::NM > 0) publish<RN::NS, RN::NM>(red, r.pout);
}

// Batched films: block L (after the XCD remap) is work item L of the batch;
// the item table names its film, whose descriptor gives the tile.  Partials
// are published at index L, so each film's partials are one contiguous range.
template <bool IS3D, int MODE, bool GEN>
__global__ __launch_bounds__(BS) void k_bstencil(BatchK bk, const double* x, double* out0,
                                                 double* out1, const double* gam, float* o32a,
                                                 float* o32b, double* pout) {
  const unsigned int L = xcd_remap(blockIdx.x, gridDim.x);
  const int f = bk.item_film[L];
  if (!bk.act[f]) return;
  const FilmDesc* fd = bk.films + f;
  const Geo g = fd->g;  // by value: held in registers for the whole item
  const Bnd b = fd->b;
  const Mat mt = fd->mt;
  const long long off = fd->off;
  const int li = (int)L - fd->item0;
  const int tiles_x = fd->tiles_x, tiles_y = fd->tiles_y;
  const int tile_x = li % tiles_x, rest = li / tiles_x;
  const int tile_y = rest % tiles_y, zc = rest / tiles_y;
  Src<double, false> s{x + off, nullptr, 0.0, g.n};
  Epi e{};
  e.out0 = out0 ? out0 + off : nullptr;
  e.out1 = out1 ? out1 + off : nullptr;
  e.gam = gam ? gam + off : nullptr;
  e.o32a = o32a ? o32a + off : nullptr;
  e.o32b = o32b ? o32b + off : nullptr;
  e.dt = bk.dtv[2 * f];
  e.hdt = bk.dtv[2 * f + 1];
  using RN = RedN<MODE, false>;
  double red[RN::NV];
#pragma unroll
  for (int a = 0; a < RN::NV; ++a) red[a] = 0.0;
  stencil_tile<IS3D, false, MODE, GEN>(g, b, mt, s, e, tile_x, tile_y, zc, fd->kz, red);
  if constexpr (RN::NS + RN::NM > 0) publish_at<RN::NS, RN::NM>(red, pout, L);
}

int stencil_blocks(const Geo& g, const StencilCfg& c) {
  const int tx = (g.nx + TX - 1) / TX, ty = (g.ny + TY - 1) / TY;
  const int tz = g.is3d ? (g.nz + c.kz - 1) / c.kz : 1;
  return tx * ty * tz;
}

template <bool IS3D, bool PERT, int MODE>
static int go_stencil(const Geo& g, const Bnd& b, const Mat& mt, const double* x, const double* p,
                      double eps, const Epi& e, const StencilCfg& cfg, const Red& r, hipStream_t st) {
  Src<double, PERT> s{x, p, eps, g.n};
  const int tx = (g.nx + TX - 1) / TX, ty = (g.ny + TY - 1) / TY;
  const int nb = stencil_blocks(g, cfg);
  if (b.bx || b.by || b.bz)
    launch((k_stencil<IS3D, PERT, MODE, true>), dim3(nb), dim3(BS), 0, st, g, b, mt, s, e, tx, ty, cfg.kz, r);
  else
    launch((k_stencil<IS3D, PERT, MODE, false>), dim3(nb), dim3(BS), 0, st, g, b, mt, s, e, tx, ty, cfg.kz, r);
  return nb;
}

template <bool IS3D>
static int go_mode(int mode, const Geo& g, const Bnd& b, const Mat& mt, const double* x,
                   const double* p, double eps, const Epi& e, const StencilCfg& cfg, const Red& r,
                   hipStream_t st) {
  switch (mode) {
    case M_RHS: return go_stencil<IS3D, false, M_RHS>(g, b, mt, x, p, eps, e, cfg, r, st);
    case M_PRED: return go_stencil<IS3D, false, M_PRED>(g, b, mt, x, p, eps, e, cfg, r, st);
    case M_NEWTON: return go_stencil<IS3D, false, M_NEWTON>(g, b, mt, x, p, eps, e, cfg, r, st);
    case M_JVP: return go_stencil<IS3D, true, M_JVP>(g, b, mt, x, p, eps, e, cfg, r, st);
    case M_EULER: return go_stencil<IS3D, false, M_EULER>(g, b, mt, x, p, eps, e, cfg, r, st);
    case M_NEWTON32: return go_stencil<IS3D, false, M_NEWTON32>(g, b, mt, x, p, eps, e, cfg, r, st);
  }
  return 0;
}

int launch_stencil(int mode, bool pert, const Geo& g, const Bnd& b, const Mat& mt, const double* x,
                   const double* p, double eps, const Epi& e, const StencilCfg& cfg, const Red& r,
                   hipStream_t st) {
  (void)pert;
  return g.is3d ? go_mode<true>(mode, g, b, mt, x, p, eps, e, cfg, r, st)
                : go_mode<false>(mode, g, b, mt, x, p, eps, e, cfg, r, st);
}

void launch_bstencil(const BatchK& bk, bool is3d, int mode, const double* x, double* out0,
                     double* out1, const double* gam, float* o32a, float* o32b, double* pout,
                     hipStream_t st) {
  const dim3 gr(bk.nitems), bl(BS);
#define TDGL_BST(D, M)                                                                              \
  do {                                                                                              \
    if (bk.gen) launch((k_bstencil<D, M, true>), gr, bl, 0, st, bk, x, out0, out1, gam, o32a, o32b, pout); \
    else launch((k_bstencil<D, M, false>), gr, bl, 0, st, bk, x, out0, out1, gam, o32a, o32b, pout);        \
  } while (0)
  if (mode == M_PRED) {
    if (is3d) TDGL_BST(true, M_PRED);
    else TDGL_BST(false, M_PRED);
  } else {
    if (is3d) TDGL_BST(true, M_NEWTON32);
    else TDGL_BST(false, M_NEWTON32);
  }
#undef TDGL_BST
}

// Per film (one block each): reduce the film's stencil partials, in item order.
template <int NS, int NM>
__global__ __launch_bounds__(kRedThreads) void k_bfin(BatchK bk, const double* pin, int slot0) {
  const int f = blockIdx.x;
  if (!bk.act[f]) return;
  const FilmDesc* fd = bk.films + f;
  double out[NS + NM];
  gather<NS, NM>(pin + (long long)fd->item0 * (NS + NM), fd->nitems, out);
  if (threadIdx.x == 0)
    for (int a = 0; a < NS + NM; ++a) bk.res[(long long)f * kResLen + slot0 + a] = out[a];
}

void launch_bfin(const BatchK& bk, int nsum, int nmax, const double* pin, int slot0, hipStream_t st) {
  (void)nsum;
  (void)nmax;
  launch((k_bfin<2, 1>), dim3(bk.nfilm), dim3(kRedThreads), 0, st, bk, pin, slot0);
}

// Host report: per film xx, rr, ginf, dxinf, r0, rk, state, basis size, iterations, breakdown.
__global__ void k_breport(BatchK bk, const int* fstat, double* out) {
  for (int f = threadIdx.x; f < bk.nfilm; f += blockDim.x) {
    const double* res = bk.res + (long long)f * kResLen;
    double* o = out + (long long)f * kReportLen;
    o[0] = res[RS_XX];
    o[1] = res[RS_RR];
    o[2] = res[RS_GINF];
    o[3] = res[RS_DXINF];
    o[4] = res[RS_R0];
    o[5] = res[RS_RK];
    o[6] = (double)fstat[4 * f];
    o[7] = (double)fstat[4 * f + 1];
    o[8] = (double)fstat[4 * f + 2];
    o[9] = res[RS_BRK];
  }
}

void launch_breport(const BatchK& bk, const int* fstat, double* out, hipStream_t st) {
  launch(k_breport, dim3(1), dim3(256), 0, st, bk, fstat, out);
}

// ---------------------------------------------------------------------------
// Node-wise vector kernels.  A node carries (psi re, psi im, phi_x, phi_y[, phi_z]).
// ---------------------------------------------------------------------------

struct V5 {
  double a, b, x, y, z;
};

template <bool IS3D>
__device__ __forceinline__ V5 ld5(const double* v, long long n, long long q) {
  const double2 p = reinterpret_cast<const double2*>(v)[q];
  V5 r{p.x, p.y, v[2 * n + q], v[3 * n + q], IS3D ? v[4 * n + q] : 0.0};
  return r;
}
template <bool IS3D>
__device__ __forceinline__ void st5(double* v, long long n, long long q, const V5& r) {
  reinterpret_cast<double2*>(v)[q] = make_double2(r.a, r.b);
  v[2 * n + q] = r.x;
  v[3 * n + q] = r.y;
  if (IS3D) v[4 * n + q] = r.z;
}
__device__ __forceinline__ double dot5(const V5& u, const V5& v) {
  return ((u.a * v.a + u.b * v.b) + u.x * v.x) + (u.y * v.y + u.z * v.z);
}
// u - c*v per component (numpy: `u - beta * v`)
__device__ __forceinline__ V5 axmy(const V5& u, double c, const V5& v) {
  return V5{u.a - c * v.a, u.b - c * v.b, u.x - c * v.x, u.y - c * v.y, u.z - c * v.z};
}
__device__ __forceinline__ V5 scal(const V5& u, double c) {
  return V5{u.a * c, u.b * c, u.x * c, u.y * c, u.z * c};
}

// Streaming grid: at most 1024 blocks (4 per CU), so consumers reduce at most
// 1024 partials in their prologue.
constexpr int kStreamCap = 1024;

int stream_blocks(const Geo& g) {
  long long nb = (g.n + kRedThreads - 1) / kRedThreads;
  return (int)(nb < kStreamCap ? nb : kStreamCap);
}

#define GRID_STRIDE(q) \
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < n; \
       q += (long long)gridDim.x * kRedThreads)

// Blocked modified Gram-Schmidt (tgcr.py:135-140, reorganised).  One pass
// subtracts the `no` old vectors a_0 .. a_{no-1} -- per element in the
// reference's order, Ap -= beta_0 a_0, then Ap -= beta_1 a_1, ... -- and then
// reduces <Ap, b_m> over the `nn` new vectors and their Gram entries
// <b_l, b_m> (l < m), or ||Ap||^2 when nn = 0.  The old betas come from the
// previous pass's sums d, G:
//     beta_m = d_m - sum_{l<m} beta_l G_lm,
// which equals <Ap - sum_{l<m} beta_l a_l, a_m>, the MGS coefficient, in
// exact arithmetic (the difference in rounding is of the order of the dot
// products' own).  Block 0 records them at res[bslot + l].  One pass moves
// (2 + no + nn) vectors for up to kMgsBlock coefficients, where one pass per
// coefficient moves 4.  NVIN: partials per block of the previous launch (1:
// the JVP's <Ap, a_0>; kMgsSums: a blocked pass).
//
// The vectors are streamed as flat arrays of `len` doubles (the real inner
// product of two states is the sum over all their doubles), as double2
// elements, kMgsGroup per thread and step with all their loads issued first.
// Every block gathers all partials of the previous pass (kMgsSums per block),
// so the pass runs few, large blocks -- kMgsThreads threads, at most one block
// per CU -- which keeps that gather at a few MB of cache traffic per pass.
// (Measured on the 128x128x16 film: 21.9 us per full pass of 10 vectors; 512
// threads x 2 elements beat 1024 x 1, 512 x 1 and 256 x 1 by 2-10 %.)
constexpr int kMgsThreads = 512;
constexpr int kMgsCap = 256;
constexpr int kMgsGroup = 2;  // elements per thread and step
int mgs_blocks(long long len) {
  const long long nb = (len / 2 + kMgsGroup * kMgsThreads - 1) / (kMgsGroup * kMgsThreads);
  return (int)(nb < 1 ? 1 : (nb < kMgsCap ? nb : kMgsCap));
}

template <int NVIN, bool LAST>
__global__ __launch_bounds__(kMgsThreads) void k_mgs_block(long long len, double* ap, MgsVecs v, int no,
                                                           int nn, int bslot, Red r) {
  if (r.res[RS_STOP] != 0.0) return;  // a queued iteration after the loop ended
  const long long n2 = len / 2;  // double2 elements; an odd last double is handled by thread 0 of block 0
  const long long e0 = (long long)blockIdx.x * kMgsThreads + threadIdx.x;
  const long long stride = (long long)gridDim.x * kMgsThreads;
  double2* X = reinterpret_cast<double2*>(ap);
  auto ld = [&](const double* p, long long e) __attribute__((always_inline)) {
    return reinterpret_cast<const double2*>(p)[e];
  };
  // A thread takes kMgsGroup elements (stride apart) per step, all loads
  // issued before any arithmetic; the first group's loads go out before the
  // partials are gathered.
  struct Grp {
    double2 x[kMgsGroup], a[kMgsGroup][kMgsBlock], b[kMgsGroup][kMgsBlock];
  };
  auto load = [&](long long e, Grp& g) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kMgsGroup; ++u) {
      const long long eu = e + u * stride;
      if (eu < n2) {
        g.x[u] = X[eu];
#pragma unroll
        for (int l = 0; l < kMgsBlock; ++l)
          if (l < no) g.a[u][l] = ld(v.a[l], eu);
#pragma unroll
        for (int m = 0; m < kMgsBlock; ++m)
          if (m < nn) g.b[u][m] = ld(v.b[m], eu);
      }
    }
  };
  Grp g0;
  load(e0, g0);
  double s[NVIN];
  gather<NVIN, 0, kMgsThreads>(r.pin, r.pcount, s);
  double beta[kMgsBlock];
  if constexpr (NVIN == 1) {
    beta[0] = s[0];
  } else {
#pragma unroll
    for (int m = 0; m < kMgsBlock; ++m) {
      double bm = s[m];
#pragma unroll
      for (int l = 0; l < m; ++l) bm -= beta[l] * s[mgs_g(l, m)];
      beta[m] = bm;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int l = 0; l < no; ++l) r.res[bslot + l] = beta[l];
  double acc[kMgsSums];
#pragma unroll
  for (int a = 0; a < kMgsSums; ++a) acc[a] = 0.0;
  // one element (two doubles): the update in the reference's order, then the sums
  auto pair = [&](long long e, double2 x, const double2 (&a)[kMgsBlock], const double2 (&b)[kMgsBlock])
                  __attribute__((always_inline)) {
#pragma unroll
    for (int l = 0; l < kMgsBlock; ++l) {
      if (l < no) {
        x.x = x.x - beta[l] * a[l].x;
        x.y = x.y - beta[l] * a[l].y;
      }
    }
    X[e] = x;
    if (LAST) {
      acc[0] += x.x * x.x;
      acc[0] += x.y * x.y;
      return;
    }
#pragma unroll
    for (int m = 0; m < kMgsBlock; ++m) {
      if (m < nn) {
        acc[m] += x.x * b[m].x;
        acc[m] += x.y * b[m].y;
#pragma unroll
        for (int l = 0; l < m; ++l) {
          acc[mgs_g(l, m)] += b[l].x * b[m].x;
          acc[mgs_g(l, m)] += b[l].y * b[m].y;
        }
      }
    }
  };
  auto proc = [&](long long e, const Grp& g) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < kMgsGroup; ++u)
      if (e + u * stride < n2) pair(e + u * stride, g.x[u], g.a[u], g.b[u]);
  };
  proc(e0, g0);
#pragma unroll 1
  for (long long e = e0 + kMgsGroup * stride; e < n2; e += kMgsGroup * stride) {
    Grp g;
    load(e, g);
    proc(e, g);
  }
  if ((len & 1) && blockIdx.x == 0 && threadIdx.x == 0) {  // the odd last double
    const long long e = len - 1;
    double x = ap[e], b[kMgsBlock];
#pragma unroll
    for (int l = 0; l < kMgsBlock; ++l)
      if (l < no) x = x - beta[l] * v.a[l][e];
#pragma unroll
    for (int m = 0; m < kMgsBlock; ++m) b[m] = m < nn ? v.b[m][e] : 0.0;
    ap[e] = x;
    if (LAST) {
      acc[0] += x * x;
    } else {
#pragma unroll
      for (int m = 0; m < kMgsBlock; ++m) {
        if (m < nn) {
          acc[m] += x * b[m];
#pragma unroll
          for (int l = 0; l < m; ++l) acc[mgs_g(l, m)] += b[l] * b[m];
        }
      }
    }
  }
  if (LAST) {
    double t[1] = {acc[0]};
    publish<1, 0, kMgsThreads>(t, r.pout);
  } else {
    publish<kMgsSums, 0, kMgsThreads>(acc, r.pout);
  }
}

// ||Ap|| from the previous partials; breakdown test (tgcr.py:140-142: ||Ap|| <
// 1e-14; right-preconditioned, ||Ap|| < 1e-14 ||r_0||, as for a unit right-hand
// side, since Ap = A M r is then of the residual's size);
// p_k = (r - sum_j beta_j p_j) / ||Ap|| ; Ap_k = Ap / ||Ap|| ; partials of alpha = <r, Ap_k>.
// The p updates of tgcr.py:137 are applied here, per node, in the same j order,
// which rounds exactly like the reference's vector-by-vector updates.
template <bool IS3D>
// zv (right preconditioning, M r): the direction's source in place of r; may alias pout.
__global__ __launch_bounds__(kRedThreads) void k_final_p(long long n, const double* rv,
                                                         const double* const* ptab, int nprev,
                                                         double* ap, double* pout, const double* zv,
                                                         Red r) {
  if (r.res[RS_STOP] != 0.0) return;  // a queued iteration after the loop ended
  double s1[1];
  gather<1, 0>(r.pin, r.pcount, s1);
  const double nrm = sqrt(s1[0]);
  const double inv = 1.0 / nrm;
  const bool brk = nrm < (zv ? 1e-14 * r.res[RS_R0] : 1e-14);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    r.res[RS_NRM] = nrm;
    r.res[RS_INV] = inv;
    r.res[RS_BRK] = brk ? 1.0 : 0.0;
  }
  if (brk) return;
  const double* beta = r.res + RS_BETA;
  double v[1] = {0.0};
  GRID_STRIDE(q) {
    const V5 rr = ld5<IS3D>(rv, n, q);
    V5 p = zv ? ld5<IS3D>(zv, n, q) : rr;
    for (int jj = 0; jj < nprev; ++jj) p = axmy(p, beta[jj], ld5<IS3D>(ptab[jj], n, q));
    const V5 a = scal(ld5<IS3D>(ap, n, q), inv);
    p = scal(p, inv);
    st5<IS3D>(ap, n, q, a);
    st5<IS3D>(pout, n, q, p);
    v[0] += dot5(rr, a);
  }
  publish<1, 0>(v, r.pout);
}

// alpha from the previous partials (block 0 records it); r -= alpha Ap;
// partials of ||r||^2   (tgcr.py:147-152)
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_update_r(long long n, double* rv, const double* ap,
                                                          int aslot, Red r) {
  if (r.res[RS_BRK] != 0.0 || r.res[RS_STOP] != 0.0) return;
  const long long q0 = (long long)blockIdx.x * kRedThreads + threadIdx.x;
  const long long stride = (long long)gridDim.x * kRedThreads;
  V5 r0{}, a0{};
  if (q0 < n) {
    r0 = ld5<IS3D>(rv, n, q0);
    a0 = ld5<IS3D>(ap, n, q0);
  }
  double a1[1];
  gather<1, 0>(r.pin, r.pcount, a1);
  const double alpha = a1[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) r.res[aslot] = alpha;
  double v[1] = {0.0};
  if (q0 < n) {
    const V5 a = axmy(r0, alpha, a0);
    st5<IS3D>(rv, n, q0, a);
    v[0] += dot5(a, a);
  }
  for (long long q = q0 + stride; q < n; q += stride) {
    const V5 a = axmy(ld5<IS3D>(rv, n, q), alpha, ld5<IS3D>(ap, n, q));
    st5<IS3D>(rv, n, q, a);
    v[0] += dot5(a, a);
  }
  publish<1, 0>(v, r.pout);
}

// dx = sum_i alpha_i p_i (the x += alpha p sequence of tgcr.py:148, per node);
// x += dx (newton.py:108); partials of ||dx||_inf (newton.py:113)
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_accum(long long n, double* x,
                                                       const double* const* ptab, int nvec, Red r) {
  const double* alpha = r.res + RS_ALPHA;
  double v[1] = {0.0};
  GRID_STRIDE(q) {
    V5 d{0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ii = 0; ii < nvec; ++ii) {
      const V5 p = ld5<IS3D>(ptab[ii], n, q);
      const double a = alpha[ii];
      d = V5{d.a + a * p.a, d.b + a * p.b, d.x + a * p.x, d.y + a * p.y, d.z + a * p.z};
    }
    const V5 xv = ld5<IS3D>(x, n, q);
    st5<IS3D>(x, n, q, V5{xv.a + d.a, xv.b + d.b, xv.x + d.x, xv.y + d.y, xv.z + d.z});
    v[0] = fmax(v[0], fmax(hypot(d.a, d.b), fmax(fabs(d.x), fmax(fabs(d.y), fabs(d.z)))));
  }
  publish<0, 1>(v, r.pout);
}

// out = x + r (x - xp) over a whole state vector (the extrapolating predictor)
__global__ __launch_bounds__(kRedThreads) void k_extrap64(const double* x, const double* xp, double r,
                                                          double* out, long long len) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride)
    out[i] = x[i] + r * (x[i] - xp[i]);
}
void launch_extrap64(const double* x, const double* xp, double ratio, double* out, long long len,
                     hipStream_t st) {
  const long long blocks = std::min<long long>((len + kRedThreads - 1) / kRedThreads, 4096);
  launch(k_extrap64, dim3((unsigned)blocks), dim3(kRedThreads), 0, st, x, xp, ratio, out, len);
}

template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_dot(long long n, const double* a, const double* b,
                                                     Red r) {
  double v[1] = {0.0};
  GRID_STRIDE(q) { v[0] += dot5(ld5<IS3D>(a, n, q), ld5<IS3D>(b, n, q)); }
  publish<1, 0>(v, r.pout);
}

template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_norm_inf(long long n, const double* a, Red r) {
  double v[1] = {0.0};
  GRID_STRIDE(q) {
    const V5 d = ld5<IS3D>(a, n, q);
    v[0] = fmax(v[0], fmax(hypot(d.a, d.b), fmax(fabs(d.x), fmax(fabs(d.y), fabs(d.z)))));
  }
  publish<0, 1>(v, r.pout);
}

template <int NS, int NM>
__global__ __launch_bounds__(kRedThreads) void k_finalize(Red r, int slot0) {
  double out[NS + NM];
  gather<NS, NM>(r.pin, r.pcount, out);
  if (threadIdx.x == 0)
    for (int a = 0; a < NS + NM; ++a) r.res[slot0 + a] = out[a];
}

#define LAUNCH_V(kern, ...)                                                                 \
  do {                                                                                      \
    if (g.is3d)                                                                             \
      launch(kern<true>, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, __VA_ARGS__); \
    else                                                                                    \
      launch(kern<false>, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, __VA_ARGS__); \
  } while (0)

// ||r||^2 into res[RS_RR]; res[RS_STOP] = the exit test of the TGCR loop
// after iteration k (tgcr.py:124, 141-142: breakdown, ||r|| / ||r_0|| <= tol,
// the iteration cap) or a non-finite residual, evaluated as the host does.
__global__ __launch_bounds__(kRedThreads) void k_tgcr_check(Red r, double r0, double tol, long long k,
                                                            long long maxk) {
  if (r.res[RS_STOP] != 0.0) return;
  double s1[1];
  gather<1, 0>(r.pin, r.pcount, s1);
  if (threadIdx.x == 0) {
    const bool brk = r.res[RS_BRK] != 0.0;
    r.res[RS_RR] = s1[0];
    const double rk = sqrt(s1[0]);
    const bool stop = brk || !isfinite(rk) || !(rk / r0 > tol) || k >= maxk;
    r.res[RS_STOP] = stop ? 1.0 : 0.0;
  }
}
void launch_tgcr_check(double r0, double tol, long long k, long long maxk, const Red& r, hipStream_t st) {
  launch(k_tgcr_check, dim3(1), dim3(kRedThreads), 0, st, r, r0, tol, k, maxk);
}

int launch_mgs_block(const Geo& g, double* ap, const MgsVecs& v, int no, int nn, int beta_slot,
                     bool first, const Red& r, hipStream_t st) {
  const long long len = (long long)(2 + g.nphi) * g.n;
  const dim3 gr(mgs_blocks(len)), bl(kMgsThreads);
  if (first) {
    if (nn) launch((k_mgs_block<1, false>), gr, bl, 0, st, len, ap, v, no, nn, beta_slot, r);
    else launch((k_mgs_block<1, true>), gr, bl, 0, st, len, ap, v, no, nn, beta_slot, r);
  } else {
    if (nn) launch((k_mgs_block<kMgsSums, false>), gr, bl, 0, st, len, ap, v, no, nn, beta_slot, r);
    else launch((k_mgs_block<kMgsSums, true>), gr, bl, 0, st, len, ap, v, no, nn, beta_slot, r);
  }
  return mgs_blocks(len);
}
int launch_final_p(const Geo& g, const double* rv, const double* const* ptab, int nprev, double* ap,
                   double* pout, const Red& r, hipStream_t st, const double* zv) {
  LAUNCH_V(k_final_p, g.n, rv, ptab, nprev, ap, pout, zv, r);
  return stream_blocks(g);
}
int launch_update_r(const Geo& g, double* rv, const double* ap, int aslot, const Red& r,
                    hipStream_t st) {
  LAUNCH_V(k_update_r, g.n, rv, ap, aslot, r);
  return stream_blocks(g);
}
int launch_accum(const Geo& g, double* x, const double* const* ptab, int nvec, const Red& r,
                 hipStream_t st) {
  LAUNCH_V(k_accum, g.n, x, ptab, nvec, r);
  return stream_blocks(g);
}
int launch_dot(const Geo& g, const double* a, const double* b, const Red& r, hipStream_t st) {
  LAUNCH_V(k_dot, g.n, a, b, r);
  return stream_blocks(g);
}
int launch_norm_inf(const Geo& g, const double* a, const Red& r, hipStream_t st) {
  LAUNCH_V(k_norm_inf, g.n, a, r);
  return stream_blocks(g);
}
void launch_finalize(int nsum, int nmax, int slot0, const Red& r, hipStream_t st) {
  if (nsum == 1 && nmax == 0)
    launch((k_finalize<1, 0>), dim3(1), dim3(kRedThreads), 0, st, r, slot0);
  else if (nsum == 0 && nmax == 1)
    launch((k_finalize<0, 1>), dim3(1), dim3(kRedThreads), 0, st, r, slot0);
  else if (nsum == 3 && nmax == 0)
    launch((k_finalize<3, 0>), dim3(1), dim3(kRedThreads), 0, st, r, slot0);
  else
    launch((k_finalize<2, 1>), dim3(1), dim3(kRedThreads), 0, st, r, slot0);
}

// ---------------------------------------------------------------------------
// Layout conversion.  Reference interior order r = ((i-1)*ny + (j-1))*nz + (k-1)
// (meshgrid 'ij', k fastest: indices.py:271-278).
// ---------------------------------------------------------------------------

__device__ __forceinline__ long long ref_of(const Geo& g, long long q) {
  const long long i1 = q % g.nx;
  const long long t = q / g.nx;
  const long long j1 = t % g.ny;
  const long long k1 = t / g.ny;
  return (i1 * g.ny + j1) * g.nz + k1;
}

__global__ void k_permute_in(Geo g, const double2* ref, double* out, int* bad) {
  const long long n = g.n;
  GRID_STRIDE(q) {
    const long long r = ref_of(g, q);
    reinterpret_cast<double2*>(out)[q] = ref[r];
    int nonreal = 0;
    for (int a = 0; a < g.nphi; ++a) {
      const double2 v = ref[(a + 1) * n + r];
      out[(2 + a) * n + q] = v.x;
      nonreal |= (v.y != 0.0);
    }
    if (nonreal) atomicOr(bad, 1);
  }
}

__global__ void k_permute_out(Geo g, const double* in, double2* ref) {
  const long long n = g.n;
  GRID_STRIDE(q) {
    const long long r = ref_of(g, q);
    ref[r] = reinterpret_cast<const double2*>(in)[q];
    for (int a = 0; a < g.nphi; ++a) ref[(a + 1) * n + r] = make_double2(in[(2 + a) * n + q], 0.0);
  }
}

void launch_permute_in(const Geo& g, const void* ref, double* st_out, int* bad, hipStream_t st) {
  launch(k_permute_in, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, g,
                     static_cast<const double2*>(ref), st_out, bad);
}
void launch_permute_out(const Geo& g, const double* st_in, void* ref, hipStream_t st) {
  launch(k_permute_out, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, g, st_in,
                     static_cast<double2*>(ref));
}

// ---------------------------------------------------------------------------
// Observables (bfield.py:58-68 with BCs; current_density.py:61-68 without)
// ---------------------------------------------------------------------------

template <bool IS3D>
__global__ void k_obs(Geo g, Bnd b, const uint8_t* flags, const double* x, double* bout,
                      double* jout) {
  const long long n = g.n;
  Src<double, false> s{x, nullptr, 0.0, n};
  GRID_STRIDE(q) {
    const int i = 1 + (int)(q % g.nx);
    const long long t = q / g.nx;
    const int j = 1 + (int)(t % g.ny);
    const int k = IS3D ? 1 + (int)(t / g.ny) : 0;
    const long long r = ref_of(g, q);
    if (bout) {
      const Cell<double> c = eff_cell<double, IS3D, false>(g, b, flags, s, i, j, k);
      const Cell<double> ci = eff_cell<double, IS3D, false>(g, b, flags, s, i + 1, j, k);
      const Cell<double> cj = eff_cell<double, IS3D, false>(g, b, flags, s, i, j + 1, k);
      double bx = 0.0, by = 0.0;
      if (IS3D) {
        const Cell<double> ck = eff_cell<double, IS3D, false>(g, b, flags, s, i, j, k + 1);
        bx = (1.0 / (b.hy * b.hz)) * (((c.y - ck.y) - c.z) + cj.z);
        by = (1.0 / (b.hz * b.hx)) * (((c.z - ci.z) - c.x) + ck.x);
      }
      const double bz = (1.0 / (b.hx * b.hy)) * (((c.x - cj.x) - c.y) + ci.y);
      bout[r] = bx;
      bout[n + r] = by;
      bout[2 * n + r] = bz;
    }
    if (jout) {
      // expanded state without BCs: ghosts are 0, interior values raw
      double pr, pi;
      s.psi(q, pr, pi);
      const double ph[3] = {s.phx(q), s.phy(q), IS3D ? s.phz(q) : 0.0};
      const int nb[3][3] = {{i + 1, j, k}, {i, j + 1, k}, {i, j, k + 1}};
      for (int a = 0; a < 3; ++a) {
        double js = 0.0;
        if (a < 2 || IS3D) {
          const int ii = nb[a][0], jj = nb[a][1], kk = nb[a][2];
          double qr = 0.0, qi = 0.0;
          if (ii <= g.Nx - 1 && jj <= g.Ny - 1 && (!IS3D || kk <= g.Nz - 1))
            s.psi(qix(g, ii, jj, kk), qr, qi);
          double sn, cs;
          sincos(ph[a], &sn, &cs);
          // conj(psi) * exp(-1j*phi) * psi1
          const double tr = pr * cs - (-pi) * (-sn);
          const double ti = pr * (-sn) + (-pi) * cs;
          js = tr * qi + ti * qr;
        }
        jout[a * n + r] = js;
      }
    }
  }
}

void launch_observables(const Geo& g, const Bnd& b, const Mat& mt, const double* x, double* bout,
                        double* jout, hipStream_t st) {
  if (g.is3d)
    launch(k_obs<true>, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, g, b, mt.flags,
                       x, bout, jout);
  else
    launch(k_obs<false>, dim3(stream_blocks(g)), dim3(kRedThreads), 0, st, g, b,
                       mt.flags, x, bout, jout);
}

}  // namespace tdgl
