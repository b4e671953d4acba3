// Mixed-precision kernels of the trapezoidal step.
//
// FP32: the Jacobian-vector product (analytic, exact derivative of eval_f)
// and the Krylov vectors p, Ap, r.  FP64: every dot product and norm is
// accumulated in double, the Gram-Schmidt / alpha coefficients are double, and
// the state update x += sum alpha_i p_i accumulates in double into the FP64
// Newton iterate.  The Newton residual itself is the FP64 stencil (tdgl_f64.hip).
//
// Orthogonalisation is classical Gram-Schmidt with one conditional
// re-orthogonalisation (the Daniel-Gragg-Kaufman-Stewart test ||Ap'|| <
// ||Ap|| / sqrt(2)), two passes over the basis instead of one per vector.

#include <hip/hip_runtime.h>
#include <math.h>

#include "tdgl_device.hpp"
#include "tdgl_launch.hpp"

namespace tdgl {

namespace {
constexpr int TX = 64, TY = 4;
constexpr int CX = TX + 2, CELLS = CX * (TY + 2);
constexpr int BS = TX * TY;
}  // namespace

// ---------------------------------------------------------------------------
// k_jvp32: Ap = p - dt/2 * J(x) p, all in FP32.
// J p is the exact derivative of eval_f (rhs.py:147-239) along p.  The
// boundary rules are affine in the state, so p's ghost values use the same
// rules with a zero applied field; hole-boundary rows are zero.
// ---------------------------------------------------------------------------
template <bool IS3D>
__global__ __launch_bounds__(BS) void k_jvp32(Geo g, Bnd b, Bnd b0, Mat mt, const float* x32,
                                              const float* p32, float* ap, float hdt, int tiles_x,
                                              int tiles_y, int kz) {
  constexpr int NS = IS3D ? 3 : 1;
  constexpr int NZ = IS3D ? 3 : 1;
  constexpr int CZ = IS3D ? CELLS : 1;
  __shared__ float X_r[NS][CELLS], X_i[NS][CELLS], X_x[NS][CELLS], X_y[NS][CELLS], X_z[NZ][CZ];
  __shared__ float P_r[NS][CELLS], P_i[NS][CELLS], P_x[NS][CELLS], P_y[NS][CELLS], P_z[NZ][CZ];

  const unsigned int L = xcd_remap(blockIdx.x, gridDim.x);
  const int tile_x = (int)(L % (unsigned)tiles_x);
  const unsigned int rest = L / (unsigned)tiles_x;
  const int tile_y = (int)(rest % (unsigned)tiles_y);
  const int zc = (int)(rest / (unsigned)tiles_y);
  const int i0 = 1 + tile_x * TX, j0 = 1 + tile_y * TY;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int i = i0 + tx, j = j0 + ty;
  const bool act = (i <= g.Nx - 1) && (j <= g.Ny - 1);
  int kb = 0, ke = 1;
  if (IS3D) {
    kb = 1 + zc * kz;
    ke = min(kb + kz, g.Nz);
  }
  const Src<float, false> sx{x32, nullptr, 0.f, g.n};
  const Src<float, false> sp{p32, nullptr, 0.f, g.n};

  auto load_plane = [&](int kp, int slot) {
    for (int c = threadIdx.x; c < CELLS; c += BS) {
      const int ci = i0 - 1 + c % CX, cj = j0 - 1 + c / CX;
      const Cell<float> v = eff_cell<float, IS3D, false>(g, b, mt.flags, sx, ci, cj, kp);
      const Cell<float> w = eff_cell<float, IS3D, false>(g, b0, mt.flags, sp, ci, cj, kp);
      X_r[slot][c] = v.pr;
      X_i[slot][c] = v.pi;
      X_x[slot][c] = v.x;
      X_y[slot][c] = v.y;
      P_r[slot][c] = w.pr;
      P_i[slot][c] = w.pi;
      P_x[slot][c] = w.x;
      P_y[slot][c] = w.y;
      if (IS3D) {
        X_z[slot][c] = v.z;
        P_z[slot][c] = w.z;
      }
    }
  };
  if (IS3D) {
    load_plane(kb - 1, (kb - 1) % 3);
    load_plane(kb, kb % 3);
  } else {
    load_plane(0, 0);
  }

  const float ihx2 = (float)mt.ihx2, ihy2 = (float)mt.ihy2, ihz2 = (float)mt.ihz2;
  const float dgpsi = (float)mt.dgpsi, inv_tau = (float)mt.inv_tau;

  for (int k = kb; k < ke; ++k) {
    if (IS3D) load_plane(k + 1, (k + 1) % 3);
    __syncthreads();
    if (act) {
      const int s0 = IS3D ? k % 3 : 0, sm = IS3D ? (k + 2) % 3 : 0, sp1 = IS3D ? (k + 1) % 3 : 0;
      const int c0 = (ty + 1) * CX + (tx + 1);
      const int cxm = c0 - 1, cxp = c0 + 1, cym = c0 - CX, cyp = c0 + CX;
      const long long q = qix(g, i, j, k);
      const float pr = X_r[s0][c0], pi = X_i[s0][c0];
      const float dr = P_r[s0][c0], di = P_i[s0][c0];
      const float ex = X_x[s0][c0], ey = X_y[s0][c0], ez = IS3D ? X_z[s0][c0] : 0.f;
      const float dex = P_x[s0][c0], dey = P_y[s0][c0], dez = IS3D ? P_z[s0][c0] : 0.f;

      float sxp, cxp_, sxm, cxm_, syp, cyp_, sym, cym_, szp = 0.f, czp = 1.f, szm = 0.f, czm = 1.f;
      sincosf(ex, &sxp, &cxp_);
      sincosf(X_x[s0][cxm], &sxm, &cxm_);
      sincosf(ey, &syp, &cyp_);
      sincosf(X_y[s0][cym], &sym, &cym_);
      if (IS3D) {
        sincosf(ez, &szp, &czp);
        sincosf(X_z[sm][c0], &szm, &czm);
      }

      // d[e^{-i t} psi(-a)] = e^{-i t}(d psi(-a) - i dt psi(-a)),  d[e^{i t} psi(+a)] = e^{i t}(d psi(+a) + i dt psi(+a))
      float ar = dgpsi * dr, ai = dgpsi * di;
      auto minus = [&](float c, float s, float dt_, float qr, float qi, float dqr, float dqi, float h) {
        const float wr = dqr + dt_ * qi, wi = dqi - dt_ * qr;
        ar += (c * wr + s * wi) * h;
        ai += (c * wi - s * wr) * h;
      };
      auto plus = [&](float c, float s, float dt_, float qr, float qi, float dqr, float dqi, float h) {
        const float wr = dqr - dt_ * qi, wi = dqi + dt_ * qr;
        ar += (c * wr - s * wi) * h;
        ai += (c * wi + s * wr) * h;
      };
      minus(cxm_, sxm, P_x[s0][cxm], X_r[s0][cxm], X_i[s0][cxm], P_r[s0][cxm], P_i[s0][cxm], ihx2);
      plus(cxp_, sxp, dex, X_r[s0][cxp], X_i[s0][cxp], P_r[s0][cxp], P_i[s0][cxp], ihx2);
      minus(cym_, sym, P_y[s0][cym], X_r[s0][cym], X_i[s0][cym], P_r[s0][cym], P_i[s0][cym], ihy2);
      plus(cyp_, syp, dey, X_r[s0][cyp], X_i[s0][cyp], P_r[s0][cyp], P_i[s0][cyp], ihy2);
      if (IS3D) {
        minus(czm, szm, P_z[sm][c0], X_r[sm][c0], X_i[sm][c0], P_r[sm][c0], P_i[sm][c0], ihz2);
        plus(czp, szp, dez, X_r[sp1][c0], X_i[sp1][c0], P_r[sp1][c0], P_i[sp1][c0], ihz2);
      }
      // d[(1-|psi|^2) psi] = (1 - 2|psi|^2) dpsi - psi^2 conj(dpsi)
      const float m2 = pr * pr + pi * pi, a2 = pr * pr - pi * pi, b2 = 2.f * pr * pi;
      float fr = (1.f - 2.f * m2) * dr - (a2 * dr + b2 * di);
      float fi = (1.f - 2.f * m2) * di - (b2 * dr - a2 * di);
      const uint8_t fl = mt.flags ? mt.flags[mfull(g, i, j, k)] : (uint8_t)0;
      if (mt.has_mat) {
        float sc;
        if (mt.sc_bits) sc = (fl & kInsulator) ? 0.f : 1.f;
        else sc = (float)mt.sc[q];
        fr = sc * fr - (1.f - sc) * dr * inv_tau;
        fi = sc * fi - (1.f - sc) * di * inv_tau;
      }
      const float jr = ar + fr, ji = ai + fi;

      float cx, cy, cz;
      if (mt.k2) {
        const float k2 = (float)mt.k2[q];
        cx = k2 / (float)mt.hx2;
        cy = k2 / (float)mt.hy2;
        cz = IS3D ? k2 / (float)mt.hz2 : 0.f;
      } else {
        cx = (float)mt.cx;
        cy = (float)mt.cy;
        cz = (float)mt.cz;
      }
      // d Im[e^{-i t} conj(psi) psi1] = Im[e^{-i t} U] - dt Re[e^{-i t} T],
      // T = conj(psi) psi1, U = conj(dpsi) psi1 + conj(psi) dpsi1
      auto dsup = [&](float c, float s, float dt_, float qr, float qi, float dqr, float dqi) {
        const float Tr = pr * qr + pi * qi, Ti = pr * qi - pi * qr;
        const float Ur = (dr * qr + di * qi) + (pr * dqr + pi * dqi);
        const float Ui = (dr * qi - di * qr) + (pr * dqi - pi * dqr);
        return (c * Ui - s * Ur) - dt_ * (c * Tr + s * Ti);
      };
      // (LPHI_y + LPHI_z) etc. on the perturbation: 2 D_aa along the link, D_bb across
      float jx = 2.f * cx * (P_x[s0][cxm] - 2.f * dex + P_x[s0][cxp]) +
                 cy * (P_x[s0][cym] - 2.f * dex + P_x[s0][cyp]);
      float jy = 2.f * cy * (P_y[s0][cym] - 2.f * dey + P_y[s0][cyp]) +
                 cx * (P_y[s0][cxm] - 2.f * dey + P_y[s0][cxp]);
      jx += cy * (((-P_y[s0][cxp] + dey) + P_y[s0][cxp - CX]) - P_y[s0][cym]);
      jy += cx * (((-P_x[s0][cyp] + dex) + P_x[s0][cyp - 1]) - P_x[s0][cxm]);
      float jz = 0.f;
      if (IS3D) {
        jx += cz * (P_x[sm][c0] - 2.f * dex + P_x[sp1][c0]);
        jy += cz * (P_y[sm][c0] - 2.f * dey + P_y[sp1][c0]);
        jx += cz * (((-P_z[s0][cxp] + dez) + P_z[sm][cxp]) - P_z[sm][c0]);
        jy += cz * (((-P_z[s0][cyp] + dez) + P_z[sm][cyp]) - P_z[sm][c0]);
        jz = 2.f * cz * (P_z[sm][c0] - 2.f * dez + P_z[sp1][c0]) +
             cx * (P_z[s0][cxm] - 2.f * dez + P_z[s0][cxp]) +
             cy * (P_z[s0][cym] - 2.f * dez + P_z[s0][cyp]);
        jz += cx * (((-P_x[sp1][c0] + dex) + P_x[sp1][cxm]) - P_x[s0][cxm]);
        jz += cy * (((-P_y[sp1][c0] + dey) + P_y[sp1][cym]) - P_y[s0][cym]);
        jz += dsup(czp, szp, dez, X_r[sp1][c0], X_i[sp1][c0], P_r[sp1][c0], P_i[sp1][c0]);
      }
      jx += dsup(cxp_, sxp, dex, X_r[s0][cxp], X_i[s0][cxp], P_r[s0][cxp], P_i[s0][cxp]);
      jy += dsup(cyp_, syp, dey, X_r[s0][cyp], X_i[s0][cyp], P_r[s0][cyp], P_i[s0][cyp]);
      if (fl & kHoleX) jx = 0.f;
      if (fl & kHoleY) jy = 0.f;
      if (IS3D && (fl & kHoleZ)) jz = 0.f;

      const long long n = g.n;
      const float2 p0 = reinterpret_cast<const float2*>(p32)[q];
      reinterpret_cast<float2*>(ap)[q] = make_float2(p0.x - hdt * jr, p0.y - hdt * ji);
      ap[2 * n + q] = p32[2 * n + q] - hdt * jx;
      ap[3 * n + q] = p32[3 * n + q] - hdt * jy;
      if (IS3D) ap[4 * n + q] = p32[4 * n + q] - hdt * jz;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// FP32 vectors, FP64 accumulation.
// ---------------------------------------------------------------------------

struct D5 {
  double a, b, x, y, z;
};

template <bool IS3D>
__device__ __forceinline__ D5 ldf(const float* v, long long n, long long q) {
  const float2 p = reinterpret_cast<const float2*>(v)[q];
  return D5{(double)p.x, (double)p.y, (double)v[2 * n + q], (double)v[3 * n + q],
            IS3D ? (double)v[4 * n + q] : 0.0};
}
template <bool IS3D>
__device__ __forceinline__ void stf(float* v, long long n, long long q, const D5& r) {
  reinterpret_cast<float2*>(v)[q] = make_float2((float)r.a, (float)r.b);
  v[2 * n + q] = (float)r.x;
  v[3 * n + q] = (float)r.y;
  if (IS3D) v[4 * n + q] = (float)r.z;
}
__device__ __forceinline__ double dotd(const D5& u, const D5& v) {
  return ((u.a * v.a + u.b * v.b) + u.x * v.x) + (u.y * v.y + u.z * v.z);
}

// Re-orthogonalise when the first pass removed more than half of ||Ap||^2.
__device__ __forceinline__ bool need_reorth(const double* res) {
  return res[RS_NRM1] <= 0.5 * res[RS_SELF];
}

constexpr int kMdotChunk = 8;

// partials[block][c] = <Ap, V_c> for c < ncols, and [ncols] = ||Ap||^2.
// skip_unless_reorth: the second Gram-Schmidt pass, executed only when needed.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_mdot32(long long n, const float* ap,
                                                        const float* const* V, int ncols,
                                                        int with_self, int second, Red r) {
  if (second && !need_reorth(r.res)) return;
  const int NC = ncols + (with_self ? 1 : 0);
  for (int c0 = 0; c0 < NC; c0 += kMdotChunk) {
    double acc[kMdotChunk];
#pragma unroll
    for (int cc = 0; cc < kMdotChunk; ++cc) acc[cc] = 0.0;
    for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < n;
         q += (long long)gridDim.x * kRedThreads) {
      const D5 a = ldf<IS3D>(ap, n, q);
#pragma unroll
      for (int cc = 0; cc < kMdotChunk; ++cc) {
        const int c = c0 + cc;
        if (c < ncols) acc[cc] += dotd(a, ldf<IS3D>(V[c], n, q));
        else if (c == ncols && with_self) acc[cc] += dotd(a, a);
      }
    }
    block_reduce<kMdotChunk, 0>(acc);
    if (threadIdx.x == 0)
      for (int cc = 0; cc < kMdotChunk && c0 + cc < NC; ++cc)
        r.pout[(long long)blockIdx.x * NC + c0 + cc] = acc[cc];
  }
}

// Column sums of partials[block][c] in block order, one block per column:
// columns c < n_coef go to res[base + c], column n_coef (if present) to res[self_slot].
__global__ __launch_bounds__(kRedThreads) void k_colsum(const double* pin, int count, int ncols,
                                                        int n_coef, double* res, int base,
                                                        int self_slot, int second) {
  const int c = blockIdx.x;
  const int dst = (c < n_coef) ? base + c : self_slot;
  if (second && !need_reorth(res)) {
    if (threadIdx.x == 0 && c < n_coef) res[dst] = 0.0;
    return;
  }
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < count; b += kRedThreads) v[0] += pin[(long long)b * ncols + c];
  block_reduce<1, 0>(v);
  if (threadIdx.x == 0) res[dst] = v[0];
}

// Ap -= sum_c coef_c V_c  (FP64 accumulation); partials of ||Ap||^2.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_cgs_upd32(long long n, float* ap,
                                                           const float* const* V, int ncols,
                                                           int coef_base, int second, Red r) {
  if (second && !need_reorth(r.res)) return;
  const double* coef = r.res + coef_base;
  double v[1] = {0.0};
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < n;
       q += (long long)gridDim.x * kRedThreads) {
    D5 s{0, 0, 0, 0, 0};
    for (int c = 0; c < ncols; ++c) {
      const D5 w = ldf<IS3D>(V[c], n, q);
      const double cc = coef[c];
      s = D5{s.a + cc * w.a, s.b + cc * w.b, s.x + cc * w.x, s.y + cc * w.y, s.z + cc * w.z};
    }
    const D5 a = ldf<IS3D>(ap, n, q);
    const D5 o{a.a - s.a, a.b - s.b, a.x - s.x, a.y - s.y, a.z - s.z};
    stf<IS3D>(ap, n, q, o);
    const D5 of = ldf<IS3D>(ap, n, q);  // norm of the stored (rounded) vector
    v[0] += dotd(of, of);
  }
  publish<1, 0>(v, r.pout);
}

// p_k = (r - sum_j (c_j + d_j) p_j) / ||Ap||, Ap_k = Ap / ||Ap||, partials of <r, Ap_k>.
// ||Ap||^2 is res[nrm_slot] when there is no previous vector, otherwise the
// norm after the last Gram-Schmidt pass that ran.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_final_p32(long long n, const float* rv,
                                                           const float* const* P, int nprev,
                                                           float* ap, float* pout, int nrm_slot,
                                                           Red r) {
  double nrm2;
  if (nprev == 0) nrm2 = r.res[nrm_slot];
  else nrm2 = need_reorth(r.res) ? r.res[RS_NRM2] : r.res[RS_NRM1];
  const double nrm = sqrt(nrm2);
  const double inv = 1.0 / nrm;
  const bool brk = nrm < 1e-14;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    r.res[RS_NRM] = nrm;
    r.res[RS_INV] = inv;
    r.res[RS_BRK] = brk ? 1.0 : 0.0;
  }
  if (brk) return;
  const double* c1 = r.res + RS_BETA;
  const double* c2 = r.res + RS_DBETA;
  const bool two = nprev > 0 && need_reorth(r.res);
  double v[1] = {0.0};
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < n;
       q += (long long)gridDim.x * kRedThreads) {
    const D5 rr = ldf<IS3D>(rv, n, q);
    D5 s{0, 0, 0, 0, 0};
    for (int jj = 0; jj < nprev; ++jj) {
      const D5 w = ldf<IS3D>(P[jj], n, q);
      const double cc = two ? c1[jj] + c2[jj] : c1[jj];
      s = D5{s.a + cc * w.a, s.b + cc * w.b, s.x + cc * w.x, s.y + cc * w.y, s.z + cc * w.z};
    }
    const D5 p{(rr.a - s.a) * inv, (rr.b - s.b) * inv, (rr.x - s.x) * inv, (rr.y - s.y) * inv,
               (rr.z - s.z) * inv};
    const D5 a0 = ldf<IS3D>(ap, n, q);
    const D5 a{a0.a * inv, a0.b * inv, a0.x * inv, a0.y * inv, a0.z * inv};
    stf<IS3D>(pout, n, q, p);
    stf<IS3D>(ap, n, q, a);
    v[0] += dotd(rr, ldf<IS3D>(ap, n, q));
  }
  publish<1, 0>(v, r.pout);
}

// r -= alpha Ap (alpha from the previous partials); partials of ||r||^2.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_update_r32(long long n, float* rv, const float* ap,
                                                            int aslot, Red r) {
  if (r.res[RS_BRK] != 0.0) return;
  double a1[1];
  gather<1, 0>(r.pin, r.pcount, a1);
  const double alpha = a1[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) r.res[aslot] = alpha;
  double v[1] = {0.0};
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < n;
       q += (long long)gridDim.x * kRedThreads) {
    const D5 x = ldf<IS3D>(rv, n, q), a = ldf<IS3D>(ap, n, q);
    stf<IS3D>(rv, n, q, D5{x.a - alpha * a.a, x.b - alpha * a.b, x.x - alpha * a.x,
                           x.y - alpha * a.y, x.z - alpha * a.z});
    const D5 o = ldf<IS3D>(rv, n, q);
    v[0] += dotd(o, o);
  }
  publish<1, 0>(v, r.pout);
}

// x (FP64) += sum_i alpha_i p_i (FP64 accumulation of FP32 p); partials of ||dx||_inf.
template <bool IS3D>
__global__ __launch_bounds__(kRedThreads) void k_accum32(long long n, double* x,
                                                         const float* const* P, int nvec, Red r) {
  const double* alpha = r.res + RS_ALPHA;
  double v[1] = {0.0};
  for (long long q = (long long)blockIdx.x * kRedThreads + threadIdx.x; q < n;
       q += (long long)gridDim.x * kRedThreads) {
    D5 d{0, 0, 0, 0, 0};
    for (int ii = 0; ii < nvec; ++ii) {
      const D5 p = ldf<IS3D>(P[ii], n, q);
      const double a = alpha[ii];
      d = D5{d.a + a * p.a, d.b + a * p.b, d.x + a * p.x, d.y + a * p.y, d.z + a * p.z};
    }
    double2* x2 = reinterpret_cast<double2*>(x);
    const double2 xv = x2[q];
    x2[q] = make_double2(xv.x + d.a, xv.y + d.b);
    x[2 * n + q] += d.x;
    x[3 * n + q] += d.y;
    if (IS3D) x[4 * n + q] += d.z;
    v[0] = fmax(v[0], fmax(hypot(d.a, d.b), fmax(fabs(d.x), fmax(fabs(d.y), fabs(d.z)))));
  }
  publish<0, 1>(v, r.pout);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

namespace {
constexpr int kStream32Cap = 512;
int blocks32(const Geo& g) {
  long long nb = (g.n + kRedThreads - 1) / kRedThreads;
  return (int)(nb < kStream32Cap ? nb : kStream32Cap);
}
}  // namespace

int stream_blocks32(const Geo& g) { return blocks32(g); }

void launch_jvp32(const Geo& g, const Bnd& b, const Mat& mt, const float* x32, const float* p32,
                  float* ap, double hdt, const StencilCfg& cfg, hipStream_t st) {
  Bnd b0 = b;
  b0.cx = b0.cy = b0.cz = 0.0;
  b0.bx = b0.by = b0.bz = nullptr;
  const int tx = (g.nx + TX - 1) / TX, ty = (g.ny + TY - 1) / TY;
  const int nb = stencil_blocks(g, cfg);
  if (g.is3d)
    hipLaunchKernelGGL(k_jvp32<true>, dim3(nb), dim3(BS), 0, st, g, b, b0, mt, x32, p32, ap,
                       (float)hdt, tx, ty, cfg.kz);
  else
    hipLaunchKernelGGL(k_jvp32<false>, dim3(nb), dim3(BS), 0, st, g, b, b0, mt, x32, p32, ap,
                       (float)hdt, tx, ty, cfg.kz);
}

#define LAUNCH_F(kern, ...)                                                                     \
  do {                                                                                          \
    if (g.is3d)                                                                                 \
      hipLaunchKernelGGL(kern<true>, dim3(blocks32(g)), dim3(kRedThreads), 0, st, __VA_ARGS__); \
    else                                                                                        \
      hipLaunchKernelGGL(kern<false>, dim3(blocks32(g)), dim3(kRedThreads), 0, st, __VA_ARGS__); \
  } while (0)

int launch_mdot32(const Geo& g, const float* ap, const float* const* V, int ncols, int with_self,
                  int second, const Red& r, hipStream_t st) {
  LAUNCH_F(k_mdot32, g.n, ap, V, ncols, with_self, second, r);
  return blocks32(g);
}
void launch_colsum(const double* pin, int count, int ncols_total, int n_coef, double* res, int base,
                   int self_slot, int second, hipStream_t st) {
  hipLaunchKernelGGL(k_colsum, dim3(ncols_total), dim3(kRedThreads), 0, st, pin, count, ncols_total,
                     n_coef, res, base, self_slot, second);
}
int launch_cgs_upd32(const Geo& g, float* ap, const float* const* V, int ncols, int coef_base,
                     int second, const Red& r, hipStream_t st) {
  LAUNCH_F(k_cgs_upd32, g.n, ap, V, ncols, coef_base, second, r);
  return blocks32(g);
}
int launch_final_p32(const Geo& g, const float* rv, const float* const* P, int nprev, float* ap,
                     float* pout, int nrm_slot, const Red& r, hipStream_t st) {
  LAUNCH_F(k_final_p32, g.n, rv, P, nprev, ap, pout, nrm_slot, r);
  return blocks32(g);
}
int launch_update_r32(const Geo& g, float* rv, const float* ap, int aslot, const Red& r,
                      hipStream_t st) {
  LAUNCH_F(k_update_r32, g.n, rv, ap, aslot, r);
  return blocks32(g);
}
int launch_accum32(const Geo& g, double* x, const float* const* P, int nvec, const Red& r,
                   hipStream_t st) {
  LAUNCH_F(k_accum32, g.n, x, P, nvec, r);
  return blocks32(g);
}

}  // namespace tdgl
