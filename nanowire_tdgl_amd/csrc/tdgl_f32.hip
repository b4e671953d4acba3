// Mixed-precision kernels of the trapezoidal step, for a batch of films.
//
// FP32: the Jacobian-vector product (analytic, the exact derivative of eval_f)
// and the Krylov vectors (residual r, basis Q).  FP64: every dot product and
// norm accumulates in double, the Gram-Schmidt coefficients and the small
// triangular system are double, and the Newton update accumulates in double
// into the FP64 iterate.  The Newton residual itself is the FP64 stencil
// (tdgl_f64.hip).
//
// TGCR without a search-direction basis.  With p_k = r_k before
// orthogonalisation (tgcr.py:125-127), A r_k = sum_{j<=k} H_jk Q_j where Q_j
// are the normalised, orthogonalised A p_j, H_jk (j < k) the Gram-Schmidt
// coefficients and H_kk = ||A p_k||.  GCR's update x = sum_k alpha_k p_k is
// then x = R y with H y = alpha, R = [r_0 .. r_{K-1}], and since
// r_k = r_0 - sum_{m<k} alpha_m Q_m,
//     x = (sum_k y_k) r_0 - sum_m alpha_m (sum_{k>m} y_k) Q_m,
// one pass over r_0 and the Q basis at the end instead of rebuilding p_k
// from the previous p_j in every iteration.
//
// Orthogonalisation is classical Gram-Schmidt with a conditional second pass
// (Daniel-Gragg-Kaufman-Stewart: repeat when ||Ap'||^2 <= ||Ap||^2 / 2).
//
// Every vector kernel covers all films of the batch in one launch: a block
// belongs to one film (the block tables), inactive films' blocks return at
// once, and per-film partials are summed in block order by one-block-per-film
// kernels (deterministic, no atomics).  Vectors are processed as flat float4
// arrays; each film's slice starts at a multiple of 64 elements and its
// padding is zero.

#include <hip/hip_runtime.h>
#include <math.h>

#include "tdgl_device.hpp"
#include "tdgl_launch.hpp"
#include "tdgl_stage32.hpp"

namespace tdgl {

namespace {

__device__ __forceinline__ void fast_sincos(float a, float& s, float& c) {
  s = __sinf(a);
  c = __cosf(a);
}
constexpr int BS = 256;  // every sweep tile shape holds 256 nodes per plane
// s_waitcnt vmcnt(0) alone (expcnt and lgkmcnt at their no-wait maxima)
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }
}  // namespace

// ---------------------------------------------------------------------------
// k_bjvp32: Ap = p - dt/2 * J(x) p, all in FP32.
// J p is the exact derivative of eval_f (rhs.py:147-239) along p.  The
// boundary rules are affine in the state, so p's ghost values use the same
// rules with a zero applied field; hole-boundary rows are zero.
// ---------------------------------------------------------------------------
//
// CHEB: the same stencil, with the epilogue of one Chebyshev step (ChebStep)
// instead of Ap.  A = I - dt/2 J is linear, so A (s p) = s A p: the source is
// staged unscaled and s enters in the epilogue.
//
// Occupancy: at most 168 VGPRs (3 waves per SIMD, which the 47.5 KB of LDS per
// block also allows: 3 blocks per CU).  The link phases use the hardware
// sine/cosine (v_sin_f32 / v_cos_f32 on x / 2 pi): absolute error ~1e-6 for the
// phases met here, far inside what the inexact Newton iteration needs from the
// JVP (its residual is FP64).
//
// TS: the tile shape (sweep_tx(TS) x sweep_ty(TS) nodes per plane); every node's
// arithmetic is the same for every shape and depth.  GEN: some film of the
// batch has general applied-field vectors; its cells are staged by the
// per-cell rules (stage32), every other film's by the mask plans (stage32u),
// which give the same bits.
// compile-time tag: whether a wave's lanes stage a second cell
template <bool B>
struct SecondCell {
  static constexpr bool value = B;
};

template <bool IS3D, bool CHEB, int TS, bool GEN>
__global__ __launch_bounds__(BS, 3) void k_bjvp32(BatchK bk, SweepPart part, const float* x32,
                                               const float* p32, float* const* Q, const int* iter,
                                               float* out_direct, ChebStep cs) {
  // TS = kSweepShapeFilm: each film's own tile (FilmDesc::tw32 x th32, at most
  // 256 nodes), the shared memory sized for the 64 x 4 tile's 66 x 6 cells
  constexpr bool RT = TS == kSweepShapeFilm;
  constexpr int TX = RT ? 64 : sweep_tx(TS), TY = RT ? 4 : sweep_ty(TS);
  static_assert(TX * TY == BS, "a sweep tile holds one node per thread and plane");
  constexpr int CX = TX + 2, CELLS = CX * (TY + 2);
  static_assert(!RT || CELLS == kSweepCells, "per-film tiles fit the 64 x 4 tile's cells");
  // Shared-memory rings of staged planes.  Plane k + 1 is committed at the top
  // of step k, over the slot of plane k + 1 - RING.  Arrays read on planes
  // k - 1..k + 1 (RING = 4) and arrays read on plane k only (NS = 3) then
  // overwrite a plane last read two steps back, behind the one barrier per
  // step; with RING = 3 (the 64x4 tile: four slots would leave room for two
  // blocks per CU instead of three) a second barrier ends every step.
  constexpr int NS = IS3D ? 3 : 1;
  constexpr int kLds4 = (8 * 4 + 2 * 3) * CELLS * 4 + 3 * CELLS;
  constexpr int RING = IS3D ? ((3 * kLds4 <= 160 * 1024) ? 4 : 3) : 1;
  constexpr int CZ = IS3D ? CELLS : 1;
  // one object, so that a cell's address in every array is one register plus
  // a constant offset
  struct Rings {
    float xr[RING][CELLS], xi[RING][CELLS], xz[RING][CZ];
    float pr[RING][CELLS], pi[RING][CELLS], px[RING][CELLS], py[RING][CELLS], pz[RING][CZ];
    float xx[NS][CELLS], xy[NS][CELLS];
    uint8_t h[NS][CELLS];  // the staged cells' flags (the own node reads its own here)
  };
  __shared__ Rings S;
  float(&X_r)[RING][CELLS] = S.xr;
  float(&X_i)[RING][CELLS] = S.xi;
  float(&X_z)[RING][CZ] = S.xz;
  float(&P_r)[RING][CELLS] = S.pr;
  float(&P_i)[RING][CELLS] = S.pi;
  float(&P_x)[RING][CELLS] = S.px;
  float(&P_y)[RING][CELLS] = S.py;
  float(&P_z)[RING][CZ] = S.pz;
  float(&X_x)[NS][CELLS] = S.xx;
  float(&X_y)[NS][CELLS] = S.xy;
  uint8_t(&H)[NS][CELLS] = S.h;

  // The block's piece of the sweep (film, tile column, z planes [kb, ke)):
  // SweepPart sizes the pieces to the sweep's work.
  const unsigned int L = xcd_remap(blockIdx.x, gridDim.x);
  const int4 seg = part.seg[L];
  const int f = seg.x;
  if (f < 0 || !bk.act[f]) return;  // an empty piece (XCD run padding) or a film at rest
  const FilmDesc* fd = bk.films + f;
  const Geo g = fd->g;  // by value: held in registers for the whole piece
  const Bnd b = fd->b;
  const Mat mt = fd->mt;
  const long long off = fd->off;
  const unsigned int n = (unsigned int)g.n;  // the host keeps 20 n below 2^32
  // the film's own Chebyshev degree: past it the film sits the sweep out; its
  // last W sweep writes v - rho to the final vector
  const double* cf = CHEB ? bk.cheb + (long long)f * kChebStride : nullptr;
  const int df = CHEB ? (int)cf[kChebDeg] : 0;
  if (CHEB && cs.m > df) return;
  const bool fin = CHEB && cs.w && cs.m == df;
  float* const obase = CHEB ? (fin ? (cs.out ? cs.out : Q[*iter]) : cs.mid) : (out_direct ? out_direct : Q[*iter]);
  const W32 out = wplanes32(obase + off, n, IS3D);
  const V32 X = planes32(x32 + off, n, IS3D), P = planes32(p32 + off, n, IS3D);
  const float hdt = (float)bk.dtv[2 * f + 1];
  float c_it = 0.f, c_om = 0.f, c_ss = 1.f, c_sp = 0.f;
  const bool has_v = CHEB && cs.v && (fin || cs.cv), has_p = CHEB && cs.prev;
  V32 CV{}, CP{};
  if (CHEB) {
    c_it = (float)cf[0];
    c_om = (float)cf[cs.m];
    c_ss = cs.src_th ? c_it : 1.f;
    c_sp = cs.prev ? (c_om - 1.f) * (cs.prev_th ? c_it : 1.f) : 0.f;
    if (has_v) CV = planes32(cs.v + off, n, IS3D);
    if (has_p) CP = planes32(cs.prev + off, n, IS3D);
  }
  const int tile_x = seg.y % fd->tiles32_x, tile_y = seg.y / fd->tiles32_x;
  // the film's tile: TXr x TYr nodes, CXr staged cells per row.  With a
  // per-film tile the lanes past TXr x TYr compute nothing (their c0 below
  // stays inside the cells: (255 / TXr + 1)(TXr + 2) + TXr < kSweepCells for
  // TXr >= 16)
  const int TXr = RT ? fd->tw32 : TX, TYr = RT ? fd->th32 : TY, CXr = TXr + 2;
  const int cellsr = RT ? CXr * (TYr + 2) : CELLS;
  const int i0 = 1 + tile_x * TXr, j0 = 1 + tile_y * TYr;
  const int tx = RT ? (int)(threadIdx.x % (unsigned)TXr) : (int)(threadIdx.x % TX);
  const int ty = RT ? (int)(threadIdx.x / (unsigned)TXr) : (int)(threadIdx.x / TX);
  const int i = i0 + tx, j = j0 + ty;
  const bool act = (i <= g.Nx - 1) && (j <= g.Ny - 1) && (!RT || ty < TYr);
  const int kb = seg.z, ke = seg.w;

  const int hc0 = threadIdx.x, hc1 = threadIdx.x + BS;
  const bool has1 = hc1 < cellsr;
  // the columns this thread stages on every plane, as mask plans (uniform
  // field); a general applied field takes the per-cell rules, the column's
  // booleans recomputed on every plane (held across the march, they would
  // occupy scalar registers as lane masks).  A plane's loads are issued a
  // plane ahead; the mask rules apply when the cells go to shared memory, so
  // the loads' latency overlaps the plane of arithmetic in between.
  const Fld32 F = fld32(b, &fd->b);
  const Plan32 pl0 = plan32(g, F, i0 - 1 + hc0 % CXr, j0 - 1 + hc0 / CXr);
  Plan32 pl1 = plan32(g, F, i0 - 1 + hc1 % CXr, j0 - 1 + hc1 / CXr);
  if (!has1) {  // no second cell: its loads all read node 0 (one line each), unused
    pl1.qij = 0u;
    pl1.mij = 0u;
  }
  const bool hf = mt.flags != nullptr;
  // live = false: the march's last step, which has no plane to fetch; it
  // issues the same loads (so that the waits count them exactly) from node 0
  // of plane kp (one line each), and the values are never committed
  auto load_into = [&](auto two, int kp, Raw32& a0, Raw32& a1, bool live = true) __attribute__((always_inline)) {
    constexpr bool H1 = decltype(two)::value;
    if (GEN && F.gen) {  // general field: the rules at once, the cells kept in the load slots
      auto pack = [](Raw32& r, const Cell<float>& v, const Cell<float>& w) {
        r.xa = make_float2(v.pr, v.pi);
        r.xx = v.x;
        r.xy = v.y;
        r.xz = v.z;
        r.pa = make_float2(w.pr, w.pi);
        r.px = w.x;
        r.py = w.y;
        r.pz = w.z;
      };
      unsigned int h0 = hc0, h1 = hc1;
      opaque(h0);
      opaque(h1);
      Cell<float> v, w;
      const Col32 c0 = col32(g, i0 - 1 + h0 % CXr, j0 - 1 + h0 / CXr);
      const unsigned int moff = IS3D ? (unsigned int)g.mk * (unsigned int)kp : 0u;
      stage32<IS3D>(g, F, mt.flags, X, P, c0, kp, v, w);
      pack(a0, v, w);
      a0.h = hf ? (unsigned int)ldu(mt.flags, c0.mij + moff) : 0u;
      if constexpr (H1) {
        const Col32 c1 = col32(g, i0 - 1 + h1 % CXr, j0 - 1 + h1 / CXr);
        stage32<IS3D>(g, F, mt.flags, X, P, c1, kp, v, w);
        pack(a1, v, w);
        a1.h = hf ? (unsigned int)ldu(mt.flags, c1.mij + moff) : 0u;
      }
      return;
    }
    // every lane issues both cells' loads: a column past the tile's cells
    // (!has1) has a plan outside the grid, whose source node and flags index
    // are valid, and its values are not written.  Loads on no branch let the
    // own node's waits below count them exactly.
    const Pl32 pp = plane32<IS3D>(g, F, kp);
    Plan32 c0 = pl0, c1 = pl1;
    if (!live) {
      c0.qij = c1.qij = 0u;
      c0.mij = c1.mij = 0u;
    }
    load32u<IS3D>(mt.flags, X, P, c0, pp, a0);
    if constexpr (H1) load32u<IS3D>(mt.flags, X, P, c1, pp, a1);
  };
  auto put = [&](int slot, int t, int c, const Cell<float>& v, const Cell<float>& w, unsigned int h) {
    H[t][c] = (uint8_t)h;
    X_r[slot][c] = v.pr;
    X_i[slot][c] = v.pi;
    X_x[t][c] = v.x;
    X_y[t][c] = v.y;
    P_r[slot][c] = w.pr;
    P_i[slot][c] = w.pi;
    P_x[slot][c] = w.x;
    P_y[slot][c] = w.y;
    if (IS3D) {
      X_z[slot][c] = v.z;
      P_z[slot][c] = w.z;
    }
  };
  // plane kp's cells, loaded by load_into, to ring slot `slot` (and slot t of
  // the arrays read on one plane only)
  auto commit_from = [&](auto two, int kp, int slot, int t, const Raw32& a0, const Raw32& a1, bool drain = true)
                         __attribute__((always_inline)) {
    constexpr bool H1 = decltype(two)::value;
    // drain: every load of the plane has landed here (all are used below);
    // waiting for all at once tells the compiler's wait tracking that no
    // load is left in flight, so no register reuse later in the plane waits
    // on one.  The prologue's commits leave the next plane's loads in flight
    // (the compiler's own waits cover the committed plane's).
    if (drain) vm_drain();
    Cell<float> v0c, w0c, v1c, w1c;
    if (GEN && F.gen) {
      auto unpack = [](const Raw32& r, Cell<float>& v, Cell<float>& w) {
        v = Cell<float>{r.xa.x, r.xa.y, r.xx, r.xy, r.xz};
        w = Cell<float>{r.pa.x, r.pa.y, r.px, r.py, r.pz};
      };
      unpack(a0, v0c, w0c);
      if constexpr (H1) unpack(a1, v1c, w1c);
    } else {
      const Pl32 pp = plane32<IS3D>(g, F, kp);
      if (IS3D && pp.fk) {
        finish32u<IS3D, true>(a0, hf, pl0, pp, v0c, w0c);
        if constexpr (H1) finish32u<IS3D, true>(a1, hf, pl1, pp, v1c, w1c);
      } else {
        finish32u<IS3D, false>(a0, hf, pl0, pp, v0c, w0c);
        if constexpr (H1) finish32u<IS3D, false>(a1, hf, pl1, pp, v1c, w1c);
      }
    }
    put(slot, t, hc0, v0c, w0c, a0.h);
    if constexpr (H1) {
      if (has1) put(slot, t, hc1, v1c, w1c, a1.h);
    }
  };

  // Ring slots: plane kp sits in slot (kp - kb + 1) % RING ((kp - kb + 1) % NS
  // for the one-plane arrays).  With RING == NS (the 64x4 tile) the march is
  // unrolled by RING, so that every copy of the step knows its slots at
  // compile time.
  auto march = [&](auto two) __attribute__((always_inline)) {
    Raw32 r0, r1;
    if (IS3D) {
      // the first three planes' loads go out together (one memory round trip
      // before the first step: the third plane's loads stay in flight while
      // the first two are committed)
      Raw32 a0, a1, b0, b1;
      load_into(two, kb - 1, a0, a1);
      load_into(two, kb, b0, b1);
      load_into(two, kb + 1, r0, r1);
      commit_from(two, kb - 1, 0, 0, a0, a1, false);
      commit_from(two, kb, 1, 1, b0, b1, false);
    } else {
      load_into(two, 0, r0, r1);
    }

    const float ihx2 = (float)mt.ihx2, ihy2 = (float)mt.ihy2, ihz2 = (float)mt.ihz2;
    const float dgpsi = (float)mt.dgpsi, inv_tau = (float)mt.inv_tau;

    // one step of the march: plane k from slots s0 (plane k), sm (k - 1) and sp1
    // (k + 1, committed first); t0 / t1: planes k / k + 1 in the one-plane arrays
    auto step_k = [&](int k, int s0, int sm, int sp1, int t0, int t1) __attribute__((always_inline)) {
      commit_from(two, IS3D ? k + 1 : 0, sp1, t1, r0, r1);
      __syncthreads();
      // own-node inputs (a lane past the grid's edge reads the nearest node and
      // stores nothing), then the halo two planes ahead
      const unsigned int q = qix32(g, min(i, g.Nx - 1), min(j, g.Ny - 1), k, IS3D);
      const int c0 = (ty + 1) * CXr + (tx + 1);
      const uint8_t fl = H[t0][c0];
      float scv = 1.f, k2v = 0.f;
      if (mt.has_mat && !mt.sc_bits) scv = (float)ldu(mt.sc, q);
      if (mt.k2) k2v = (float)ldu(mt.k2, q);
      // p at the own node is its staged cell (read below): the rules keep psi and
      // zero a link only on the last interior column / row / plane (the normal
      // zeroing) or on a hole link; there the raw value is loaded
      const bool rx = i == g.Nx - 1 || (fl & kHoleX), ry = j == g.Ny - 1 || (fl & kHoleY);
      const bool rz = IS3D && (k == g.Nz - 1 || (fl & kHoleZ));
      float lx = 0.f, ly = 0.f, lz = 0.f;
      if (rx) lx = P.ph(2, q);
      if (ry) ly = P.ph(3, q);
      if (rz) lz = P.ph(4, q);
      float2 v0 = make_float2(0.f, 0.f), w0 = make_float2(0.f, 0.f);
      float v_x = 0.f, v_y = 0.f, v_z = 0.f, w_x = 0.f, w_y = 0.f, w_z = 0.f;
      if (has_v) {
        v0 = CV.psi(q);
        const float3 l = CV.links<IS3D>(q);
        v_x = l.x;
        v_y = l.y;
        v_z = l.z;
      }
      if (has_p) {
        w0 = CP.psi(q);
        const float3 l = CP.links<IS3D>(q);
        w_x = l.x;
        w_y = l.y;
        w_z = l.z;
      }
      // the halo two planes ahead (on the last step, the one-line stand-in)
      if (IS3D) load_into(two, min(k + 2, ke), r0, r1, k + 2 <= ke);

      if (act) {
        const int cxm = c0 - 1, cxp = c0 + 1, cym = c0 - CXr, cyp = c0 + CXr;
        const float pr = X_r[s0][c0], pi = X_i[s0][c0];
        const float dr = P_r[s0][c0], di = P_i[s0][c0];
        const float ex = X_x[t0][c0], ey = X_y[t0][c0], ez = IS3D ? X_z[s0][c0] : 0.f;
        const float dex = P_x[s0][c0], dey = P_y[s0][c0], dez = IS3D ? P_z[s0][c0] : 0.f;
        const float2 p0 = make_float2(dr, di);
        const float praw_x = rx ? lx : dex, praw_y = ry ? ly : dey, praw_z = rz ? lz : dez;

        float sxp, cxp_, sxm, cxm_, syp, cyp_, sym, cym_, szp = 0.f, czp = 1.f, szm = 0.f, czm = 1.f;
        fast_sincos(ex, sxp, cxp_);
        fast_sincos(X_x[t0][cxm], sxm, cxm_);
        fast_sincos(ey, syp, cyp_);
        fast_sincos(X_y[t0][cym], sym, cym_);
        if (IS3D) {
          fast_sincos(ez, szp, czp);
          fast_sincos(X_z[sm][c0], szm, czm);
        }

        // d[e^{-i t} psi(-a)] = e^{-i t}(d psi(-a) - i dt psi(-a)),
        // d[e^{i t} psi(+a)]  = e^{i t}(d psi(+a) + i dt psi(+a))
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
        if (mt.has_mat) {
          const float sc = mt.sc_bits ? ((fl & kInsulator) ? 0.f : 1.f) : scv;
          fr = sc * fr - (1.f - sc) * dr * inv_tau;
          fi = sc * fi - (1.f - sc) * di * inv_tau;
        }
        const float jr = ar + fr, ji = ai + fi;

        float cx, cy, cz;
        if (mt.k2) {
          cx = k2v * (float)mt.ihx2;
          cy = k2v * (float)mt.ihy2;
          cz = IS3D ? k2v * (float)mt.ihz2 : 0.f;
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
        // link operators on the perturbation: 2 D_aa along the link, D_bb across
        float jx = 2.f * cx * (P_x[s0][cxm] - 2.f * dex + P_x[s0][cxp]) +
                   cy * (P_x[s0][cym] - 2.f * dex + P_x[s0][cyp]);
        float jy = 2.f * cy * (P_y[s0][cym] - 2.f * dey + P_y[s0][cyp]) +
                   cx * (P_y[s0][cxm] - 2.f * dey + P_y[s0][cxp]);
        jx += cy * (((-P_y[s0][cxp] + dey) + P_y[s0][cxp - CXr]) - P_y[s0][cym]);
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

        if (CHEB) {
          const float cvf = cs.cv ? 1.f : 0.f;
          auto step = [&](float pv, float jv, float vv, float wv) {
            const float y = c_ss * pv;
            const float ay = c_ss * (pv - hdt * jv);
            const float res = c_om * (y + (cvf * vv - ay) * c_it) - c_sp * wv;
            return fin ? vv - res : res;
          };
          out.psi(q, make_float2(step(p0.x, jr, v0.x, w0.x), step(p0.y, ji, v0.y, w0.y)));
          out.links<IS3D>(q, step(praw_x, jx, v_x, w_x), step(praw_y, jy, v_y, w_y),
                          IS3D ? step(praw_z, jz, v_z, w_z) : 0.f);
        } else {
          out.psi(q, make_float2(p0.x - hdt * jr, p0.y - hdt * ji));
          out.links<IS3D>(q, praw_x - hdt * jx, praw_y - hdt * jy, IS3D ? praw_z - hdt * jz : 0.f);
        }
      }
      if (IS3D && RING == 3) __syncthreads();
    };
    if constexpr (IS3D && RING == NS) {
      static_assert(RING == 3, "the unrolled march has three slots");
      for (int k = kb; k < ke; k += 3) {
        step_k(k, 1, 0, 2, 1, 2);
        if (k + 1 >= ke) break;
        step_k(k + 1, 2, 1, 0, 2, 0);
        if (k + 2 >= ke) break;
        step_k(k + 2, 0, 2, 1, 0, 1);
      }
    } else {
      for (int k = kb; k < ke; ++k) {
        const int r = k - kb + 1;
        step_k(k, r % RING, (r + RING - 1) % RING, (r + 1) % RING, r % NS, (r + 1) % NS);
      }
    }
  };
  // a wave none of whose lanes stages a second cell runs the march without
  // the second cell's loads (the branch is uniform per wave; every wave still
  // passes the same barriers)
  const unsigned int wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63u;
  if (wave0 + BS < (unsigned int)cellsr) march(SecondCell<true>{});
  else march(SecondCell<false>{});
}

// ---------------------------------------------------------------------------
// Flat FP32 vectors with FP64 accumulation.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double dot4(const float4& a, const float4& b) {
  return ((double)a.x * (double)b.x + (double)a.y * (double)b.y) +
         ((double)a.z * (double)b.z + (double)a.w * (double)b.w);
}

// Re-orthogonalise when the first pass removed more than half of ||Ap||^2.
__device__ __forceinline__ bool need_reorth(const double* res) {
  return res[RS_NRM1] <= 0.5 * res[RS_SELF];
}

#define FILM_STREAM(bk)                                                       \
  const int f = (bk).sblk_film[blockIdx.x];                                   \
  if (!(bk).act[f]) return;                                                   \
  const FilmDesc* fd = (bk).films + f;                                        \
  const long long off = fd->off, n4 = (fd->len + 3) >> 2;                     \
  const long long e0 = (long long)(blockIdx.x - fd->sb0) * kRedThreads + threadIdx.x; \
  const long long estride = (long long)fd->nsb * kRedThreads;                 \
  double* res = (bk).res + (long long)f * kResLen;                            \
  (void)res;                                                                  \
  (void)off

constexpr int kMdotChunk = 8;

// Column partials pm[c * nsblk + block] of
//   PASS 0 (first iteration): ||W||^2, <r, W>
//   PASS 1: <W, Q_j> for j < i, then ||W||^2
//   PASS 2 (only when re-orthogonalising): <W, Q_j> for j < i
// with W = Q_i, the new Jacobian-vector product.
template <int PASS>
__global__ __launch_bounds__(kRedThreads) void k_bmdot(BatchK bk, float* const* Q, const float* rv,
                                                       const int* iter, double* pm) {
  FILM_STREAM(bk);
  const int i = *iter;
  if (PASS == 2 && !need_reorth(res)) return;
  const float4* W = reinterpret_cast<const float4*>(Q[i] + off);
  const float4* R = reinterpret_cast<const float4*>(rv + off);
  const int nq = (PASS == 0) ? 0 : i;
  const int ncol = (PASS == 0) ? 2 : (PASS == 1 ? i + 1 : i);
  for (int c0 = 0; c0 < ncol; c0 += kMdotChunk) {
    double acc[kMdotChunk];
    const float4* V[kMdotChunk];
#pragma unroll
    for (int cc = 0; cc < kMdotChunk; ++cc) {
      acc[cc] = 0.0;
      const int c = c0 + cc;
      V[cc] = (c < nq) ? reinterpret_cast<const float4*>(Q[c] + off) : nullptr;
    }
    for (long long e = e0; e < n4; e += estride) {
      const float4 w = W[e];
#pragma unroll
      for (int cc = 0; cc < kMdotChunk; ++cc) {
        const int c = c0 + cc;
        if (c < nq) acc[cc] += dot4(w, V[cc][e]);
        else if (PASS != 2 && c == nq) acc[cc] += dot4(w, w);
        else if (PASS == 0 && c == 1) acc[cc] += dot4(w, R[e]);
      }
    }
    block_reduce<kMdotChunk, 0>(acc);
    if (threadIdx.x == 0)
      for (int cc = 0; cc < kMdotChunk && c0 + cc < ncol; ++cc)
        pm[(long long)(c0 + cc) * bk.nsblk + blockIdx.x] = acc[cc];
  }
}

// W -= sum_{j<i} coef_j Q_j (coefficients from the pass's column sums), then
// partials of ||W||^2 and <r, W> (columns 0 and 1).
template <int PASS>
__global__ __launch_bounds__(kRedThreads) void k_bupd(BatchK bk, float* const* Q, const float* rv,
                                                      const int* iter, double* pm) {
  FILM_STREAM(bk);
  const int i = *iter;
  if (PASS == 2 && !need_reorth(res)) return;
  const double* coef = res + (PASS == 1 ? RS_BETA : RS_DBETA);
  float4* W = reinterpret_cast<float4*>(Q[i] + off);
  const float4* R = reinterpret_cast<const float4*>(rv + off);
  double v[2] = {0.0, 0.0};
  for (long long e = e0; e < n4; e += estride) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int c = 0;
    for (; c + 4 <= i; c += 4) {
      const float4 a = reinterpret_cast<const float4*>(Q[c] + off)[e];
      const float4 bq = reinterpret_cast<const float4*>(Q[c + 1] + off)[e];
      const float4 cq = reinterpret_cast<const float4*>(Q[c + 2] + off)[e];
      const float4 dq = reinterpret_cast<const float4*>(Q[c + 3] + off)[e];
      const double ka = coef[c], kb2 = coef[c + 1], kc = coef[c + 2], kd = coef[c + 3];
      s0 += ((ka * a.x + kb2 * bq.x) + (kc * cq.x + kd * dq.x));
      s1 += ((ka * a.y + kb2 * bq.y) + (kc * cq.y + kd * dq.y));
      s2 += ((ka * a.z + kb2 * bq.z) + (kc * cq.z + kd * dq.z));
      s3 += ((ka * a.w + kb2 * bq.w) + (kc * cq.w + kd * dq.w));
    }
    for (; c < i; ++c) {
      const float4 a = reinterpret_cast<const float4*>(Q[c] + off)[e];
      const double ka = coef[c];
      s0 += ka * a.x;
      s1 += ka * a.y;
      s2 += ka * a.z;
      s3 += ka * a.w;
    }
    const float4 w = W[e];
    const float4 o = make_float4((float)((double)w.x - s0), (float)((double)w.y - s1),
                                 (float)((double)w.z - s2), (float)((double)w.w - s3));
    W[e] = o;
    v[0] += dot4(o, o);
    v[1] += dot4(R[e], o);
  }
  block_reduce<2, 0>(v);
  if (threadIdx.x == 0) {
    pm[blockIdx.x] = v[0];
    pm[(long long)bk.nsblk + blockIdx.x] = v[1];
  }
}

// Q_i = W / ||W||, alpha_i = <r, Q_i>, r -= alpha_i Q_i; H column i; partials
// of ||r||^2.  Stops the film at ||W|| < 1e-14 (tgcr.py:141-142).
__global__ __launch_bounds__(kRedThreads) void k_bfinish(BatchK bk, float* const* Q, float* rv,
                                                         const int* iter, double* H, double* pm) {
  FILM_STREAM(bk);
  const int i = *iter;
  const bool two = i > 0 && need_reorth(res);
  const double nrm2 = two ? res[RS_NRM2] : res[RS_NRM1];
  const double rw = two ? res[RS_RW2] : res[RS_RW1];
  const double nrm = sqrt(nrm2);
  const bool brk = nrm < 1e-14;
  const double inv = 1.0 / nrm;
  const double alpha = rw * inv;
  if (blockIdx.x == fd->sb0) {
    if (threadIdx.x == 0) {
      res[RS_NRM] = nrm;
      res[RS_BRK] = brk ? 1.0 : 0.0;
      if (!brk) res[RS_ALPHA + i] = alpha;
    }
    if (!brk) {
      double* col = H + fd->hoff + (long long)i * (i + 1) / 2;
      for (int jj = threadIdx.x; jj < i; jj += kRedThreads)
        col[jj] = two ? res[RS_BETA + jj] + res[RS_DBETA + jj] : res[RS_BETA + jj];
      if (threadIdx.x == 0) col[i] = nrm;
    }
  }
  if (brk) return;
  float4* W = reinterpret_cast<float4*>(Q[i] + off);
  float4* R = reinterpret_cast<float4*>(rv + off);
  double v[1] = {0.0};
  for (long long e = e0; e < n4; e += estride) {
    const float4 w = W[e];
    const float4 qv = make_float4((float)(w.x * inv), (float)(w.y * inv), (float)(w.z * inv),
                                  (float)(w.w * inv));
    W[e] = qv;
    const float4 r = R[e];
    const float4 rn = make_float4((float)((double)r.x - alpha * qv.x), (float)((double)r.y - alpha * qv.y),
                                  (float)((double)r.z - alpha * qv.z), (float)((double)r.w - alpha * qv.w));
    R[e] = rn;
    v[0] += dot4(rn, rn);
  }
  block_reduce<1, 0>(v);
  if (threadIdx.x == 0) pm[blockIdx.x] = v[0];
}

// Per-film column sums of pm in block order (grid: films x G).
__global__ __launch_bounds__(kRedThreads) void k_bcolsum(BatchK bk, const int* iter, const double* pm,
                                                         int kind) {
  const int f = blockIdx.x;
  if (!bk.act[f]) return;
  const FilmDesc* fd = bk.films + f;
  double* res = bk.res + (long long)f * kResLen;
  const int i = *iter;
  int ncol = 1;
  bool is_max = false;
  switch (kind) {
    case CS_MD0: ncol = 2; break;
    case CS_MD1: ncol = i + 1; break;
    case CS_MD2: if (!need_reorth(res)) return; ncol = i; break;
    case CS_UP1: ncol = 2; break;
    case CS_UP2: if (!need_reorth(res)) return; ncol = 2; break;
    case CS_DXINF: ncol = 1; is_max = true; break;
  }
  for (int c = blockIdx.y; c < ncol; c += gridDim.y) {
    double v[1] = {0.0};
    const double* col = pm + (long long)c * bk.nsblk + fd->sb0;
    for (int bb = threadIdx.x; bb < fd->nsb; bb += kRedThreads)
      v[0] = is_max ? fmax(v[0], col[bb]) : v[0] + col[bb];
    if (is_max) block_reduce<0, 1>(v);
    else block_reduce<1, 0>(v);
    if (threadIdx.x == 0) {
      int slot = 0;
      switch (kind) {
        case CS_MD0: slot = c == 0 ? RS_NRM1 : RS_RW1; break;
        case CS_MD1: slot = c < i ? RS_BETA + c : RS_SELF; break;
        case CS_MD2: slot = RS_DBETA + c; break;
        case CS_UP1: slot = c == 0 ? RS_NRM1 : RS_RW1; break;
        case CS_UP2: slot = c == 0 ? RS_NRM2 : RS_RW2; break;
        case CS_DXINF: slot = RS_DXINF; break;
      }
      res[slot] = v[0];
    }
  }
}

// End of a TGCR iteration: ||r||, the stop tests of tgcr.py:125 and :141,
// and the film's status.  One block per film.
__global__ __launch_bounds__(kRedThreads) void k_bstatus(BatchK bk, int* fstat, const int* iter,
                                                         double tol, const double* pm) {
  const int f = blockIdx.x;
  if (!bk.act[f]) return;
  const FilmDesc* fd = bk.films + f;
  double* res = bk.res + (long long)f * kResLen;
  const int i = *iter;
  int* st = fstat + 4 * f;
  if (res[RS_BRK] != 0.0) {
    if (threadIdx.x == 0) {
      st[0] = TG_BREAK;
      st[1] = i;
      st[2] = i + 1;
      bk.act[f] = 0;
    }
    return;
  }
  double v[1] = {0.0};
  for (int bb = threadIdx.x; bb < fd->nsb; bb += kRedThreads) v[0] += pm[fd->sb0 + bb];
  block_reduce<1, 0>(v);
  if (threadIdx.x == 0) {
    const double rk = sqrt(v[0]);
    res[RS_RR] = v[0];
    res[RS_RK] = rk;
    st[1] = i + 1;
    st[2] = i + 1;
    if (!isfinite(rk)) {
      st[0] = TG_NONFINITE;
      bk.act[f] = 0;
    } else if (!(rk / res[RS_R0] > tol)) {
      st[0] = TG_CONVERGED;
      bk.act[f] = 0;
    } else if (i + 1 >= st[3]) {
      st[0] = TG_MAXITER;
      bk.act[f] = 0;
    }
  }
}

// Advance the device iteration counter and count the films still iterating.
__global__ void k_bnext(BatchK bk, int* iter, int* nact) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int c = 0;
  for (int f = threadIdx.x; f < bk.nfilm; f += blockDim.x) c += bk.act[f] ? 1 : 0;
  atomicAdd(&cnt, c);
  __syncthreads();
  if (threadIdx.x == 0) {
    *iter = *iter + 1;
    *nact = cnt;
  }
}

// Start of TGCR for the active films: r0 = ||r||_2 from the Newton residual's
// reduction (tgcr.py:116-121).  maxk: per-film iteration caps.
__global__ void k_btinit(BatchK bk, int* fstat, const int* maxk, int* iter) {
  for (int f = threadIdx.x; f < bk.nfilm; f += blockDim.x) {
    if (!bk.act[f]) continue;
    double* res = bk.res + (long long)f * kResLen;
    const double r0 = sqrt(res[RS_RR]);
    res[RS_R0] = r0;
    res[RS_RK] = r0;
    res[RS_BRK] = 0.0;
    int* st = fstat + 4 * f;
    st[0] = TG_RUNNING;
    st[1] = 0;
    st[2] = 0;
    st[3] = maxk[f];
    if (r0 == 0.0) {
      st[0] = TG_CONVERGED;
      bk.act[f] = 0;
    } else if (!isfinite(r0)) {
      st[0] = TG_NONFINITE;
      bk.act[f] = 0;
    }
  }
  if (threadIdx.x == 0) *iter = 0;
}

// y = H^{-1} alpha (column-oriented back substitution), then the update's
// coefficients: res[RS_GAMMA] for r_0 and res[RS_BETA + m] for Q_m, m < K-1.
__global__ __launch_bounds__(1024) void k_btrisolve(BatchK bk, const int* fstat, const double* H) {
  __shared__ double rhs[kMaxKrylov], ys[kMaxKrylov];
  const int f = blockIdx.x;
  if (!bk.act[f]) return;
  const FilmDesc* fd = bk.films + f;
  double* res = bk.res + (long long)f * kResLen;
  const int K = fstat[4 * f + 1];
  const double* Hf = H + fd->hoff;
  for (int jj = threadIdx.x; jj < K; jj += blockDim.x) rhs[jj] = res[RS_ALPHA + jj];
  for (int k = K - 1; k >= 0; --k) {
    __syncthreads();
    const double* col = Hf + (long long)k * (k + 1) / 2;
    const double yk = rhs[k] / col[k];
    if (threadIdx.x == 0) ys[k] = yk;
    for (int jj = threadIdx.x; jj < k; jj += blockDim.x) rhs[jj] = rhs[jj] - col[jj] * yk;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;  // s = sum_{k > m} y_k while m runs down
    for (int m = K - 1; m >= 0; --m) {
      if (m < K - 1) res[RS_BETA + m] = -res[RS_ALPHA + m] * s;
      s += ys[m];
    }
    res[RS_GAMMA] = K > 0 ? s : 0.0;
  }
}

// x (FP64, internal layout) += d for FP32 elements el .. el + 3 of the mixed
// layout (psi elements in place, link elements to their node's phi planes,
// the film's padding in place).
__device__ __forceinline__ void add4_internal(double* x, long long el, long long n, int nphi, double d0,
                                              double d1, double d2, double d3) {
  if (el + 3 < 2 * n) {
    double2* X = reinterpret_cast<double2*>(x + el);
    const double2 xa = X[0], xb = X[1];
    X[0] = make_double2(xa.x + d0, xa.y + d1);
    X[1] = make_double2(xb.x + d2, xb.y + d3);
    return;
  }
  const double d[4] = {d0, d1, d2, d3};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const long long i = internal_of_l32(el + t, n, nphi);
    x[i] += d[t];
  }
}

// x (FP64) += gamma r_0 + sum_{m<K-1} coef_m Q_m; partials of ||dx||_inf
// (complex modulus on psi, |.| on phi).
__global__ __launch_bounds__(kRedThreads) void k_baccum(BatchK bk, double* x, const float* r0v,
                                                        float* const* Q, const int* fstat, double* pm) {
  FILM_STREAM(bk);
  const int K = fstat[4 * f + 1];
  const double gam = res[RS_GAMMA];
  const double* coef = res + RS_BETA;
  const float4* R0 = reinterpret_cast<const float4*>(r0v + off);
  double* X = x + off;
  const long long npsi = 2 * fd->g.n;  // psi occupies elements [0, 2n)
  double v[1] = {0.0};
  for (long long e = e0; e < n4; e += estride) {
    const float4 r = R0[e];
    double d0 = gam * r.x, d1 = gam * r.y, d2 = gam * r.z, d3 = gam * r.w;
    for (int m = 0; m < K - 1; ++m) {
      const float4 a = reinterpret_cast<const float4*>(Q[m] + off)[e];
      const double c = coef[m];
      d0 += c * a.x;
      d1 += c * a.y;
      d2 += c * a.z;
      d3 += c * a.w;
    }
    const long long el = 4 * e;
    add4_internal(X, el, fd->g.n, fd->g.nphi, d0, d1, d2, d3);
    const double ma = (el < npsi) ? hypot(d0, d1) : fmax(fabs(d0), fabs(d1));
    const double mb = (el + 2 < npsi) ? hypot(d2, d3) : fmax(fabs(d2), fabs(d3));
    v[0] = fmax(v[0], fmax(ma, mb));
  }
  block_reduce<0, 1>(v);
  if (threadIdx.x == 0) pm[blockIdx.x] = v[0];
}

// u = gamma r_0 + sum_{m<K-1} coef_m Q_m in FP32: with the Chebyshev
// preconditioner M the Newton update is x += M u (GCR with p_k = M r_k and a
// fixed polynomial M gives x = M R y).  u may alias r0v (same element, same
// thread).
__global__ __launch_bounds__(kRedThreads) void k_baccum_u(BatchK bk, float* u, const float* r0v,
                                                          float* const* Q, const int* fstat) {
  FILM_STREAM(bk);
  const int K = fstat[4 * f + 1];
  const double gam = res[RS_GAMMA];
  const double* coef = res + RS_BETA;
  const float4* R0 = reinterpret_cast<const float4*>(r0v + off);
  float4* U = reinterpret_cast<float4*>(u + off);
  for (long long e = e0; e < n4; e += estride) {
    const float4 r = R0[e];
    double d0 = gam * r.x, d1 = gam * r.y, d2 = gam * r.z, d3 = gam * r.w;
    for (int m = 0; m < K - 1; ++m) {
      const float4 a = reinterpret_cast<const float4*>(Q[m] + off)[e];
      const double c = coef[m];
      d0 += c * a.x;
      d1 += c * a.y;
      d2 += c * a.z;
      d3 += c * a.w;
    }
    U[e] = make_float4((float)d0, (float)d1, (float)d2, (float)d3);
  }
}

// x (FP64) += z_f, the film's Chebyshev iterate z_{d_f}: u / theta for d_f = 1
// (M = 1/theta), else the buffer t[d_f % 2] its last z sweep wrote; partials
// of ||z_f||_inf (complex modulus on psi, |.| on phi).
__global__ __launch_bounds__(kRedThreads) void k_baddz(BatchK bk, double* x, const float* u,
                                                       const float* t0, const float* t1, double* pm) {
  FILM_STREAM(bk);
  const double* cf = bk.cheb + (long long)f * kChebStride;
  const int df = (int)cf[kChebDeg];
  const double s = df < 2 ? cf[0] : 1.0;
  const float* z = df < 2 ? u : ((df & 1) ? t1 : t0);
  const float4* Z = reinterpret_cast<const float4*>(z + off);
  double* X = x + off;
  const long long npsi = 2 * fd->g.n;
  double v[1] = {0.0};
  for (long long e = e0; e < n4; e += estride) {
    const float4 a = Z[e];
    const double d0 = s * a.x, d1 = s * a.y, d2 = s * a.z, d3 = s * a.w;
    const long long el = 4 * e;
    add4_internal(X, el, fd->g.n, fd->g.nphi, d0, d1, d2, d3);
    const double ma = (el < npsi) ? hypot(d0, d1) : fmax(fabs(d0), fabs(d1));
    const double mb = (el + 2 < npsi) ? hypot(d2, d3) : fmax(fabs(d2), fabs(d3));
    v[0] = fmax(v[0], fmax(ma, mb));
  }
  block_reduce<0, 1>(v);
  if (threadIdx.x == 0) pm[blockIdx.x] = v[0];
}

// out = x + r (x - xp) with r = ratio[f]: Newton's start extrapolated from the
// film's last accepted step (r <= 0: out keeps the reference predictor).
__global__ __launch_bounds__(kRedThreads) void k_bextrap(BatchK bk, const double* x, const double* xp,
                                                         const double* ratio, double* out) {
  FILM_STREAM(bk);
  (void)n4;
  const double r = ratio[f];
  if (!(r > 0.0)) return;
  // the film's own elements only: the caller's padding is never read
  const long long len = fd->len;
  const double* X = x + off;
  const double* XP = xp + off;
  double* O = out + off;
  for (long long i = e0; i < len; i += estride) O[i] = X[i] + r * (X[i] - XP[i]);
}

// dst = src on the films selected by sel (the film's own elements; padding is
// never touched).
__global__ __launch_bounds__(kRedThreads) void k_bcopy_films(BatchK bk, const int* sel, const double* src,
                                                             double* dst) {
  const int f = bk.sblk_film[blockIdx.x];
  if (!sel[f]) return;
  const FilmDesc* fd = bk.films + f;
  const long long off = fd->off, len = fd->len;
  const long long e0 = (long long)(blockIdx.x - fd->sb0) * kRedThreads + threadIdx.x;
  const long long estride = (long long)fd->nsb * kRedThreads;
  for (long long i = e0; i < len; i += estride) dst[off + i] = src[off + i];
}

// One film's FP32 vector between the internal layout and the mixed method's
// (to_l32: out = the mixed layout of in; else the internal layout of in).
__global__ __launch_bounds__(256) void k_l32_permute(const float* in, float* out, long long n, int nphi,
                                                     int to_l32) {
  const long long len = (2 + nphi) * n;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
       e += (long long)gridDim.x * blockDim.x)
    out[e] = in[to_l32 ? internal_of_l32(e, n, nphi) : l32_of_internal(e, n, nphi)];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

void launch_l32_permute(const float* in, float* out, long long n, int nphi, bool to_l32, hipStream_t st) {
  const long long len = (2 + nphi) * n;
  const int blocks = (int)std::min<long long>((len + 255) / 256, 4096);
  launch(k_l32_permute, dim3(blocks), dim3(256), 0, st, in, out, n, nphi, to_l32 ? 1 : 0);
}

void launch_bcopy_films(const BatchK& bk, const int* sel, const double* src, double* dst, hipStream_t st) {
  launch(k_bcopy_films, dim3(bk.nsblk), dim3(kRedThreads), 0, st, bk, sel, src, dst);
}

void launch_bextrap(const BatchK& bk, const double* x, const double* xp, const double* ratio, double* out,
                    hipStream_t st) {
  launch(k_bextrap, dim3(bk.nsblk), dim3(kRedThreads), 0, st, bk, x, xp, ratio, out);
}

template <bool IS3D, bool CHEB, bool GEN>
void launch_sweep_g(const BatchK& bk, const SweepPart& part, const float* x32, const float* p32,
                    float* const* Q, const int* iter, float* out, const ChebStep& cs, hipStream_t st) {
  const dim3 gr(part.nblocks), bl(BS);
  switch (bk.shape) {
    case 1: launch((k_bjvp32<IS3D, CHEB, 1, GEN>), gr, bl, 0, st, bk, part, x32, p32, Q, iter, out, cs); break;
    case 2: launch((k_bjvp32<IS3D, CHEB, 2, GEN>), gr, bl, 0, st, bk, part, x32, p32, Q, iter, out, cs); break;
    case kSweepShapeFilm:
      launch((k_bjvp32<IS3D, CHEB, kSweepShapeFilm, GEN>), gr, bl, 0, st, bk, part, x32, p32, Q, iter, out, cs);
      break;
    default: launch((k_bjvp32<IS3D, CHEB, 0, GEN>), gr, bl, 0, st, bk, part, x32, p32, Q, iter, out, cs); break;
  }
}
template <bool IS3D, bool CHEB>
void launch_sweep(const BatchK& bk, const SweepPart& part, const float* x32, const float* p32,
                  float* const* Q, const int* iter, float* out, const ChebStep& cs, hipStream_t st) {
  if (part.nblocks <= 0) return;
  if (bk.gen) launch_sweep_g<IS3D, CHEB, true>(bk, part, x32, p32, Q, iter, out, cs, st);
  else launch_sweep_g<IS3D, CHEB, false>(bk, part, x32, p32, Q, iter, out, cs, st);
}
void launch_bjvp32(const BatchK& bk, const SweepPart& part, bool is3d, const float* x32, const float* p32,
                   float* const* Q, const int* iter, float* out_direct, hipStream_t st) {
  const ChebStep none{};
  if (is3d) launch_sweep<true, false>(bk, part, x32, p32, Q, iter, out_direct, none, st);
  else launch_sweep<false, false>(bk, part, x32, p32, Q, iter, out_direct, none, st);
}
void launch_bcheb32(const BatchK& bk, const SweepPart& part, bool is3d, const float* x32, const float* p32,
                    float* const* Q, const int* iter, const ChebStep& cs, hipStream_t st) {
  if (is3d) launch_sweep<true, true>(bk, part, x32, p32, Q, iter, nullptr, cs, st);
  else launch_sweep<false, true>(bk, part, x32, p32, Q, iter, nullptr, cs, st);
}
int sweep_blocks_per_cu(bool is3d, int shape, bool cheb) {
  int n = 0;
  hipError_t e;
  auto q = [&](auto kern) { e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, BS, 0); };
  if (is3d) {
    if (cheb) {
      if (shape == 1) q(k_bjvp32<true, true, 1, false>);
      else if (shape == 2) q(k_bjvp32<true, true, 2, false>);
      else if (shape == kSweepShapeFilm) q(k_bjvp32<true, true, kSweepShapeFilm, false>);
      else q(k_bjvp32<true, true, 0, false>);
    } else {
      if (shape == 1) q(k_bjvp32<true, false, 1, false>);
      else if (shape == 2) q(k_bjvp32<true, false, 2, false>);
      else if (shape == kSweepShapeFilm) q(k_bjvp32<true, false, kSweepShapeFilm, false>);
      else q(k_bjvp32<true, false, 0, false>);
    }
  } else {
    q(k_bjvp32<false, true, 0, false>);
  }
  return (e == hipSuccess && n > 0) ? n : 1;
}
void launch_baccum_u(const BatchK& bk, float* u, const float* r0v, float* const* Q, const int* fstat,
                     hipStream_t st) {
  launch(k_baccum_u, dim3(bk.nsblk), dim3(kRedThreads), 0, st, bk, u, r0v, Q, fstat);
}
void launch_baddz(const BatchK& bk, double* x, const float* u, const float* t0, const float* t1, double* pm,
                  hipStream_t st) {
  launch(k_baddz, dim3(bk.nsblk), dim3(kRedThreads), 0, st, bk, x, u, t0, t1, pm);
}
void launch_bmdot(const BatchK& bk, int pass, float* const* Q, const float* rv, const int* iter,
                  double* pm, hipStream_t st) {
  const dim3 gr(bk.nsblk), bl(kRedThreads);
  if (pass == 0) launch(k_bmdot<0>, gr, bl, 0, st, bk, Q, rv, iter, pm);
  else if (pass == 1) launch(k_bmdot<1>, gr, bl, 0, st, bk, Q, rv, iter, pm);
  else launch(k_bmdot<2>, gr, bl, 0, st, bk, Q, rv, iter, pm);
}
void launch_bupd(const BatchK& bk, int pass, float* const* Q, const float* rv, const int* iter,
                 double* pm, hipStream_t st) {
  const dim3 gr(bk.nsblk), bl(kRedThreads);
  if (pass == 1) launch(k_bupd<1>, gr, bl, 0, st, bk, Q, rv, iter, pm);
  else launch(k_bupd<2>, gr, bl, 0, st, bk, Q, rv, iter, pm);
}
void launch_bfinish(const BatchK& bk, float* const* Q, float* rv, const int* iter, double* H,
                    double* pm, hipStream_t st) {
  launch(k_bfinish, dim3(bk.nsblk), dim3(kRedThreads), 0, st, bk, Q, rv, iter, H, pm);
}
void launch_bcolsum(const BatchK& bk, const int* iter, const double* pm, int kind, int gy,
                    hipStream_t st) {
  launch(k_bcolsum, dim3(bk.nfilm, gy), dim3(kRedThreads), 0, st, bk, iter, pm, kind);
}
void launch_bstatus(const BatchK& bk, int* fstat, const int* iter, double tol, const double* pm,
                    hipStream_t st) {
  launch(k_bstatus, dim3(bk.nfilm), dim3(kRedThreads), 0, st, bk, fstat, iter, tol, pm);
}
void launch_bnext(const BatchK& bk, int* iter, int* nact, hipStream_t st) {
  launch(k_bnext, dim3(1), dim3(256), 0, st, bk, iter, nact);
}
void launch_btinit(const BatchK& bk, int* fstat, const int* maxk, int* iter, hipStream_t st) {
  launch(k_btinit, dim3(1), dim3(256), 0, st, bk, fstat, maxk, iter);
}
void launch_btrisolve(const BatchK& bk, const int* fstat, const double* H, hipStream_t st) {
  launch(k_btrisolve, dim3(bk.nfilm), dim3(1024), 0, st, bk, fstat, H);
}
void launch_baccum(const BatchK& bk, double* x, const float* r0v, float* const* Q, const int* fstat,
                   double* pm, hipStream_t st) {
  launch(k_baccum, dim3(bk.nsblk), dim3(kRedThreads), 0, st, bk, x, r0v, Q, fstat, pm);
}

}  // namespace tdgl
