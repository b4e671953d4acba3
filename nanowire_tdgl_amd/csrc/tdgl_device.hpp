// Device-side building blocks shared by the TDGL kernels (gfx950 / CDNA4).
//
// Layout (internal): interior nodes only, i fastest:
//   q = (i-1) + nx*((j-1) + ny*(k-1))        (3-D; 2-D uses k = 0 and no k term)
// A state buffer of doubles is [psi re/im interleaved: 2n][phi_x: n][phi_y: n][phi_z: n].
// Ghost (boundary) values are never stored: every kernel that needs them
// rebuilds them from the interior values with the boundary rules of
// tdgl3d's _apply_boundary_conditions (src/tdgl3d/physics/rhs.py:54-139).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "tdgl_types.hpp"

namespace tdgl {

constexpr int kRedThreads = 256;  // block size of every reducing kernel

// Node indices are formed in 32-bit arithmetic: the host rejects grids whose
// full-grid size or state length reaches 2^31.
__host__ __device__ __forceinline__ long long qix(const Geo& g, int i, int j, int k) {
  return (long long)((i - 1) + g.nx * ((j - 1) + g.ny * (g.is3d ? k - 1 : 0)));
}

__host__ __device__ __forceinline__ long long mfull(const Geo& g, int i, int j, int k) {
  return (long long)(i + g.mj * (j + (g.Ny + 1) * (g.is3d ? k : 0)));
}

// Raw interior values of a state, optionally perturbed: x + eps*p, the
// argument of f(x_lin + epsilon*pk) in tgcr.py:130.  The two roundings of the
// reference (eps*p, then +x) are kept: this file is compiled with
// -ffp-contract=off.
template <typename T, bool PERT>
struct Src {
  const T* x;
  const T* p;
  T eps;
  long long n;
  __host__ __device__ __forceinline__ T ld(long long off) const {
    T v = x[off];
    if (PERT) v = v + eps * p[off];
    return v;
  }
  // psi(q) as one vector load (re, im are adjacent; buffers are 8-byte aligned per pair)
  __host__ __device__ __forceinline__ void psi(long long q, T& re, T& im) const {
    using V2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    const V2 a = reinterpret_cast<const V2*>(x)[q];
    re = a.x;
    im = a.y;
    if (PERT) {
      const V2 d = reinterpret_cast<const V2*>(p)[q];
      re = re + eps * d.x;
      im = im + eps * d.y;
    }
  }
  __host__ __device__ __forceinline__ T phx(long long q) const { return ld(2 * n + q); }
  __host__ __device__ __forceinline__ T phy(long long q) const { return ld(3 * n + q); }
  __host__ __device__ __forceinline__ T phz(long long q) const { return ld(4 * n + q); }
};

// Effective (post-boundary-condition) values at one full-grid node.
template <typename T>
struct Cell {
  T pr, pi, x, y, z;
};

// Field terms that the ghost rules add: cz = (Bz*hx)*hy, cy = (By*hz)*hx,
// cx = (Bx*hy)*hz, exactly as rhs.py:95-127 groups them.
template <typename T>
__host__ __device__ __forceinline__ T fld_cz(const Geo& g, const Bnd& b, long long m) {
  return b.bz ? (T)((b.bz[m] * b.hx) * b.hy) : (T)b.cz;
}
template <typename T>
__host__ __device__ __forceinline__ T fld_cy(const Geo& g, const Bnd& b, long long m) {
  return b.by ? (T)((b.by[m] * b.hz) * b.hx) : (T)b.cy;
}
template <typename T>
__host__ __device__ __forceinline__ T fld_cx(const Geo& g, const Bnd& b, long long m) {
  return b.bx ? (T)((b.bx[m] * b.hy) * b.hz) : (T)b.cx;
}

// The value tdgl3d's eval_f sees at full-grid node (i, j, k) after
// _apply_boundary_conditions:
//  (1) phi_x := 0 on i in {0, Nx-1}, phi_y on j in {0, Ny-1}, phi_z on k in {0, Nz-1}
//      (rhs.py:73-76; this also hits the interior planes i = Nx-1 etc.);
//  (2) face-inner ghosts from the zeroed copies (rhs.py:85-127): psi copied
//      from the first interior layer (the e^{-+i phi} factor is e^0 there),
//      transverse links shifted by the applied-field flux, or the periodic
//      image;
//  (3) edges and corners stay 0;
//  (4) hole-boundary links := 0 last (rhs.py:132-137).
template <typename T, bool IS3D, bool PERT>
__host__ __device__ Cell<T> eff_cell(const Geo& g, const Bnd& b, const uint8_t* flags, const Src<T, PERT>& s,
                            int i, int j, int k) {
  Cell<T> c{T(0), T(0), T(0), T(0), T(0)};
  if (i < 0 || i > g.Nx || j < 0 || j > g.Ny) return c;
  if (IS3D && (k < 0 || k > g.Nz)) return c;
  const bool bi = (i == 0) || (i == g.Nx);
  const bool bj = (j == 0) || (j == g.Ny);
  const bool bk = IS3D && ((k == 0) || (k == g.Nz));
  const int nb = (int)bi + (int)bj + (int)bk;
  if (nb > 1) return c;  // edge or corner: never assigned
  if (nb == 0) {
    const long long q = qix(g, i, j, k);
    s.psi(q, c.pr, c.pi);
    c.x = (i == g.Nx - 1) ? T(0) : s.phx(q);
    c.y = (j == g.Ny - 1) ? T(0) : s.phy(q);
    if (IS3D) c.z = (k == g.Nz - 1) ? T(0) : s.phz(q);
  } else {
    const long long m = mfull(g, i, j, k);
    if (bi) {
      if (g.per_x) {  // rhs.py:85-89
        const long long q = qix(g, (i == 0) ? g.Nx - 1 : 1, j, k);
        s.psi(q, c.pr, c.pi);
        c.x = (i == 0 || g.Nx - 1 == 1) ? T(0) : s.phx(q);
      } else {        // rhs.py:91-98
        const int si = (i == 0) ? 1 : g.Nx - 1;
        const long long q = qix(g, si, j, k);
        s.psi(q, c.pr, c.pi);
        const T cz = fld_cz<T>(g, b, m);
        const T sy = (j == g.Ny - 1) ? T(0) : s.phy(q);
        c.y = (i == 0) ? (-cz + sy) : (cz + sy);
        if (IS3D) {
          const T cy = fld_cy<T>(g, b, m);
          const T sz = (k == g.Nz - 1) ? T(0) : s.phz(q);
          c.z = (i == 0) ? (cy + sz) : (-cy + sz);
        }
      }
    } else if (bj) {
      if (g.per_y) {  // rhs.py:101-105
        const long long q = qix(g, i, (j == 0) ? g.Ny - 1 : 1, k);
        s.psi(q, c.pr, c.pi);
        c.y = (j == 0 || g.Ny - 1 == 1) ? T(0) : s.phy(q);
      } else {        // rhs.py:107-112
        const int sj = (j == 0) ? 1 : g.Ny - 1;
        const long long q = qix(g, i, sj, k);
        s.psi(q, c.pr, c.pi);
        const T cz = fld_cz<T>(g, b, m);
        const T sx = (i == g.Nx - 1) ? T(0) : s.phx(q);
        c.x = (j == 0) ? (cz + sx) : (-cz + sx);
        if (IS3D) {
          const T cx = fld_cx<T>(g, b, m);
          const T sz = (k == g.Nz - 1) ? T(0) : s.phz(q);
          c.z = (j == 0) ? (-cx + sz) : (cx + sz);
        }
      }
    } else {  // bk (3-D only)
      if (g.per_z) {  // rhs.py:116-120
        const long long q = qix(g, i, j, (k == 0) ? g.Nz - 1 : 1);
        s.psi(q, c.pr, c.pi);
        c.z = (k == 0 || g.Nz - 1 == 1) ? T(0) : s.phz(q);
      } else {        // rhs.py:122-127
        const int sk = (k == 0) ? 1 : g.Nz - 1;
        const long long q = qix(g, i, j, sk);
        s.psi(q, c.pr, c.pi);
        const T cy = fld_cy<T>(g, b, m);
        const T cx = fld_cx<T>(g, b, m);
        const T sx = (i == g.Nx - 1) ? T(0) : s.phx(q);
        const T sy = (j == g.Ny - 1) ? T(0) : s.phy(q);
        c.x = (k == 0) ? (-cy + sx) : (cy + sx);
        c.y = (k == 0) ? (cx + sy) : (-cx + sy);
      }
    }
  }
  if (flags) {  // rhs.py:132-137 (applied last)
    const uint8_t f = flags[mfull(g, i, j, k)];
    if (f & kHoleX) c.x = T(0);
    if (f & kHoleY) c.y = T(0);
    if (IS3D && (f & kHoleZ)) c.z = T(0);
  }
  return c;
}

// eff_cell for a state x (applied field b) and a perturbation p (zero field)
// at the same node: the boundary rules are affine, so p's ghosts follow the
// same rules without the field terms.  One pass over the index logic.
template <typename T, bool IS3D>
__host__ __device__ __forceinline__ void eff_cell_pair(const Geo& g, const Bnd& b, const uint8_t* flags,
                                              const Src<T, false>& sx, const Src<T, false>& sp,
                                              int i, int j, int k, Cell<T>& cx, Cell<T>& cp) {
  cx = Cell<T>{T(0), T(0), T(0), T(0), T(0)};
  cp = cx;
  if (i < 0 || i > g.Nx || j < 0 || j > g.Ny) return;
  if (IS3D && (k < 0 || k > g.Nz)) return;
  const bool bi = (i == 0) || (i == g.Nx);
  const bool bj = (j == 0) || (j == g.Ny);
  const bool bk = IS3D && ((k == 0) || (k == g.Nz));
  const int nb = (int)bi + (int)bj + (int)bk;
  if (nb > 1) return;
  if (nb == 0) {
    const long long q = qix(g, i, j, k);
    sx.psi(q, cx.pr, cx.pi);
    sp.psi(q, cp.pr, cp.pi);
    if (i != g.Nx - 1) {
      cx.x = sx.phx(q);
      cp.x = sp.phx(q);
    }
    if (j != g.Ny - 1) {
      cx.y = sx.phy(q);
      cp.y = sp.phy(q);
    }
    if (IS3D && k != g.Nz - 1) {
      cx.z = sx.phz(q);
      cp.z = sp.phz(q);
    }
  } else {
    const long long m = mfull(g, i, j, k);
    if (bi) {
      if (g.per_x) {
        const long long q = qix(g, (i == 0) ? g.Nx - 1 : 1, j, k);
        sx.psi(q, cx.pr, cx.pi);
        sp.psi(q, cp.pr, cp.pi);
        if (!(i == 0 || g.Nx - 1 == 1)) {
          cx.x = sx.phx(q);
          cp.x = sp.phx(q);
        }
      } else {
        const long long q = qix(g, (i == 0) ? 1 : g.Nx - 1, j, k);
        sx.psi(q, cx.pr, cx.pi);
        sp.psi(q, cp.pr, cp.pi);
        const T cz = fld_cz<T>(g, b, m);
        const T yx = (j == g.Ny - 1) ? T(0) : sx.phy(q);
        cp.y = (j == g.Ny - 1) ? T(0) : sp.phy(q);
        cx.y = (i == 0) ? (-cz + yx) : (cz + yx);
        if (IS3D) {
          const T cy = fld_cy<T>(g, b, m);
          const T zx = (k == g.Nz - 1) ? T(0) : sx.phz(q);
          cp.z = (k == g.Nz - 1) ? T(0) : sp.phz(q);
          cx.z = (i == 0) ? (cy + zx) : (-cy + zx);
        }
      }
    } else if (bj) {
      if (g.per_y) {
        const long long q = qix(g, i, (j == 0) ? g.Ny - 1 : 1, k);
        sx.psi(q, cx.pr, cx.pi);
        sp.psi(q, cp.pr, cp.pi);
        if (!(j == 0 || g.Ny - 1 == 1)) {
          cx.y = sx.phy(q);
          cp.y = sp.phy(q);
        }
      } else {
        const long long q = qix(g, i, (j == 0) ? 1 : g.Ny - 1, k);
        sx.psi(q, cx.pr, cx.pi);
        sp.psi(q, cp.pr, cp.pi);
        const T cz = fld_cz<T>(g, b, m);
        const T xx = (i == g.Nx - 1) ? T(0) : sx.phx(q);
        cp.x = (i == g.Nx - 1) ? T(0) : sp.phx(q);
        cx.x = (j == 0) ? (cz + xx) : (-cz + xx);
        if (IS3D) {
          const T cxf = fld_cx<T>(g, b, m);
          const T zx = (k == g.Nz - 1) ? T(0) : sx.phz(q);
          cp.z = (k == g.Nz - 1) ? T(0) : sp.phz(q);
          cx.z = (j == 0) ? (-cxf + zx) : (cxf + zx);
        }
      }
    } else {
      if (g.per_z) {
        const long long q = qix(g, i, j, (k == 0) ? g.Nz - 1 : 1);
        sx.psi(q, cx.pr, cx.pi);
        sp.psi(q, cp.pr, cp.pi);
        if (!(k == 0 || g.Nz - 1 == 1)) {
          cx.z = sx.phz(q);
          cp.z = sp.phz(q);
        }
      } else {
        const long long q = qix(g, i, j, (k == 0) ? 1 : g.Nz - 1);
        sx.psi(q, cx.pr, cx.pi);
        sp.psi(q, cp.pr, cp.pi);
        const T cy = fld_cy<T>(g, b, m);
        const T cxf = fld_cx<T>(g, b, m);
        const T xx = (i == g.Nx - 1) ? T(0) : sx.phx(q);
        const T yx = (j == g.Ny - 1) ? T(0) : sx.phy(q);
        cp.x = (i == g.Nx - 1) ? T(0) : sp.phx(q);
        cp.y = (j == g.Ny - 1) ? T(0) : sp.phy(q);
        cx.x = (k == 0) ? (-cy + xx) : (cy + xx);
        cx.y = (k == 0) ? (cxf + yx) : (-cxf + yx);
      }
    }
  }
  if (flags) {
    const uint8_t f = flags[mfull(g, i, j, k)];
    if (f & kHoleX) cx.x = cp.x = T(0);
    if (f & kHoleY) cx.y = cp.y = T(0);
    if (IS3D && (f & kHoleZ)) cx.z = cp.z = T(0);
  }
}

// ---------------------------------------------------------------------------
// Deterministic grid reductions without atomics.
//
// A producing kernel reduces its values inside each block in a fixed tree
// (wave shuffles, then the waves in order) and stores one partial per block.
// The kernel boundary orders those stores before the next kernel, which
// reduces the partials in block order in its prologue -- every consuming
// block computes the same bits.  (A single ticket counter serialised ~1000
// same-address atomics per launch: ~40 us at 242k nodes.)
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// NS sums then NM maxima over the block (blockDim.x == BT); the result is
// valid in every thread.  Wave results are combined in wave order by one
// thread per quantity.
template <int NS, int NM, int BT = kRedThreads>
__device__ __forceinline__ void block_reduce(double (&v)[NS + NM]) {
  constexpr int NV = NS + NM;
  static_assert(NV <= BT, "one thread per reduced quantity");
  __shared__ double sh[BT / 64][NV];
  __shared__ double tot[NV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NV; ++a) {
    const double r = (a < NS) ? wave_sum(v[a]) : wave_max(v[a]);
    if (lane == 0) sh[wid][a] = r;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int a = threadIdx.x;
    double r = sh[0][a];
    for (int w = 1; w < BT / 64; ++w) r = (a < NS) ? r + sh[w][a] : fmax(r, sh[w][a]);
    tot[a] = r;
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < NV; ++a) v[a] = tot[a];
  __syncthreads();
}

// Store this block's partials: pout[blockIdx.x * NV + a].
template <int NS, int NM, int BT = kRedThreads>
__device__ __forceinline__ void publish(double (&v)[NS + NM], double* pout) {
  block_reduce<NS, NM, BT>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < NS + NM; ++a) pout[(long long)blockIdx.x * (NS + NM) + a] = v[a];
  }
}

// Store this block's partials at pout[idx * NV + a].
template <int NS, int NM>
__device__ __forceinline__ void publish_at(double (&v)[NS + NM], double* pout, long long idx) {
  block_reduce<NS, NM>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < NS + NM; ++a) pout[idx * (NS + NM) + a] = v[a];
  }
}

// Reduce `count` published partials (block order); valid in every thread.
template <int NS, int NM, int BT = kRedThreads>
__device__ __forceinline__ void gather(const double* pin, int count, double (&out)[NS + NM]) {
  constexpr int NV = NS + NM;
#pragma unroll
  for (int a = 0; a < NV; ++a) out[a] = 0.0;
  for (int b = threadIdx.x; b < count; b += BT) {
#pragma unroll
    for (int a = 0; a < NV; ++a) {
      const double r = pin[(long long)b * NV + a];
      out[a] = (a < NS) ? out[a] + r : fmax(out[a], r);
    }
  }
  block_reduce<NS, NM, BT>(out);
}

// Block-index remap so that consecutive logical tiles land on the same XCD
// (hardware dispatches block b to XCD b % 8), keeping stencil halos inside one L2.
__device__ __forceinline__ unsigned int xcd_remap(unsigned int b, unsigned int nblocks) {
  const unsigned int nx = 8;
  const unsigned int per = (nblocks + nx - 1) / nx;
  const unsigned int xcd = b % nx, slot = b / nx;
  const unsigned int full = nblocks % nx == 0 ? nx : nblocks % nx;  // XCDs holding `per` blocks
  unsigned int base = (xcd < full) ? xcd * per : full * per + (xcd - full) * (per - 1);
  return base + slot;
}

}  // namespace tdgl
