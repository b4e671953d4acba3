// Staging of one grid cell for the FP32 stencil sweeps: the values tdgl3d's
// eval_f sees at a full-grid node after _apply_boundary_conditions
// (src/tdgl3d/physics/rhs.py:54-139), for a state x (with the applied field)
// and a perturbation p (the same affine rules without the field), as
// eff_cell_pair computes them, in a form with few branches and 32-bit indices.
//
// Every cell takes its values from one source node -- itself, its face-inner
// neighbour (ghost cells) or its periodic image -- and then, per link
// component: the normal zeroing at the source node (rhs.py:73-76), the field
// flux of the ghost rules (rhs.py:95-127), the components the ghost rules do
// not assign (0), and the hole-boundary zeroing at the cell (rhs.py:132-137).
// Edges, corners and cells outside the grid are 0.
#pragma once

#include "tdgl_device.hpp"

namespace tdgl {

// Element e of an array from a uniform base with a 32-bit byte offset (the
// global_load saddr form: no 64-bit address arithmetic per lane).  Callers keep
// arrays below 2^30 bytes / sizeof(T) elements.
template <typename T>
__host__ __device__ __forceinline__ T ldu(const T* base, unsigned int e) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + e * (unsigned int)sizeof(T));
}
template <typename T>
__host__ __device__ __forceinline__ void stu(T* base, unsigned int e, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + e * (unsigned int)sizeof(T)) = v;
}

// The four planes of one FP32 state-shaped vector of n nodes.
struct V32 {
  const float2* psi;
  const float* px;
  const float* py;
  const float* pz;
};
__host__ __device__ __forceinline__ V32 planes32(const float* v, unsigned int n) {
  return V32{reinterpret_cast<const float2*>(v), v + 2u * n, v + 3u * n, v + 4u * n};
}

__host__ __device__ __forceinline__ unsigned int qix32(const Geo& g, int i, int j, int k, bool is3d) {
  return (unsigned int)(i - 1) + (unsigned int)g.nx * ((unsigned int)(j - 1) +
                                                       (is3d ? (unsigned int)g.ny * (unsigned int)(k - 1) : 0u));
}

template <bool IS3D>
__host__ __device__ __forceinline__ void cell_pair32(const Geo& g, const Bnd& b, const uint8_t* flags,
                                                     const V32& X, const V32& P, int i, int j, int k,
                                                     Cell<float>& cx, Cell<float>& cp) {
  const bool out = i < 0 || i > g.Nx || j < 0 || j > g.Ny || (IS3D && (k < 0 || k > g.Nz));
  const bool fi = (i == 0) || (i == g.Nx);
  const bool fj = (j == 0) || (j == g.Ny);
  const bool fk = IS3D && ((k == 0) || (k == g.Nz));
  const bool ok = !out && ((int)fi + (int)fj + (int)fk) <= 1;
  // source node
  int si = fi ? (g.per_x ? (i == 0 ? g.Nx - 1 : 1) : (i == 0 ? 1 : g.Nx - 1)) : i;
  int sj = fj ? (g.per_y ? (j == 0 ? g.Ny - 1 : 1) : (j == 0 ? 1 : g.Ny - 1)) : j;
  int sk = IS3D ? (fk ? (g.per_z ? (k == 0 ? g.Nz - 1 : 1) : (k == 0 ? 1 : g.Nz - 1)) : k) : 1;
  if (!ok) si = sj = sk = 1;  // any node in range; the cell is zeroed below
  const unsigned int q = qix32(g, si, sj, sk, IS3D);
  const float2 xa = ldu(X.psi, q), pa = ldu(P.psi, q);
  float xx = ldu(X.px, q), xy = ldu(X.py, q), xz = IS3D ? ldu(X.pz, q) : 0.f;
  float px = ldu(P.px, q), py = ldu(P.py, q), pz = IS3D ? ldu(P.pz, q) : 0.f;
  // normal zeroing at the source node
  if (si == g.Nx - 1) xx = px = 0.f;
  if (sj == g.Ny - 1) xy = py = 0.f;
  if (IS3D && sk == g.Nz - 1) xz = pz = 0.f;
  const long long m = ok ? mfull(g, i, j, k) : 0;
  // field flux of the ghost rules (ghost cells only)
  float fz = 0.f, fy = 0.f, fx = 0.f;
  if (ok && (fi || fj || fk)) {
    fz = fld_cz<float>(g, b, m);
    if (IS3D) {
      fy = fld_cy<float>(g, b, m);
      fx = fld_cx<float>(g, b, m);
    }
  }
  float gx = xx, gy = xy, gz = xz, hx = px, hy = py, hz = pz;
  if (fi) {
    gx = g.per_x ? xx : 0.f;
    hx = g.per_x ? px : 0.f;
    gy = g.per_x ? 0.f : xy + (i == 0 ? -fz : fz);
    hy = g.per_x ? 0.f : py;
    gz = g.per_x ? 0.f : xz + (i == 0 ? fy : -fy);
    hz = g.per_x ? 0.f : pz;
  } else if (fj) {
    gx = g.per_y ? 0.f : xx + (j == 0 ? fz : -fz);
    hx = g.per_y ? 0.f : px;
    gy = g.per_y ? xy : 0.f;
    hy = g.per_y ? py : 0.f;
    gz = g.per_y ? 0.f : xz + (j == 0 ? -fx : fx);
    hz = g.per_y ? 0.f : pz;
  } else if (fk) {
    gx = g.per_z ? 0.f : xx + (k == 0 ? -fy : fy);
    hx = g.per_z ? 0.f : px;
    gy = g.per_z ? 0.f : xy + (k == 0 ? fx : -fx);
    hy = g.per_z ? 0.f : py;
    gz = g.per_z ? xz : 0.f;
    hz = g.per_z ? pz : 0.f;
  }
  if (ok && flags) {  // hole-boundary links at the cell, applied last
    const uint8_t f = flags[m];
    if (f & kHoleX) gx = hx = 0.f;
    if (f & kHoleY) gy = hy = 0.f;
    if (IS3D && (f & kHoleZ)) gz = hz = 0.f;
  }
  if (!ok) {
    cx = Cell<float>{0.f, 0.f, 0.f, 0.f, 0.f};
    cp = cx;
    return;
  }
  cx = Cell<float>{xa.x, xa.y, gx, gy, IS3D ? gz : 0.f};
  cp = Cell<float>{pa.x, pa.y, hx, hy, IS3D ? hz : 0.f};
}

}  // namespace tdgl
