// Staging of one grid cell for the FP32 stencil sweeps: the values tdgl3d's
// eval_f sees at a full-grid node after _apply_boundary_conditions
// (src/tdgl3d/physics/rhs.py:54-139), for a state x (with the applied field)
// and a perturbation p (the same affine rules without the field), as
// eff_cell_pair computes them, split into a per-column plan (computed once per
// thread) and a per-plane step, on 32-bit offsets.
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

// One FP32 state-shaped vector of n nodes: a base pointer and the byte size of
// one phi plane (4 n); psi of node q is at byte 8 q, phi_x at 2 (4n) + 4 q,
// phi_y at 3 (4n) + 4 q, phi_z at 4 (4n) + 4 q.  One 64-bit base per vector
// keeps the kernels' uniform state small; the host keeps 5 (4n) below 2^32.
struct V32 {
  const char* base;
  unsigned int pl;  // 4 n
  __host__ __device__ __forceinline__ float2 psi(unsigned int q) const {
    return *reinterpret_cast<const float2*>(base + 8u * q);
  }
  __host__ __device__ __forceinline__ float ph(int a, unsigned int q) const {  // a = 2, 3, 4
    return *reinterpret_cast<const float*>(base + (unsigned int)a * pl + 4u * q);
  }
};
__host__ __device__ __forceinline__ V32 planes32(const float* v, unsigned int n) {
  return V32{reinterpret_cast<const char*>(v), 4u * n};
}
// The same for an output vector.
struct W32 {
  char* base;
  unsigned int pl;
  __host__ __device__ __forceinline__ void psi(unsigned int q, float2 v) const {
    *reinterpret_cast<float2*>(base + 8u * q) = v;
  }
  __host__ __device__ __forceinline__ void ph(int a, unsigned int q, float v) const {
    *reinterpret_cast<float*>(base + (unsigned int)a * pl + 4u * q) = v;
  }
};
__host__ __device__ __forceinline__ W32 wplanes32(float* v, unsigned int n) {
  return W32{reinterpret_cast<char*>(v), 4u * n};
}

__host__ __device__ __forceinline__ unsigned int qix32(const Geo& g, int i, int j, int k, bool is3d) {
  return (unsigned int)(i - 1) + (unsigned int)g.nx * ((unsigned int)(j - 1) +
                                                       (is3d ? (unsigned int)g.ny * (unsigned int)(k - 1) : 0u));
}

// Field terms of the ghost rules in FP32: uniform values converted once, or the
// general BoundaryVectors (rhs.py:32-40), read through `gen` per ghost cell.
struct Fld32 {
  float cx, cy, cz;  // (float)b.cx ... (uniform field)
  const Bnd* gen;    // the film's Bnd when it holds general field vectors, else nullptr
};
__host__ __device__ __forceinline__ Fld32 fld32(const Bnd& b, const Bnd* where) {
  return Fld32{(float)b.cx, (float)b.cy, (float)b.cz, b.bz ? where : nullptr};
}

// The part of a cell's staging that depends only on its column (i, j): a
// thread stages the same columns on every plane of its z march.
struct Col32 {
  unsigned int qij;  // (si - 1) + nx (sj - 1): source column in the interior numbering
  unsigned int mij;  // i + mj j: the cell's column in the full-grid numbering
  bool okij;         // inside the grid and not an xy edge or corner
  bool fi, fj;       // on an x / y face
  bool lo_i, lo_j;   // on the low x / y face (i == 0, j == 0)
  bool zx, zy;       // the source's phi_x / phi_y is zeroed (normal planes)
};

__host__ __device__ __forceinline__ Col32 col32(const Geo& g, int i, int j) {
  Col32 c;
  c.fi = (i == 0) || (i == g.Nx);
  c.fj = (j == 0) || (j == g.Ny);
  c.lo_i = i == 0;
  c.lo_j = j == 0;
  c.okij = i >= 0 && i <= g.Nx && j >= 0 && j <= g.Ny && !(c.fi && c.fj);
  int si = c.fi ? (g.per_x ? (i == 0 ? g.Nx - 1 : 1) : (i == 0 ? 1 : g.Nx - 1)) : i;
  int sj = c.fj ? (g.per_y ? (j == 0 ? g.Ny - 1 : 1) : (j == 0 ? 1 : g.Ny - 1)) : j;
  if (!c.okij) si = sj = 1;  // any node in range; the cell is zeroed
  c.qij = (unsigned int)(si - 1) + (unsigned int)g.nx * (unsigned int)(sj - 1);
  c.mij = c.okij ? (unsigned int)i + (unsigned int)g.mj * (unsigned int)j : 0u;
  c.zx = si == g.Nx - 1;
  c.zy = sj == g.Ny - 1;
  return c;
}

// Where the cell of column c on plane k takes its values from.
struct Src32 {
  unsigned int q;  // source node (interior numbering; any valid node when !ok)
  unsigned int m;  // the cell in the full-grid numbering (0 when !ok)
  bool ok;         // inside the grid and not an edge or corner
  bool fk;         // on a z face
  bool zz;         // the source's phi_z is zeroed (normal plane)
};

template <bool IS3D>
__host__ __device__ __forceinline__ Src32 src32(const Geo& g, const Col32& c, int k) {
  Src32 r;
  r.fk = IS3D && ((k == 0) || (k == g.Nz));
  r.ok = c.okij && (!IS3D || (k >= 0 && k <= g.Nz)) && !(r.fk && (c.fi || c.fj));
  int sk = IS3D ? (r.fk ? (g.per_z ? (k == 0 ? g.Nz - 1 : 1) : (k == 0 ? 1 : g.Nz - 1)) : k) : 1;
  if (!r.ok) sk = 1;
  r.q = c.qij + (IS3D ? (unsigned int)g.nx * (unsigned int)g.ny * (unsigned int)(sk - 1) : 0u);
  r.m = r.ok ? c.mij + (IS3D ? (unsigned int)g.mk * (unsigned int)k : 0u) : 0u;
  r.zz = IS3D && sk == g.Nz - 1;
  return r;
}

// The perturbation's link values of a cell from its source node's raw values:
// normal zeroing at the source, the face rules without field (components the
// ghost rules do not assign are 0; a periodic face copies its normal link),
// and the cell's hole bits.  Valid for ok cells.
template <bool IS3D>
__host__ __device__ __forceinline__ void fix_p32(const Geo& g, const Col32& c, const Src32& r,
                                                 uint8_t hole, float& px, float& py, float& pz) {
  if (c.zx) px = 0.f;
  if (c.zy) py = 0.f;
  if (r.zz) pz = 0.f;
  if (c.fi) {
    px = g.per_x ? px : 0.f;
    py = g.per_x ? 0.f : py;
    pz = g.per_x ? 0.f : pz;
  } else if (c.fj) {
    px = g.per_y ? 0.f : px;
    py = g.per_y ? py : 0.f;
    pz = g.per_y ? 0.f : pz;
  } else if (r.fk) {
    px = g.per_z ? 0.f : px;
    py = g.per_z ? 0.f : py;
    pz = g.per_z ? pz : 0.f;
  }
  if (hole & kHoleX) px = 0.f;
  if (hole & kHoleY) py = 0.f;
  if (IS3D && (hole & kHoleZ)) pz = 0.f;
  if (!IS3D) pz = 0.f;
}

// Stage the cell of column c on plane k (k = 0 in 2-D).
template <bool IS3D>
__host__ __device__ __forceinline__ void stage32(const Geo& g, const Fld32& F, const uint8_t* flags,
                                                 const V32& X, const V32& P, const Col32& c, int k,
                                                 Cell<float>& cx, Cell<float>& cp) {
  const Src32 r = src32<IS3D>(g, c, k);
  const unsigned int q = r.q, m = r.m;
  const float2 xa = X.psi(q), pa = P.psi(q);
  float xx = X.ph(2, q), xy = X.ph(3, q), xz = IS3D ? X.ph(4, q) : 0.f;
  float px = P.ph(2, q), py = P.ph(3, q), pz = IS3D ? P.ph(4, q) : 0.f;
  uint8_t hole = 0;
  if (r.ok && flags) hole = ldu(flags, m);
  fix_p32<IS3D>(g, c, r, hole, px, py, pz);
  // the state's links: the same rules plus the field flux of the ghost rules (rhs.py:95-127)
  if (c.zx) xx = 0.f;
  if (c.zy) xy = 0.f;
  if (r.zz) xz = 0.f;
  float fz = F.cz, fy = F.cy, fx = F.cx;
  if (F.gen && r.ok && (c.fi || c.fj || r.fk)) {
    const Bnd& B = *F.gen;
    fz = (float)((B.bz[m] * B.hx) * B.hy);
    fy = (float)((B.by[m] * B.hz) * B.hx);
    fx = (float)((B.bx[m] * B.hy) * B.hz);
  }
  float gx = xx, gy = xy, gz = xz;
  if (c.fi) {
    gx = g.per_x ? xx : 0.f;
    gy = g.per_x ? 0.f : xy + (c.lo_i ? -fz : fz);
    gz = g.per_x ? 0.f : xz + (c.lo_i ? fy : -fy);
  } else if (c.fj) {
    gx = g.per_y ? 0.f : xx + (c.lo_j ? fz : -fz);
    gy = g.per_y ? xy : 0.f;
    gz = g.per_y ? 0.f : xz + (c.lo_j ? -fx : fx);
  } else if (r.fk) {
    gx = g.per_z ? 0.f : xx + (k == 0 ? -fy : fy);
    gy = g.per_z ? 0.f : xy + (k == 0 ? fx : -fx);
    gz = g.per_z ? xz : 0.f;
  }
  if (hole & kHoleX) gx = 0.f;
  if (hole & kHoleY) gy = 0.f;
  if (IS3D && (hole & kHoleZ)) gz = 0.f;
  if (!r.ok) {
    cx = Cell<float>{0.f, 0.f, 0.f, 0.f, 0.f};
    cp = cx;
    return;
  }
  cx = Cell<float>{xa.x, xa.y, gx, gy, IS3D ? gz : 0.f};
  cp = Cell<float>{pa.x, pa.y, px, py, IS3D ? pz : 0.f};
}

// One cell, column plan included (the CPU check compares this with
// eff_cell_pair; kernels build the plan once per column).
template <bool IS3D>
__host__ __device__ __forceinline__ void cell_pair32(const Geo& g, const Bnd& b, const uint8_t* flags,
                                                     const V32& X, const V32& P, int i, int j, int k,
                                                     Cell<float>& cx, Cell<float>& cp) {
  stage32<IS3D>(g, fld32(b, &b), flags, X, P, col32(g, i, j), k, cx, cp);
}

}  // namespace tdgl
