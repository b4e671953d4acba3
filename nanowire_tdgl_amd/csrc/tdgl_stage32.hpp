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

// Device code accesses the vectors through global-address-space pointers, so
// that loads and stores take the global forms (a flat access, which a pointer
// read from memory otherwise gets, also waits with the LDS traffic).  The CPU
// build of these helpers (the staging check) uses plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define TDGL_GLOBAL __attribute__((address_space(1)))
#else
#define TDGL_GLOBAL
#endif

// A T at byte offset `off` from a uniform base.  The offset is one 32-bit sum
// (the global_load saddr form: a 32-bit offset per lane, no 64-bit address
// arithmetic, and no per-plane 64-bit bases hoisted into scalar registers).
template <typename T>
__host__ __device__ __forceinline__ T ldb(const void* base, unsigned int off) {
  return *(const TDGL_GLOBAL T*)((const TDGL_GLOBAL char*)base + off);
}
template <typename T>
__host__ __device__ __forceinline__ void stb(void* base, unsigned int off, T v) {
  *(TDGL_GLOBAL T*)((TDGL_GLOBAL char*)base + off) = v;
}
// Element e of an array (callers keep arrays below 2^32 bytes).
template <typename T>
__host__ __device__ __forceinline__ T ldu(const T* base, unsigned int e) {
  return ldb<T>(base, e * (unsigned int)sizeof(T));
}
template <typename T>
__host__ __device__ __forceinline__ void stu(T* base, unsigned int e, T v) {
  stb<T>(base, e * (unsigned int)sizeof(T), v);
}

// One FP32 state-shaped vector of n nodes, the mixed method's layout: psi of
// node q at byte 8 q (the first 2 n floats, as in the internal layout), then
// each node's link components together: phi_x, phi_y (, phi_z) of node q at
// byte 8 n + ls q, ls = 4 nphi.  A node's links are then one load (dwordx3,
// dwordx2 in 2-D) where three planes took three.  One 64-bit base per vector
// keeps the kernels' uniform state small; the host keeps 5 (4n) below 2^32.
struct V32 {
  const char* base;
  unsigned int lk;  // 8 n: where the links start
  unsigned int ls;  // 4 nphi: bytes of one node's links
  __host__ __device__ __forceinline__ float2 psi(unsigned int q) const { return ldb<float2>(base, 8u * q); }
  __host__ __device__ __forceinline__ float ph(int a, unsigned int q) const {  // a = 2, 3, 4
    return ldb<float>(base, lk + ls * q + 4u * (unsigned int)(a - 2));
  }
  // phi_x, phi_y, phi_z of node q (phi_z = 0 in 2-D)
  template <bool IS3D>
  __host__ __device__ __forceinline__ float3 links(unsigned int q) const {
    if (IS3D) return ldb<float3>(base, lk + 12u * q);
    const float2 v = ldb<float2>(base, lk + 8u * q);
    return make_float3(v.x, v.y, 0.f);
  }
};
__host__ __device__ __forceinline__ V32 planes32(const float* v, unsigned int n, bool is3d) {
  return V32{reinterpret_cast<const char*>(v), 8u * n, is3d ? 12u : 8u};
}
// The same for an output vector.
struct W32 {
  char* base;
  unsigned int lk, ls;
  __host__ __device__ __forceinline__ void psi(unsigned int q, float2 v) const { stb<float2>(base, 8u * q, v); }
  __host__ __device__ __forceinline__ void ph(int a, unsigned int q, float v) const {
    stb<float>(base, lk + ls * q + 4u * (unsigned int)(a - 2), v);
  }
  template <bool IS3D>
  __host__ __device__ __forceinline__ void links(unsigned int q, float x, float y, float z) const {
    if (IS3D) stb<float3>(base, lk + 12u * q, make_float3(x, y, z));
    else stb<float2>(base, lk + 8u * q, make_float2(x, y));
  }
};
__host__ __device__ __forceinline__ W32 wplanes32(float* v, unsigned int n, bool is3d) {
  return W32{reinterpret_cast<char*>(v), 8u * n, is3d ? 12u : 8u};
}
// Element index in the mixed method's FP32 layout of element e of the internal
// layout (psi in [0, 2n) as is; phi_a of node q, internal (2 + a) n + q, at
// 2n + nphi q + a) -- the maps the FP32 <-> FP64 kernels use.
__host__ __device__ __forceinline__ long long l32_of_internal(long long e, long long n, int nphi) {
  if (e < 2 * n || e >= (2 + nphi) * n) return e;
  const long long a = (e - 2 * n) / n, q = (e - 2 * n) - a * n;
  return 2 * n + nphi * q + a;
}
__host__ __device__ __forceinline__ long long internal_of_l32(long long e, long long n, int nphi) {
  if (e < 2 * n || e >= (2 + nphi) * n) return e;
  const long long q = (e - 2 * n) / nphi, a = (e - 2 * n) - q * nphi;
  return 2 * n + a * n + q;
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
    fz = (float)((ldu(B.bz, m) * B.hx) * B.hy);
    fy = (float)((ldu(B.by, m) * B.hz) * B.hx);
    fx = (float)((ldu(B.bx, m) * B.hy) * B.hz);
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

// ---------------------------------------------------------------------------
// The same staging for a uniform applied field, as masks.  Every rule above
// either keeps a value, zeroes it, or adds a field flux, so a cell's link
// component is g = (((x & s) + a) & k) bit for bit: s the source's normal
// zeroing, a the flux (-0.f when none: x + -0.f == x for every x), k the face
// rule, the hole bits and !ok.  The column's part of s, a and k is computed
// once per column as 32-bit words and floats: lane values that stay in vector
// registers across the z march, where the booleans of Col32 would be held as
// 64-bit lane masks in scalar registers.  Planes on a z face replace the
// column's face rule by the z one (only cells on no x / y face stay).
// ---------------------------------------------------------------------------

__host__ __device__ __forceinline__ float fmask(float x, unsigned int m) {
  return __builtin_bit_cast(float, __builtin_bit_cast(unsigned int, x) & m);
}
__host__ __device__ __forceinline__ unsigned int keep(bool b) { return b ? ~0u : 0u; }
// Hide a lane value's origin from the optimiser, which would otherwise turn
// x & keep(b) back into a select on b and hold b as a lane mask.
__host__ __device__ __forceinline__ void opaque(unsigned int& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(v));
#endif
}

struct Plan32 {
  unsigned int qij, mij;    // as in Col32
  unsigned int ok;          // keep(okij): psi is kept
  unsigned int sx, sy;      // keep(!zx), keep(!zy)
  unsigned int kx, ky, kz;  // the x / y face rule and ok, per component
  float ax, ay, az;         // flux the x / y face rule adds to the state's link (-0.f: none)
};

__host__ __device__ __forceinline__ Plan32 plan32(const Geo& g, const Fld32& F, int i, int j) {
  const Col32 c = col32(g, i, j);
  Plan32 p;
  p.qij = c.qij;
  p.mij = c.mij;
  p.ok = keep(c.okij);
  p.sx = keep(!c.zx);
  p.sy = keep(!c.zy);
  bool kx = true, ky = true, kz = true;
  p.ax = p.ay = p.az = -0.f;
  if (c.fi) {
    kx = g.per_x;
    ky = kz = !g.per_x;
    if (!g.per_x) {
      p.ay = c.lo_i ? -F.cz : F.cz;
      p.az = c.lo_i ? F.cy : -F.cy;
    }
  } else if (c.fj) {
    ky = g.per_y;
    kx = kz = !g.per_y;
    if (!g.per_y) {
      p.ax = c.lo_j ? F.cz : -F.cz;
      p.az = c.lo_j ? -F.cx : F.cx;
    }
  }
  p.kx = keep(c.okij && kx);
  p.ky = keep(c.okij && ky);
  p.kz = keep(c.okij && kz);
  opaque(p.ok);
  opaque(p.sx);
  opaque(p.sy);
  opaque(p.kx);
  opaque(p.ky);
  opaque(p.kz);
  return p;
}

// The plane's part (uniform over a block): source plane, z-normal zeroing, and
// on a z face its rule (rhs.py:73-76, 95-127).  Planes 0 <= k <= Nz (3-D; the
// sweeps stage no others), k = 0 in 2-D.
struct Pl32 {
  unsigned int qoff, moff;  // source plane offset (interior numbering), cell plane offset (full grid)
  unsigned int sz;          // keep(!zz)
  bool fk;                  // a z face
  unsigned int kx, ky, kz;  // the z face rule per component (fk only)
  float ax, ay;             // its flux (fk only)
};

template <bool IS3D>
__host__ __device__ __forceinline__ Pl32 plane32(const Geo& g, const Fld32& F, int k) {
  Pl32 p;
  p.fk = IS3D && ((k == 0) || (k == g.Nz));
  const int sk = IS3D ? (p.fk ? (g.per_z ? (k == 0 ? g.Nz - 1 : 1) : (k == 0 ? 1 : g.Nz - 1)) : k) : 1;
  p.qoff = IS3D ? (unsigned int)g.nx * (unsigned int)g.ny * (unsigned int)(sk - 1) : 0u;
  p.moff = IS3D ? (unsigned int)g.mk * (unsigned int)k : 0u;
  p.sz = keep(!(IS3D && sk == g.Nz - 1));
  p.kx = p.ky = keep(!g.per_z);
  p.kz = keep(g.per_z);
  p.ax = g.per_z ? -0.f : (k == 0 ? -F.cy : F.cy);
  p.ay = g.per_z ? -0.f : (k == 0 ? F.cx : -F.cx);
  return p;
}

// One staged cell's loads, before any rule: the state and the perturbation at
// the source node and the cell's hole bits (0 without flags).  A sweep issues
// them a plane ahead and applies the rules (finish32u) when it writes the cell
// to shared memory, so the loads' latency overlaps a plane of arithmetic.
struct Raw32 {
  float2 xa, pa;
  float xx, xy, xz, px, py, pz;
  unsigned int h;
};

template <bool IS3D>
__host__ __device__ __forceinline__ void load32u(const uint8_t* flags, const V32& X, const V32& P,
                                                 const Plan32& c, const Pl32& p, Raw32& r) {
  const unsigned int q = c.qij + p.qoff;
  r.xa = X.psi(q);
  r.pa = P.psi(q);
  const float3 xl = X.links<IS3D>(q), pk = P.links<IS3D>(q);
  r.xx = xl.x;
  r.xy = xl.y;
  r.xz = xl.z;
  r.px = pk.x;
  r.py = pk.y;
  r.pz = pk.z;
  r.h = flags ? (unsigned int)ldu(flags, c.mij + p.moff) : 0u;
}

// The rules of column plan c on plane p applied to a cell's loads (uniform
// field).  FK: p is a z face (the caller branches on p.fk, which is uniform).
// HF: the film has hole flags (r.h was loaded).
template <bool IS3D, bool FK>
__host__ __device__ __forceinline__ void finish32u(const Raw32& r, bool hf, const Plan32& c,
                                                   const Pl32& p, Cell<float>& cx, Cell<float>& cp) {
  unsigned int kx, ky, kz, ok;
  float ax, ay, az;
  if (FK) {  // only cells on no x / y face stay: their x / y rule keeps every component
    ok = c.kx & c.ky & c.kz;
    kx = ok & p.kx;
    ky = ok & p.ky;
    kz = ok & p.kz;
    ax = p.ax;
    ay = p.ay;
    az = -0.f;
  } else {
    ok = c.ok;
    kx = c.kx;
    ky = c.ky;
    kz = c.kz;
    ax = c.ax;
    ay = c.ay;
    az = c.az;
  }
  if (hf) {  // keep(!(h & bit)) as (bit of h) - 1: all ones when clear, zero when set
    unsigned int hx = (r.h & 1u) - 1u, hy = ((r.h >> 1) & 1u) - 1u, hz = ((r.h >> 2) & 1u) - 1u;
    static_assert(kHoleX == 1 && kHoleY == 2 && kHoleZ == 4, "hole bits");
    opaque(hx);
    opaque(hy);
    opaque(hz);
    kx &= hx;
    ky &= hy;
    if (IS3D) kz &= hz;
  }
  cx.pr = fmask(r.xa.x, ok);
  cx.pi = fmask(r.xa.y, ok);
  cx.x = fmask(fmask(r.xx, c.sx) + ax, kx);
  cx.y = fmask(fmask(r.xy, c.sy) + ay, ky);
  cx.z = IS3D ? fmask(fmask(r.xz, p.sz) + az, kz) : 0.f;
  cp.pr = fmask(r.pa.x, ok);
  cp.pi = fmask(r.pa.y, ok);
  cp.x = fmask(r.px, c.sx & kx);
  cp.y = fmask(r.py, c.sy & ky);
  cp.z = IS3D ? fmask(r.pz, p.sz & kz) : 0.f;
}

// Stage the cell of column plan c on plane p (uniform field): loads and rules
// together.  FK: p is a z face.
template <bool IS3D, bool FK>
__host__ __device__ __forceinline__ void stage32u(const uint8_t* flags, const V32& X, const V32& P,
                                                  const Plan32& c, const Pl32& p, Cell<float>& cx,
                                                  Cell<float>& cp) {
  Raw32 r;
  load32u<IS3D>(flags, X, P, c, p, r);
  finish32u<IS3D, FK>(r, flags != nullptr, c, p, cx, cp);
}

// ---------------------------------------------------------------------------
// FP64 (the plain path's stencils): one state -- or x + eps p, the FD JVP's
// argument (tgcr.py:130) -- staged with the same masks as eff_cell's rules
// give bit for bit; a mask applies to both words of a double.
// ---------------------------------------------------------------------------

__host__ __device__ __forceinline__ double fmask(double x, unsigned int m) {
  const unsigned long long mm = ((unsigned long long)m << 32) | m;
  return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) & mm);
}

// One FP64 state-shaped vector of n nodes on 32-bit byte offsets: psi of node q
// at 16 q, phi_x at 2 (8n) + 8 q, ...; the host keeps 5 (8n) below 2^32.
struct V64 {
  const char* base;
  unsigned int pl;  // 8 n
  __host__ __device__ __forceinline__ double2 psi(unsigned int q) const { return ldb<double2>(base, 16u * q); }
  __host__ __device__ __forceinline__ double ph(int a, unsigned int q) const {
    return ldb<double>(base, (unsigned int)a * pl + 8u * q);
  }
};
__host__ __device__ __forceinline__ V64 planes64(const double* v, unsigned int n) {
  return V64{reinterpret_cast<const char*>(v), 8u * n};
}

struct Plan64 {
  unsigned int qij, mij, ok, sx, sy, kx, ky, kz;  // as in Plan32
  double ax, ay, az;
};
__host__ __device__ __forceinline__ Plan64 plan64(const Geo& g, const Bnd& b, int i, int j) {
  const Col32 c = col32(g, i, j);
  Plan64 p;
  p.qij = c.qij;
  p.mij = c.mij;
  p.ok = keep(c.okij);
  p.sx = keep(!c.zx);
  p.sy = keep(!c.zy);
  bool kx = true, ky = true, kz = true;
  p.ax = p.ay = p.az = -0.0;
  if (c.fi) {
    kx = g.per_x;
    ky = kz = !g.per_x;
    if (!g.per_x) {  // rhs.py:91-98
      p.ay = c.lo_i ? -b.cz : b.cz;
      p.az = c.lo_i ? b.cy : -b.cy;
    }
  } else if (c.fj) {
    ky = g.per_y;
    kx = kz = !g.per_y;
    if (!g.per_y) {  // rhs.py:107-112
      p.ax = c.lo_j ? b.cz : -b.cz;
      p.az = c.lo_j ? -b.cx : b.cx;
    }
  }
  p.kx = keep(c.okij && kx);
  p.ky = keep(c.okij && ky);
  p.kz = keep(c.okij && kz);
  opaque(p.ok);
  opaque(p.sx);
  opaque(p.sy);
  opaque(p.kx);
  opaque(p.ky);
  opaque(p.kz);
  return p;
}

struct Pl64 {
  unsigned int qoff, moff, sz;
  bool fk;
  unsigned int kx, ky, kz;
  double ax, ay;
};
template <bool IS3D>
__host__ __device__ __forceinline__ Pl64 plane64(const Geo& g, const Bnd& b, int k) {
  Pl64 p;
  p.fk = IS3D && ((k == 0) || (k == g.Nz));
  const int sk = IS3D ? (p.fk ? (g.per_z ? (k == 0 ? g.Nz - 1 : 1) : (k == 0 ? 1 : g.Nz - 1)) : k) : 1;
  p.qoff = IS3D ? (unsigned int)g.nx * (unsigned int)g.ny * (unsigned int)(sk - 1) : 0u;
  p.moff = IS3D ? (unsigned int)g.mk * (unsigned int)k : 0u;
  p.sz = keep(!(IS3D && sk == g.Nz - 1));
  p.kx = p.ky = keep(!g.per_z);
  p.kz = keep(g.per_z);
  p.ax = g.per_z ? -0.0 : (k == 0 ? -b.cy : b.cy);  // rhs.py:122-127
  p.ay = g.per_z ? -0.0 : (k == 0 ? b.cx : -b.cx);
  return p;
}

// The raw loads of a staged FP64 cell (x, p and the hole flags at the cell), and
// the column's and plane's rules applied to them (stage64u below is the two).
struct Raw64 {
  double2 xa, pa;
  double xx, xy, xz, px, py, pz;
  unsigned int h;
};
template <bool IS3D, bool PERT>
__host__ __device__ __forceinline__ void load64u(const uint8_t* flags, const V64& X, const V64& P,
                                                 const Plan64& c, const Pl64& p, Raw64& r) {
  const unsigned int q = c.qij + p.qoff;
  r.xa = X.psi(q);
  r.xx = X.ph(2, q);
  r.xy = X.ph(3, q);
  r.xz = IS3D ? X.ph(4, q) : 0.0;
  if (PERT) {
    r.pa = P.psi(q);
    r.px = P.ph(2, q);
    r.py = P.ph(3, q);
    if (IS3D) r.pz = P.ph(4, q);
  }
  r.h = flags ? (unsigned int)ldu(flags, c.mij + p.moff) : 0u;
}
template <bool IS3D, bool FK, bool PERT>
__host__ __device__ __forceinline__ void finish64u(const Raw64& r, bool hf, double eps, const Plan64& c,
                                                   const Pl64& p, Cell<double>& cx) {
  unsigned int kx, ky, kz, ok;
  double ax, ay, az;
  if (FK) {
    ok = c.kx & c.ky & c.kz;
    kx = ok & p.kx;
    ky = ok & p.ky;
    kz = ok & p.kz;
    ax = p.ax;
    ay = p.ay;
    az = -0.0;
  } else {
    ok = c.ok;
    kx = c.kx;
    ky = c.ky;
    kz = c.kz;
    ax = c.ax;
    ay = c.ay;
    az = c.az;
  }
  double2 xa = r.xa;
  double xx = r.xx, xy = r.xy, xz = r.xz;
  if (PERT) {
    xa.x = xa.x + eps * r.pa.x;
    xa.y = xa.y + eps * r.pa.y;
    xx = xx + eps * r.px;
    xy = xy + eps * r.py;
    if (IS3D) xz = xz + eps * r.pz;
  }
  if (hf) {
    kx &= keep(!(r.h & kHoleX));
    ky &= keep(!(r.h & kHoleY));
    if (IS3D) kz &= keep(!(r.h & kHoleZ));
  }
  cx.pr = fmask(xa.x, ok);
  cx.pi = fmask(xa.y, ok);
  cx.x = fmask(fmask(xx, c.sx) + ax, kx);
  cx.y = fmask(fmask(xy, c.sy) + ay, ky);
  cx.z = IS3D ? fmask(fmask(xz, p.sz) + az, kz) : 0.0;
}

// Stage the cell of column plan c on plane p (0 <= k <= Nz; uniform field):
// x, or x + eps p when PERT (two roundings, as Src::ld).
template <bool IS3D, bool FK, bool PERT>
__host__ __device__ __forceinline__ void stage64u(const uint8_t* flags, const V64& X, const V64& P,
                                                  double eps, const Plan64& c, const Pl64& p,
                                                  Cell<double>& cx) {
  Raw64 r;
  load64u<IS3D, PERT>(flags, X, P, c, p, r);
  finish64u<IS3D, FK, PERT>(r, flags != nullptr, eps, c, p, cx);
}

// FP64 cell with its plan (the CPU check compares this with eff_cell).
template <bool IS3D, bool PERT>
__host__ __device__ __forceinline__ Cell<double> cell64(const Geo& g, const Bnd& b, const uint8_t* flags,
                                                        const V64& X, const V64& P, double eps, int i,
                                                        int j, int k) {
  const Plan64 c = plan64(g, b, i, j);
  const Pl64 p = plane64<IS3D>(g, b, k);
  Cell<double> r;
  if (p.fk)
    stage64u<IS3D, true, PERT>(flags, X, P, eps, c, p, r);
  else
    stage64u<IS3D, false, PERT>(flags, X, P, eps, c, p, r);
  return r;
}

// One cell, column plan included (the CPU check compares this with
// eff_cell_pair; kernels build the plan once per column).  `masks` picks the
// uniform-field form (stage32u), which needs a uniform field.
template <bool IS3D>
__host__ __device__ __forceinline__ void cell_pair32(const Geo& g, const Bnd& b, const uint8_t* flags,
                                                     const V32& X, const V32& P, int i, int j, int k,
                                                     Cell<float>& cx, Cell<float>& cp, bool masks = false) {
  const Fld32 F = fld32(b, &b);
  if (masks && !F.gen) {
    const Plan32 c = plan32(g, F, i, j);
    const Pl32 p = plane32<IS3D>(g, F, k);
    if (p.fk)
      stage32u<IS3D, true>(flags, X, P, c, p, cx, cp);
    else
      stage32u<IS3D, false>(flags, X, P, c, p, cx, cp);
    return;
  }
  stage32<IS3D>(g, F, flags, X, P, col32(g, i, j), k, cx, cp);
}

}  // namespace tdgl
