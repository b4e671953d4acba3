// C ABI of libtdglhip.so and the host driver of the trapezoidal step.
//
// The driver keeps tdgl3d's control flow exactly (integrators.py:151-172,
// newton.py:97-120, tgcr.py:106-156) and reads scalars back only where the
// reference branches: once per TGCR iteration (stop test + breakdown test)
// and once per Newton iteration (convergence test).

#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/tdgl_hip.h"
#include "tdgl_ctx.hpp"
#include "tdgl_launch.hpp"

using namespace tdgl;

namespace tdgl {
KernelTiming& kernel_timing() {
  thread_local KernelTiming t;
  return t;
}
}  // namespace tdgl

namespace {

int fail(tdgl_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(ctx, TDGL_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define LAUNCHCHK() HIPCHK(hipGetLastError())

int alloc(tdgl_ctx* ctx, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(ctx, TDGL_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return TDGL_OK;
}

// Brackets of one kernel launch are timed on the kernel itself.
int prof_begin(tdgl_ctx* ctx, int cls, double bytes, int launches = 1) {
  return ctx->prof.begin(ctx->st, cls, bytes, launches, launches == 1);
}
void prof_end(tdgl_ctx* ctx, int tok) { ctx->prof.end(ctx->st, tok); }
void prof_harvest(tdgl_ctx* ctx) { ctx->prof.harvest(); }

// Reduction plumbing for the next launch: consume the last published partials,
// publish into the other buffer.
Red red_next(tdgl_ctx* ctx) {
  Red r;
  r.pin = ctx->d_pb[ctx->pb_cur];
  r.pcount = ctx->pb_count;
  r.pout = ctx->d_pb[ctx->pb_cur ^ 1];
  r.res = ctx->d_res;
  return r;
}
void red_commit(tdgl_ctx* ctx, int nblocks) {
  ctx->pb_cur ^= 1;
  ctx->pb_count = nblocks;
}
// Reduce the last published partials into res[slot0 ...] (one block).
void red_finalize(tdgl_ctx* ctx, int nsum, int nmax, int slot0) {
  launch_finalize(nsum, nmax, slot0, red_next(ctx), ctx->st);
}

// Scalars back to the host (synchronises the stream).
int readback(tdgl_ctx* ctx, int slot, int count) {
  HIPCHK(hipMemcpyAsync(ctx->h_res + slot, ctx->d_res + slot, sizeof(double) * count,
                        hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (!ctx->prof.pending.empty()) prof_harvest(ctx);
  return TDGL_OK;
}

int ensure_work(tdgl_ctx* ctx) {
  if (ctx->w_x) return TDGL_OK;
  const size_t bytes = sizeof(double) * (size_t)ctx->len;
  int rc;
  if ((rc = alloc(ctx, (void**)&ctx->w_x, bytes))) return rc;
  if ((rc = alloc(ctx, (void**)&ctx->w_gam, bytes))) return rc;
  if ((rc = alloc(ctx, (void**)&ctx->w_f, bytes))) return rc;
  if ((rc = alloc(ctx, (void**)&ctx->w_r, bytes))) return rc;
  return TDGL_OK;
}

// Krylov slot i (p_i and Ap_i) allocated on first use; the device table of
// p pointers is refreshed when it grows.
int ensure_vec(tdgl_ctx* ctx, int i) {
  if (i >= kMaxKrylov)
    return fail(ctx, TDGL_E_INVALID, "TGCR needs more than 2048 Krylov vectors");
  const size_t bytes = sizeof(double) * (size_t)ctx->len;
  while ((int)ctx->P.size() <= i) {
    double *p = nullptr, *ap = nullptr;
    int rc;
    if ((rc = alloc(ctx, (void**)&p, bytes))) return rc;
    if ((rc = alloc(ctx, (void**)&ap, bytes))) {
      hipFree(p);
      return rc;
    }
    ctx->P.push_back(p);
    ctx->AP.push_back(ap);
  }
  if ((int)ctx->P.size() > ctx->ptab_cap) {
    int cap = std::max(64, ctx->ptab_cap * 2);
    while (cap < (int)ctx->P.size()) cap *= 2;
    HIPCHK(hipStreamSynchronize(ctx->st));  // no kernel may still read the old table
    if (ctx->d_ptab) HIPCHK(hipFree(ctx->d_ptab));
    ctx->d_ptab = nullptr;
    int rc;
    if ((rc = alloc(ctx, (void**)&ctx->d_ptab, sizeof(double*) * cap))) return rc;
    ctx->ptab_cap = cap;
    ctx->ptab_len = 0;
  }
  if ((int)ctx->P.size() > ctx->ptab_len) {  // the table only grows: copy the new entries
    HIPCHK(hipMemcpyAsync(ctx->d_ptab + ctx->ptab_len, ctx->P.data() + ctx->ptab_len,
                          sizeof(double*) * (ctx->P.size() - ctx->ptab_len), hipMemcpyHostToDevice,
                          ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));  // the host vector may reallocate later
    ctx->ptab_len = (int)ctx->P.size();
  }
  return TDGL_OK;
}

void set_uniform_coeffs(tdgl_ctx* ctx, double kappa) {
  ctx->kappa_uniform = kappa;
  Mat& m = ctx->mt;
  const double k2 = kappa * kappa;  // sparse_operators.py:131 (kappa_m**2)
  ctx->k2max = k2;
  m.cx = k2 / m.hx2;
  m.cy = k2 / m.hy2;
  m.cz = ctx->g.is3d ? k2 / m.hz2 : 0.0;
  m.dgx = (-2.0 * (m.cx + m.cz)) + (-2.0 * (m.cx + m.cy));
  m.dgy = (-2.0 * (m.cy + m.cz)) + (-2.0 * (m.cx + m.cy));
  m.dgz = (-2.0 * (m.cy + m.cz)) + (-2.0 * (m.cx + m.cz));
}

double state_bytes(const tdgl_ctx* ctx) { return 8.0 * (double)ctx->len; }

}  // namespace

extern "C" {

int tdgl_ctx_create(const tdgl_grid* grid, tdgl_ctx** out) {
  tdgl_ctx* ctx = nullptr;
  if (!grid || !out) return TDGL_E_INVALID;
  *out = nullptr;
  if (grid->Nx < 2 || grid->Ny < 2 || grid->Nz < 1) return TDGL_E_INVALID;  // parameters.py:47-52
  if (!(grid->hx > 0) || !(grid->hy > 0) || !(grid->hz > 0)) return TDGL_E_INVALID;
  ctx = new tdgl_ctx();
  ctx->grid = *grid;
  HIPCHK(hipGetDevice(&ctx->device));
  {
    const int rc0 = noalloc_init_device(ctx->device, &ctx->err);
    if (rc0 != TDGL_OK) {
      delete ctx;
      return rc0;
    }
  }
  Geo& g = ctx->g;
  g.Nx = grid->Nx;
  g.Ny = grid->Ny;
  g.Nz = grid->Nz;
  g.is3d = grid->Nz > 1;
  g.nx = grid->Nx - 1;
  g.ny = grid->Ny - 1;
  g.nz = g.is3d ? grid->Nz - 1 : 1;
  g.per_x = grid->periodic_x != 0;
  g.per_y = grid->periodic_y != 0;
  g.per_z = grid->periodic_z != 0;
  g.mj = grid->Nx + 1;
  g.mk = (long long)(grid->Nx + 1) * (grid->Ny + 1);
  g.n = (long long)g.nx * g.ny * g.nz;
  g.nphi = g.is3d ? 3 : 2;
  ctx->len = (2 + g.nphi) * g.n;
  // node indices are 32-bit on the device, and the stencils and sweeps address
  // a state vector with 32-bit byte offsets (8 B per FP64 entry)
  if (g.mk * (long long)(grid->Nz + 1) >= (1LL << 31) || ctx->len >= (1LL << 31) ||
      8 * ctx->len >= (1LL << 32)) {
    delete ctx;
    return TDGL_E_INVALID;
  }

  Bnd& b = ctx->b;
  b.hx = grid->hx;
  b.hy = grid->hy;
  b.hz = grid->hz;
  Mat& m = ctx->mt;
  m.hx2 = grid->hx * grid->hx;
  m.hy2 = grid->hy * grid->hy;
  m.hz2 = grid->hz * grid->hz;
  m.ihx2 = 1.0 / m.hx2;
  m.ihy2 = 1.0 / m.hy2;
  m.ihz2 = 1.0 / m.hz2;
  m.dgpsi = (-2.0 * m.ihx2) + (-2.0 * m.ihy2);
  if (g.is3d) m.dgpsi = m.dgpsi + (-2.0 * m.ihz2);
  m.inv_tau = 1.0 / 0.1;
  set_uniform_coeffs(ctx, grid->kappa);

  // z planes per block: enough blocks to fill 256 CUs several times over
  if (g.is3d) {
    const long long tiles = (long long)((g.nx + 63) / 64) * ((g.ny + 3) / 4);
    long long kz = tiles * g.nz / 1024;
    ctx->cfg.kz = (int)std::max(1LL, std::min<long long>(kz, g.nz));
  }

  const int maxb = std::max({stencil_blocks(g, ctx->cfg), stream_blocks(g), mgs_blocks((2LL + g.nphi) * g.n)});
  ctx->pb_blocks = maxb;
  int rc;
  if ((rc = alloc(ctx, (void**)&ctx->d_pb[0], sizeof(double) * kPartialStride * (size_t)maxb))) goto bad;
  if ((rc = alloc(ctx, (void**)&ctx->d_pb[1], sizeof(double) * kPartialStride * (size_t)maxb))) goto bad;
  if ((rc = alloc(ctx, (void**)&ctx->d_res, sizeof(double) * kResLen))) goto bad;
  if ((rc = alloc(ctx, (void**)&ctx->d_bad, sizeof(int)))) goto bad;
  if (hipMemset(ctx->d_pb[0], 0, sizeof(double) * kPartialStride * (size_t)maxb) != hipSuccess ||
      hipMemset(ctx->d_pb[1], 0, sizeof(double) * kPartialStride * (size_t)maxb) != hipSuccess ||
      hipMemset(ctx->d_res, 0, sizeof(double) * kResLen) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_res, sizeof(double) * kResLen, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_la, sizeof(double) * 8, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_la[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_la[1], hipEventDisableTiming) != hipSuccess) {
    rc = TDGL_E_HIP;
    goto bad;
  }
  *out = ctx;
  return TDGL_OK;
bad:
  tdgl_ctx_destroy(ctx);
  return rc;
}

int tdgl_ctx_destroy(tdgl_ctx* ctx) {
  if (!ctx) return TDGL_OK;
  if (ctx->st) hipStreamSynchronize(ctx->st);
  else hipDeviceSynchronize();
  for (void* p : {(void*)ctx->d_flags, (void*)ctx->d_k2, (void*)ctx->d_sc, (void*)ctx->d_bfull,
                  (void*)ctx->d_pb[0], (void*)ctx->d_pb[1], (void*)ctx->d_res,
                  (void*)ctx->d_bad, (void*)ctx->w_x, (void*)ctx->w_gam, (void*)ctx->w_f,
                  (void*)ctx->w_r, (void*)ctx->d_ptab})
    if (p) hipFree(p);
  for (double* p : ctx->P) hipFree(p);
  for (double* p : ctx->AP) hipFree(p);
  if (ctx->own_batch) tdgl_batch_destroy(ctx->own_batch);
  if (ctx->d_acell) hipFree(ctx->d_acell);
  if (ctx->d_vjpb) hipFree(ctx->d_vjpb);
  for (double* p : ctx->d_adjt)
    if (p) hipFree(p);
  ctx->prof.destroy();
  if (ctx->h_res) hipHostFree(ctx->h_res);
  if (ctx->h_la) hipHostFree(ctx->h_la);
  for (hipEvent_t e : ctx->ev_la)
    if (e) hipEventDestroy(e);
  delete ctx;
  return TDGL_OK;
}

int tdgl_set_stencil_kz(tdgl_ctx* ctx, int kz) {
  if (!ctx || kz < 1) return TDGL_E_INVALID;
  const Geo& g = ctx->g;
  HIPCHK(hipStreamSynchronize(ctx->st));
  StencilCfg cfg{g.is3d ? std::min(kz, g.nz) : 1};
  const int maxb = std::max({stencil_blocks(g, cfg), stream_blocks(g), mgs_blocks((2LL + g.nphi) * g.n)});
  if (maxb > ctx->pb_blocks) {  // more work items than partial slots: grow both buffers
    double* nb[2] = {nullptr, nullptr};
    for (int h = 0; h < 2; ++h) {
      int rc;
      if ((rc = alloc(ctx, (void**)&nb[h], sizeof(double) * kPartialStride * (size_t)maxb))) {
        if (nb[0]) hipFree(nb[0]);
        return rc;
      }
      HIPCHK(hipMemset(nb[h], 0, sizeof(double) * kPartialStride * (size_t)maxb));
    }
    HIPCHK(hipFree(ctx->d_pb[0]));
    HIPCHK(hipFree(ctx->d_pb[1]));
    ctx->d_pb[0] = nb[0];
    ctx->d_pb[1] = nb[1];
    ctx->pb_count = 0;
    ctx->pb_blocks = maxb;
  }
  ctx->cfg = cfg;
  ctx->kz_user = cfg.kz;
  if (ctx->own_batch) return tdgl_batch_set_kz(ctx->own_batch, 0, cfg.kz);
  return TDGL_OK;
}

int tdgl_get_stencil_kz(const tdgl_ctx* ctx) { return ctx ? ctx->cfg.kz : -1; }

int tdgl_set_sweep_tile(tdgl_ctx* ctx, int shape) {
  if (!ctx || shape < 0 || shape >= kSweepShapes) return TDGL_E_INVALID;
  ctx->tile_user = shape;
  if (ctx->own_batch) return tdgl_batch_set_sweep_tile(ctx->own_batch, shape);
  return TDGL_OK;
}

int tdgl_get_sweep_tile(const tdgl_ctx* ctx) { return ctx ? ctx->tile_user : -1; }

int tdgl_set_stream(tdgl_ctx* ctx, void* s) {
  if (!ctx) return TDGL_E_INVALID;
  ctx->st = (hipStream_t)s;
  return TDGL_OK;
}

const char* tdgl_last_error(const tdgl_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }
int64_t tdgl_state_len(const tdgl_ctx* ctx) { return ctx ? ctx->len : -1; }
int64_t tdgl_n_interior(const tdgl_ctx* ctx) { return ctx ? ctx->g.n : -1; }

int tdgl_set_material(tdgl_ctx* ctx, const uint8_t* link_bits_full, const double* kappa_int,
                      const double* sc_int) {
  if (!ctx) return TDGL_E_INVALID;
  const Geo& g = ctx->g;
  const long long dimx = g.mk * (g.is3d ? (g.Nz + 1) : 1);
  const long long n = g.n;
  Mat& m = ctx->mt;
  HIPCHK(hipStreamSynchronize(ctx->st));
  // reference interior order -> internal order
  auto int_of_ref = [&](long long r) {
    const long long k1 = r % g.nz, t = r / g.nz, j1 = t % g.ny, i1 = t / g.ny;
    return i1 + g.nx * (j1 + g.ny * k1);
  };
  auto full_of_ref = [&](long long r) {
    const long long k1 = r % g.nz, t = r / g.nz, j1 = t % g.ny, i1 = t / g.ny;
    return (i1 + 1) + (long long)g.mj * (j1 + 1) + (g.is3d ? g.mk * (k1 + 1) : 0LL);
  };
  std::vector<uint8_t> flags;
  bool any_flags = false;
  if (link_bits_full) {
    flags.assign(link_bits_full, link_bits_full + dimx);
    for (auto& f : flags) f &= (kHoleX | kHoleY | kHoleZ);
    any_flags = true;
  }
  // kappa: uniform -> scalar coefficients, otherwise kappa^2 per node
  if (m.k2) {
    hipFree(ctx->d_k2);
    ctx->d_k2 = nullptr;
    m.k2 = nullptr;
  }
  if (kappa_int) {
    bool uniform = true;
    for (long long r = 1; r < n && uniform; ++r) uniform = (kappa_int[r] == kappa_int[0]);
    if (uniform) {
      set_uniform_coeffs(ctx, kappa_int[0]);
    } else {
      std::vector<double> k2(n);
      for (long long r = 0; r < n; ++r) k2[int_of_ref(r)] = kappa_int[r] * kappa_int[r];
      ctx->k2max = *std::max_element(k2.begin(), k2.end());
      int rc;
      if ((rc = alloc(ctx, (void**)&ctx->d_k2, sizeof(double) * n))) return rc;
      HIPCHK(hipMemcpy(ctx->d_k2, k2.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      m.k2 = ctx->d_k2;
    }
  } else {
    set_uniform_coeffs(ctx, ctx->grid.kappa);
  }
  // sc: binary -> insulator bits, otherwise per node
  if (ctx->d_sc) {
    hipFree(ctx->d_sc);
    ctx->d_sc = nullptr;
  }
  m.sc = nullptr;
  m.sc_bits = 0;
  m.has_mat = sc_int != nullptr;
  if (sc_int) {
    bool binary = true;
    for (long long r = 0; r < n && binary; ++r) binary = (sc_int[r] == 0.0 || sc_int[r] == 1.0);
    if (binary) {
      if (!any_flags) {
        flags.assign(dimx, 0);
        any_flags = true;
      }
      for (long long r = 0; r < n; ++r)
        if (sc_int[r] == 0.0) flags[full_of_ref(r)] |= kInsulator;
      m.sc_bits = 1;
    } else {
      std::vector<double> sc(n);
      for (long long r = 0; r < n; ++r) sc[int_of_ref(r)] = sc_int[r];
      int rc;
      if ((rc = alloc(ctx, (void**)&ctx->d_sc, sizeof(double) * n))) return rc;
      HIPCHK(hipMemcpy(ctx->d_sc, sc.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      m.sc = ctx->d_sc;
    }
  }
  if (ctx->d_flags) {
    hipFree(ctx->d_flags);
    ctx->d_flags = nullptr;
  }
  m.flags = nullptr;
  if (any_flags) {
    int rc;
    if ((rc = alloc(ctx, (void**)&ctx->d_flags, (size_t)dimx))) return rc;
    HIPCHK(hipMemcpy(ctx->d_flags, flags.data(), (size_t)dimx, hipMemcpyHostToDevice));
    m.flags = ctx->d_flags;
  }
  return TDGL_OK;
}

int tdgl_set_field(tdgl_ctx* ctx, double bx, double by, double bz) {
  if (!ctx) return TDGL_E_INVALID;
  Bnd& b = ctx->b;
  b.cx = (bx * b.hy) * b.hz;
  b.cy = (by * b.hz) * b.hx;
  b.cz = (bz * b.hx) * b.hy;
  b.bx = b.by = b.bz = nullptr;
  return TDGL_OK;
}

int tdgl_set_field_vectors(tdgl_ctx* ctx, const double* bx, const double* by, const double* bz) {
  if (!ctx || !bx || !by || !bz) return TDGL_E_INVALID;
  const Geo& g = ctx->g;
  const long long dimx = g.mk * (g.is3d ? (g.Nz + 1) : 1);
  if (!ctx->d_bfull) {
    int rc;
    if ((rc = alloc(ctx, (void**)&ctx->d_bfull, sizeof(double) * 3 * dimx))) return rc;
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipMemcpy(ctx->d_bfull, bx, sizeof(double) * dimx, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_bfull + dimx, by, sizeof(double) * dimx, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_bfull + 2 * dimx, bz, sizeof(double) * dimx, hipMemcpyHostToDevice));
  ctx->b.bx = ctx->d_bfull;
  ctx->b.by = ctx->d_bfull + dimx;
  ctx->b.bz = ctx->d_bfull + 2 * dimx;
  return TDGL_OK;
}

int tdgl_permute_in(tdgl_ctx* ctx, const void* ref, double* out) {
  if (!ctx || !ref || !out) return TDGL_E_INVALID;
  HIPCHK(hipMemsetAsync(ctx->d_bad, 0, sizeof(int), ctx->st));
  launch_permute_in(ctx->g, ref, out, ctx->d_bad, ctx->st);
  LAUNCHCHK();
  int bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, ctx->d_bad, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (bad) return fail(ctx, TDGL_E_COMPLEX_PHI, "link variables must be real (Im(phi) == 0)");
  return TDGL_OK;
}

int tdgl_permute_out(tdgl_ctx* ctx, const double* in, void* ref) {
  if (!ctx || !ref || !in) return TDGL_E_INVALID;
  launch_permute_out(ctx->g, in, ref, ctx->st);
  LAUNCHCHK();
  return TDGL_OK;
}

int tdgl_rhs(tdgl_ctx* ctx, const double* x, double* f) {
  if (!ctx || !x || !f) return TDGL_E_INVALID;
  Epi e{};
  e.out0 = f;
  launch_stencil(M_RHS, false, ctx->g, ctx->b, ctx->mt, x, nullptr, 0.0, e, ctx->cfg, red_next(ctx),
                 ctx->st);
  LAUNCHCHK();
  return TDGL_OK;
}

int tdgl_jvp_trap_fd(tdgl_ctx* ctx, const double* x, const double* f0, const double* p, double eps,
                     double dt, double* ap) {
  if (!ctx || !x || !f0 || !p || !ap) return TDGL_E_INVALID;
  Epi e{};
  e.out0 = ap;
  e.f0 = f0;
  e.ap0 = nullptr;
  e.hdt = dt / 2.0;
  e.inv_eps = 1.0 / eps;
  red_commit(ctx, launch_stencil(M_JVP, true, ctx->g, ctx->b, ctx->mt, x, p, eps, e, ctx->cfg,
                                 red_next(ctx), ctx->st));
  LAUNCHCHK();
  return TDGL_OK;
}

int tdgl_jvp_trap32(tdgl_ctx* ctx, const float* x32, const float* p32, double dt, float* ap32) {
  if (!ctx || !x32 || !p32 || !ap32) return TDGL_E_INVALID;
  return jvp32_single(ctx, x32, p32, dt, ap32);
}

int tdgl_precond_cheb32(tdgl_ctx* ctx, const float* x32, const float* v32, double dt, int degree,
                        float* w32, float* z32, double* info) {
  if (!ctx || !x32 || !v32 || !w32 || !z32 || !info || !(dt > 0.0)) return TDGL_E_INVALID;
  if (degree > kMaxChebDeg) return fail(ctx, TDGL_E_INVALID, "Chebyshev degree above 64");
  return precond_single(ctx, x32, v32, dt, degree, w32, z32, info);
}

}  // extern "C"

namespace {

// The cell adjoint; every kernel that fills it (k_cells_gather, k_bz_cells)
// writes every cell and component.
int ensure_acell(tdgl_ctx* ctx) {
  const Geo& g = ctx->g;
  const long long dimx = g.mk * (g.is3d ? (long long)(g.Nz + 1) : 1LL);
  if (!ctx->d_acell) {
    int rc;
    if ((rc = alloc(ctx, (void**)&ctx->d_acell, sizeof(double) * 5 * (size_t)dimx))) return rc;
  }
  return TDGL_OK;
}
// The per-node planes of the VJP (every entry written by each launch_vjp_cells,
// read once by the launch_cells_gather after it).
int ensure_vjpb(tdgl_ctx* ctx) {
  if (ctx->d_vjpb) return TDGL_OK;
  const size_t bytes = sizeof(double) * (size_t)vjp_planes() * (size_t)ctx->g.n;
  int rc;
  if ((rc = alloc(ctx, (void**)&ctx->d_vjpb, bytes))) return rc;
  HIPCHK(hipMemsetAsync(ctx->d_vjpb, 0, bytes, ctx->st));
  return TDGL_OK;
}

// d<w, .>/d(Bx, By, Bz) from the cell adjoint (uniform applied field only).
int field_grad(tdgl_ctx* ctx, double* dfield) {
  if (ctx->b.bx || ctx->b.by || ctx->b.bz)
    return fail(ctx, TDGL_E_INVALID, "field gradients need a uniform applied field (tdgl_set_field)");
  red_commit(ctx, launch_field_grad(ctx->g, ctx->mt.flags, ctx->d_acell, red_next(ctx).pout, ctx->st));
  red_finalize(ctx, 3, 0, RS_BETA);
  LAUNCHCHK();
  int rc;
  if ((rc = readback(ctx, RS_BETA, 3))) return rc;
  const Bnd& b = ctx->b;
  // cx = (Bx hy) hz, cy = (By hz) hx, cz = (Bz hx) hy
  dfield[0] = ctx->h_res[RS_BETA] * (b.hy * b.hz);
  dfield[1] = ctx->h_res[RS_BETA + 1] * (b.hz * b.hx);
  dfield[2] = ctx->h_res[RS_BETA + 2] * (b.hx * b.hy);
  return TDGL_OK;
}

}  // namespace

extern "C" {

int tdgl_vjp_f(tdgl_ctx* ctx, const double* x, const double* w, double* gx, double* dfield,
               double* dkappa, double* dsc) {
  if (!ctx || !x || !w || !gx) return TDGL_E_INVALID;
  int rc;
  if ((rc = ensure_acell(ctx)) || (rc = ensure_vjpb(ctx))) return rc;
  Red r = red_next(ctx);
  launch_vjp_cells(ctx->g, ctx->b, ctx->mt, x, w, ctx->d_vjpb, dsc, ctx->kappa_uniform,
                   dkappa ? r.pout : nullptr, ctx->st);
  launch_cells_gather(ctx->g, ctx->d_vjpb, ctx->d_acell, ctx->st);
  LAUNCHCHK();
  if (dkappa) {
    red_commit(ctx, stream_blocks(ctx->g));
    red_finalize(ctx, 1, 0, RS_DOT);
    LAUNCHCHK();
    if ((rc = readback(ctx, RS_DOT, 1))) return rc;
    *dkappa = ctx->h_res[RS_DOT];
  }
  launch_bc_gather(ctx->g, ctx->mt.flags, ctx->d_acell, gx, ctx->st);
  LAUNCHCHK();
  if (dfield && (rc = field_grad(ctx, dfield))) return rc;
  return TDGL_OK;
}

int tdgl_vjp_bz(tdgl_ctx* ctx, const double* gbz, double* gx, double* dfield) {
  if (!ctx || !gbz || !gx) return TDGL_E_INVALID;
  int rc;
  if ((rc = ensure_acell(ctx))) return rc;
  launch_bz_cells(ctx->g, ctx->b, gbz, ctx->d_acell, ctx->st);
  launch_bc_gather(ctx->g, ctx->mt.flags, ctx->d_acell, gx, ctx->st);
  LAUNCHCHK();
  if (dfield && (rc = field_grad(ctx, dfield))) return rc;
  return TDGL_OK;
}

int tdgl_dot(tdgl_ctx* ctx, const double* a, const double* b, double* out) {
  if (!ctx || !a || !b || !out) return TDGL_E_INVALID;
  red_commit(ctx, launch_dot(ctx->g, a, b, red_next(ctx), ctx->st));
  red_finalize(ctx, 1, 0, RS_DOT);
  LAUNCHCHK();
  int rc;
  if ((rc = readback(ctx, RS_DOT, 1))) return rc;
  *out = ctx->h_res[RS_DOT];
  return TDGL_OK;
}

int tdgl_norm_inf(tdgl_ctx* ctx, const double* a, double* out) {
  if (!ctx || !a || !out) return TDGL_E_INVALID;
  red_commit(ctx, launch_norm_inf(ctx->g, a, red_next(ctx), ctx->st));
  red_finalize(ctx, 0, 1, RS_DOT);
  LAUNCHCHK();
  int rc;
  if ((rc = readback(ctx, RS_DOT, 1))) return rc;
  *out = ctx->h_res[RS_DOT];
  return TDGL_OK;
}

int tdgl_euler_step(tdgl_ctx* ctx, double* X, double dt) {
  if (!ctx || !X) return TDGL_E_INVALID;
  int rc;
  if ((rc = ensure_work(ctx))) return rc;
  Epi e{};
  e.out0 = ctx->w_x;
  e.dt = dt;
  launch_stencil(M_EULER, false, ctx->g, ctx->b, ctx->mt, X, nullptr, 0.0, e, ctx->cfg, red_next(ctx),
                 ctx->st);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(X, ctx->w_x, sizeof(double) * ctx->len, hipMemcpyDeviceToDevice, ctx->st));
  return TDGL_OK;
}

int tdgl_observables(tdgl_ctx* ctx, const double* x, double* bout, double* jout) {
  if (!ctx || !x) return TDGL_E_INVALID;
  launch_observables(ctx->g, ctx->b, ctx->mt, x, bout, jout, ctx->st);
  LAUNCHCHK();
  return TDGL_OK;
}

}  // extern "C"

namespace {

// Newton residual at the current iterate w_x: f, r = -g, ||x||^2, ||r||^2, ||g||_inf.
int newton_eval(tdgl_ctx* ctx, double hdt, tdgl_step_stats* s) {
  Epi e{};
  e.out0 = ctx->w_f;
  e.out1 = ctx->w_r;
  e.gam = ctx->w_gam;
  e.hdt = hdt;
  const double bytes = 4.0 * state_bytes(ctx);
  const int tok = prof_begin(ctx, TDGL_K_NEWTON, bytes);
  red_commit(ctx, launch_stencil(M_NEWTON, false, ctx->g, ctx->b, ctx->mt, ctx->w_x, nullptr, 0.0, e,
                                 ctx->cfg, red_next(ctx), ctx->st));
  prof_end(ctx, tok);
  red_finalize(ctx, 2, 1, RS_XX);  // ||x||^2, ||r||^2, ||g||_inf
  LAUNCHCHK();
  s->rhs_evals++;
  s->bytes_moved += bytes;
  return TDGL_OK;
}

// Vectors the blocked Gram-Schmidt passes of TGCR iteration i move, in units
// of S (each pass: Ap, its old and new vectors, Ap again).
double mgs_vectors(int i) {
  double v = 0.0;
  for (int j = 0, no = 1; i > 0;) {
    const int nb = j + no, nn = std::min(kMgsBlock, i - nb);
    v += 2.0 + no + nn;
    if (nn == 0) break;
    j = nb;
    no = nn;
  }
  return v;
}

// The operator TGCR applies: the reference's finite-difference JVP of f at w_x
// (default), or the adjoint step operator A^T = I - hdt J(x1)^T.
struct TgcrOp {
  int vjp = 0;
  const double* x1 = nullptr;   // vjp: the linearisation point
  double* acc_out = nullptr;    // x += sum alpha p accumulates here (default w_x)
  // right preconditioner M = the degree-d Chebyshev iterate for A z = r (A^T z
  // = r for vjp) on [a, b] (p_k = M r_k, the direction's source in k_final_p);
  // 0: none
  int cheb = 0;
  double thinv = 0.0;           // 1 / theta
  double om[kMaxChebDeg + 1] = {0.0};
};

// The Chebyshev right preconditioner on [1 - hdt, b] (the mixed method's
// interval: J^T has J's spectrum); request < 0: the automatic degree (the
// spectrum within 0.25 of 1), > 0: that degree; off (op.cheb = 0) when the
// interval nears 0.  Allocates the recurrence's three vectors.
int cheb_op(tdgl_ctx* ctx, double hdt, int request, TgcrOp& op) {
  op.cheb = 0;
  const double lo = 1.0 - hdt, hi = cheb_upper_bound(ctx, hdt);
  if (request == 0 || !(lo >= 0.1) || !(hi > lo)) return TDGL_OK;
  const double th = 0.5 * (hi + lo), de = 0.5 * (hi - lo), sig = th / de, s2 = sig * sig;
  int d = request > 0 ? request : (int)ceil(acosh(4.0) / acosh(sig));
  d = std::max(1, std::min(d, kMaxChebDeg));
  op.cheb = d;
  op.thinv = 1.0 / th;
  op.om[1] = 1.0;
  if (d >= 2) op.om[2] = 2.0 * s2 / (2.0 * s2 - 1.0);
  for (int m = 3; m <= d; ++m) op.om[m] = 1.0 / (1.0 - op.om[m - 1] / (4.0 * s2));
  const size_t bytes = sizeof(double) * (size_t)ctx->len;
  int rc;
  for (double*& p : ctx->d_adjt)
    if (!p && (rc = alloc(ctx, (void**)&p, bytes))) return rc;
  return TDGL_OK;
}

// tgcr_matrix_free_trap (tgcr.py:96-156) on (I - dt/2 J) dx = r, r = -g held in
// w_r, linearised at w_x with f(w_x) in w_f.  On success w_x += dx.
// Returns TDGL_OK and sets *ok; *ok = 0 is the reference's empty-array return.
int tgcr(tdgl_ctx* ctx, double hdt, const tdgl_solver_opts* o, double rr, double xx, int* ok,
         tdgl_step_stats* s, const TgcrOp& op = TgcrOp{}) {
  const Geo& g = ctx->g;
  const double S = state_bytes(ctx);
  const long long N = (1 + g.nphi) * g.n;  // complex entries of the reference state
  long long maxk = o->tgcr_max_iter;
  if (maxk <= 0) maxk = std::max<long long>(N, llround(0.2 * (double)N));
  const double r0 = sqrt(rr);
  double rk = r0;
  int nvec = 0, k = 0, rc;
  *ok = 1;
  s->tgcr_r0 = r0;
  s->tgcr_rk = r0;
  if (r0 == 0.0) {  // tgcr.py:118-119: x = 0
    s->tgcr_iters_last = 0;
    return TDGL_OK;
  }
  const double xn = sqrt(xx);
  if (op.cheb >= 1)  // the preconditioned breakdown test is relative to ||r_0||
    HIPCHK(hipMemcpyAsync(ctx->d_res + RS_R0, &r0, sizeof(double), hipMemcpyHostToDevice, ctx->st));
  // Lookahead: without profiling, iteration k + 1 is queued before iteration
  // k's residual comes back, so the GPU never waits for the host.  The device
  // evaluates the loop's exit test itself (k_tgcr_check -> res[RS_STOP]); the
  // kernels of an iteration queued after the exit return at once, so a
  // speculative iteration leaves no trace.  eps (tgcr.py:129) is formed on the
  // device from ||r||^2 with the host's expression.
  const bool ahead = !ctx->prof.on;
  HIPCHK(hipMemsetAsync(ctx->d_res + RS_STOP, 0, sizeof(double), ctx->st));
  const double eps_num = o->eps_mf * (1.0 + xn);
  long long queued = 0;
  auto enqueue = [&](long long kk) -> int {
    const int i = (int)(kk - 1);  // every earlier iteration added a vector
    int rc2;
    if ((rc2 = ensure_vec(ctx, i))) return rc2;
    int tok;
    const double* zsrc = nullptr;  // the direction's source when preconditioned
    if (op.vjp) {  // Ap = z - hdt J(x1)^T z (z = M r, or r), then <Ap, Ap_0> (or ||Ap||^2)
      const double* stop = ctx->d_res + RS_STOP;
      // J(x1)^T v into AP[i], or (ch.z) the recurrence step formed from it
      auto jt = [&](const double* v, const AdjCheb& ch) {
        launch_vjp_cells(g, ctx->b, ctx->mt, op.x1, v, ctx->d_vjpb, nullptr, 0.0, nullptr, ctx->st, stop);
        launch_planes_gather(g, ctx->mt.flags, ctx->d_vjpb, ctx->AP[i], ctx->st, stop, ch);
      };
      tok = prof_begin(ctx, TDGL_K_JVP, (i > 0 ? 5.0 : 4.0) * S, 4);
      const double* z = ctx->w_r;
      if (op.cheb >= 1) {
        double* const* T = ctx->d_adjt;
        launch_scale(g, T[1], op.thinv, ctx->w_r, ctx->st);  // z_1 = r / theta
        for (int m = 2; m <= op.cheb; ++m) {
          AdjCheb ch;
          ch.z = T[m % 3];
          ch.zp = T[(m - 1) % 3];
          ch.r = ctx->w_r;
          ch.zpp = m >= 3 ? T[(m - 2) % 3] : nullptr;
          ch.om = op.om[m];
          ch.thinv = op.thinv;
          ch.hdt = hdt;
          jt(ch.zp, ch);
        }
        z = zsrc = T[op.cheb % 3];
      }
      jt(z, AdjCheb{});
      red_commit(ctx, launch_adj_combine(g, z, ctx->AP[i], hdt, (i > 0) ? ctx->AP[0] : nullptr,
                                         red_next(ctx), ctx->st));
      prof_end(ctx, tok);
    } else {
      Epi e{};
      e.out0 = ctx->AP[i];
      e.f0 = ctx->w_f;
      e.ap0 = (i > 0) ? ctx->AP[0] : nullptr;
      e.hdt = hdt;
      e.tgcr = 1;
      e.eps_num = eps_num;
      const double* z = ctx->w_r;
      if (op.cheb >= 1) {  // z = M r by the recurrence, A v = v - hdt J v by the FD JVP (eps from ||r||)
        double* const* T = ctx->d_adjt;
        tok = prof_begin(ctx, TDGL_K_JVP, 4.0 * S * op.cheb, op.cheb);
        launch_scale(g, T[1], op.thinv, ctx->w_r, ctx->st);  // z_1 = r / theta
        Epi ei = e;
        ei.ap0 = nullptr;
        ei.cr = ctx->w_r;
        ei.cthinv = op.thinv;
        for (int m = 2; m <= op.cheb; ++m) {
          // z_m from A z_{m-1} in the stencil's epilogue (its partials land in the
          // unread buffer and are not committed)
          ei.cz = T[m % 3];
          ei.czpp = m >= 3 ? T[(m - 2) % 3] : nullptr;
          ei.com = op.om[m];
          launch_stencil(M_JVP, true, g, ctx->b, ctx->mt, ctx->w_x, T[(m - 1) % 3], 0.0, ei, ctx->cfg,
                         red_next(ctx), ctx->st);
        }
        prof_end(ctx, tok);
        z = zsrc = T[op.cheb % 3];
      }
      // bytes: read x, p, f0 (+ Ap_0 for beta_0), write Ap
      tok = prof_begin(ctx, TDGL_K_JVP, (i > 0 ? 5.0 : 4.0) * S);
      red_commit(ctx, launch_stencil(M_JVP, true, g, ctx->b, ctx->mt, ctx->w_x, z, 0.0, e,
                                     ctx->cfg, red_next(ctx), ctx->st));
      prof_end(ctx, tok);
    }
    // Gram-Schmidt against Ap_0 .. Ap_{i-1} (tgcr.py:135-140) in blocked
    // passes: each subtracts the previous pass's new vectors (the first: Ap_0,
    // whose beta the JVP reduced) and reduces the dots with the next
    // kMgsBlock vectors; the last reduces ||Ap||^2.
    for (int j = 0, no = 1, first = 1; i > 0; first = 0) {
      const int nb = j + no;
      const int nn = std::min(kMgsBlock, i - nb);
      MgsVecs v{};
      for (int l = 0; l < no; ++l) v.a[l] = ctx->AP[j + l];
      for (int m = 0; m < nn; ++m) v.b[m] = ctx->AP[nb + m];
      tok = prof_begin(ctx, nn ? TDGL_K_MGS : TDGL_K_MGS_LAST, (2.0 + no + nn) * S);  // Ap, old, new; Ap
      red_commit(ctx, launch_mgs_block(g, ctx->AP[i], v, no, nn, RS_BETA + j, first, red_next(ctx),
                                       ctx->st));
      prof_end(ctx, tok);
      if (nn == 0) break;
      j = nb;
      no = nn;
    }
    // read r, p_0..p_{i-1}, Ap; write p_i, Ap
    tok = prof_begin(ctx, TDGL_K_FINAL_P, (double)(i + 4) * S);
    red_commit(ctx, launch_final_p(g, ctx->w_r, ctx->d_ptab, i, ctx->AP[i], ctx->P[i], red_next(ctx),
                                   ctx->st, zsrc));
    prof_end(ctx, tok);
    tok = prof_begin(ctx, TDGL_K_UPDATE_R, 3.0 * S);  // read r, Ap; write r
    red_commit(ctx, launch_update_r(g, ctx->w_r, ctx->AP[i], RS_ALPHA + i, red_next(ctx), ctx->st));
    prof_end(ctx, tok);
    launch_tgcr_check(r0, o->tol_gcr, kk, maxk, red_next(ctx), ctx->st);
    LAUNCHCHK();
    double* h = ctx->h_la + 4 * (kk & 1);
    HIPCHK(hipMemcpyAsync(h, ctx->d_res + RS_BRK, sizeof(double) * 3, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipEventRecord(ctx->ev_la[kk & 1], ctx->st));
    queued = kk;
    return TDGL_OK;
  };
  while (rk / r0 > o->tol_gcr && k < maxk) {
    ++k;
    const int i = nvec;
    if (queued < k && (rc = enqueue(k))) return rc;
    if (ahead && k < maxk && (rc = enqueue(k + 1))) return rc;
    HIPCHK(hipEventSynchronize(ctx->ev_la[k & 1]));
    if (!ctx->prof.pending.empty()) prof_harvest(ctx);
    const double* h = ctx->h_la + 4 * (k & 1);  // brk, (xx), rr
    s->jvp_evals += std::max(1, op.cheb);
    s->bytes_moved += ((i > 0 ? 5.0 : 4.0) + mgs_vectors(i)) * S;
    if (h[0] != 0.0) break;  // tgcr.py:141-142
    s->bytes_moved += (double)(i + 4) * S + 3.0 * S;
    rk = sqrt(h[2]);
    nvec = i + 1;
    if (!std::isfinite(rk)) return fail(ctx, TDGL_E_NONFINITE, "TGCR residual is not finite");
  }
  s->tgcr_iters_last = k;
  s->tgcr_iters_total += k;
  s->tgcr_rk = rk;
  if (rk > o->tol_gcr * r0) {  // tgcr.py:154-155
    *ok = 0;
    return TDGL_OK;
  }
  const int tok = prof_begin(ctx, TDGL_K_ACCUM, (double)(nvec + 2) * S);  // read x, p_i; write x
  red_commit(ctx, launch_accum(g, op.acc_out ? op.acc_out : ctx->w_x, ctx->d_ptab, nvec, red_next(ctx),
                               ctx->st));
  prof_end(ctx, tok);
  red_finalize(ctx, 0, 1, RS_DXINF);
  LAUNCHCHK();
  s->bytes_moved += (double)(nvec + 2) * S;
  return TDGL_OK;
}

// The Newton loop of newton_gcr_trap (newton.py:94-120) on w_x (the start,
// then each iterate) and w_gam.  On return w_x holds the last iterate:
// the converged one, the one whose TGCR failed (newton.py:102-106), or the
// last after max_iter iterations.
int newton_solve(tdgl_ctx* ctx, double hdt, const tdgl_solver_opts* o, tdgl_step_stats* s,
                 const TgcrOp& op = TgcrOp{}) {
  int rc;
  if ((rc = newton_eval(ctx, hdt, s))) return rc;
  if ((rc = readback(ctx, RS_XX, 3))) return rc;  // xx, rr, ginf
  for (int it = 1; it <= o->newton_max_iter; ++it) {
    const double rr = ctx->h_res[RS_RR], xx = ctx->h_res[RS_XX];
    int ok = 0;
    if ((rc = tgcr(ctx, hdt, o, rr, xx, &ok, s, op))) return rc;
    if (!ok) {  // newton.py:102-106
      s->tgcr_failed = 1;
      s->newton_iters = it;
      s->converged = 0;
      return TDGL_OK;
    }
    if ((rc = newton_eval(ctx, hdt, s))) return rc;  // newton.py:109-113
    if ((rc = readback(ctx, RS_XX, 4))) return rc;  // xx, rr, ginf, dxinf
    s->g_inf = ctx->h_res[RS_GINF];
    s->dx_inf = ctx->h_res[RS_DXINF];
    if (!std::isfinite(s->g_inf) || !std::isfinite(s->dx_inf))
      return fail(ctx, TDGL_E_NONFINITE, "Newton residual is not finite");
    if (s->g_inf <= o->newton_tol_f && s->dx_inf <= o->newton_tol_dx) {  // newton.py:118
      s->converged = 1;
      s->newton_iters = it;
      return TDGL_OK;
    }
  }
  s->converged = 0;
  s->newton_iters = o->newton_max_iter;
  return TDGL_OK;
}

}  // namespace

extern "C" int tdgl_trap_step(tdgl_ctx* ctx, double* X, double dt, const tdgl_solver_opts* o,
                              tdgl_step_stats* s) {
  if (!ctx || !X || !o || !s) return TDGL_E_INVALID;
  if (o->precision != TDGL_PREC_F64 && o->precision != TDGL_PREC_MIXED)
    return fail(ctx, TDGL_E_INVALID, "precision must be TDGL_PREC_F64 or TDGL_PREC_MIXED");
  if (o->x_prev && !o->dt_prev) return fail(ctx, TDGL_E_INVALID, "x_prev needs dt_prev");
  if (o->precision == TDGL_PREC_MIXED) return mixed_step_single(ctx, X, dt, o, s);
  if (o->f64_precond < -1 || o->f64_precond > kMaxChebDeg)
    return fail(ctx, TDGL_E_INVALID, "f64_precond must be -1 (automatic), 0 (none) or a degree up to 64");
  memset(s, 0, sizeof(*s));
  int rc;
  if ((rc = ensure_work(ctx))) return rc;
  if ((rc = ensure_vec(ctx, 0))) return rc;
  const double hdt = dt / 2.0;
  const double S = state_bytes(ctx);
  // predictor + gamma (integrators.py:156-158)
  Epi e{};
  e.out0 = ctx->w_x;
  e.out1 = ctx->w_gam;
  e.dt = dt;
  e.hdt = hdt;
  const int tok = prof_begin(ctx, TDGL_K_PRED, 3.0 * S);  // read X; write x_pred, gamma
  launch_stencil(M_PRED, false, ctx->g, ctx->b, ctx->mt, X, nullptr, 0.0, e, ctx->cfg, red_next(ctx),
                 ctx->st);
  prof_end(ctx, tok);
  LAUNCHCHK();
  s->rhs_evals++;
  s->bytes_moved += 3.0 * S;
  if (o->x_prev && o->dt_prev[0] > 0.0) {  // Newton starts from the extrapolated state
    const int tk = prof_begin(ctx, TDGL_K_PRED, 3.0 * S);
    launch_extrap64(X, o->x_prev, dt / o->dt_prev[0], ctx->w_x, ctx->len, ctx->st);
    prof_end(ctx, tk);
    LAUNCHCHK();
    s->bytes_moved += 3.0 * S;
  }
  TgcrOp op;  // the FP64 TGCR's optional Chebyshev right preconditioner
  if (o->f64_precond && (rc = cheb_op(ctx, hdt, o->f64_precond, op))) return rc;
  if ((rc = newton_solve(ctx, hdt, o, s, op))) return rc;
  if (s->converged) {
    HIPCHK(hipMemcpyAsync(X, ctx->w_x, sizeof(double) * ctx->len, hipMemcpyDeviceToDevice, ctx->st));
    s->bytes_moved += 2.0 * S;
  }
  return TDGL_OK;
}

// newton_gcr_trap (newton.py:76-120) as an entry point of its own: Newton from
// x (in: x0, out: the last iterate, also when Newton fails) on
// g(x) = x - dt/2 f(x) - gamma.
extern "C" int tdgl_newton_trap(tdgl_ctx* ctx, double* x, const double* gamma, double dt,
                                const tdgl_solver_opts* o, tdgl_step_stats* s) {
  if (!ctx || !x || !gamma || !o || !s) return TDGL_E_INVALID;
  if (o->precision != TDGL_PREC_F64)
    return fail(ctx, TDGL_E_INVALID, "tdgl_newton_trap runs the plain FP64 method only");
  memset(s, 0, sizeof(*s));
  int rc;
  if ((rc = ensure_work(ctx)) || (rc = ensure_vec(ctx, 0))) return rc;
  const size_t bytes = sizeof(double) * (size_t)ctx->len;
  HIPCHK(hipMemcpyAsync(ctx->w_x, x, bytes, hipMemcpyDeviceToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(ctx->w_gam, gamma, bytes, hipMemcpyDeviceToDevice, ctx->st));
  if ((rc = newton_solve(ctx, dt / 2.0, o, s))) return rc;
  HIPCHK(hipMemcpyAsync(x, ctx->w_x, bytes, hipMemcpyDeviceToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return TDGL_OK;
}

// tgcr_matrix_free_trap (tgcr.py:96-156) as an entry point of its own: dx with
// (I - dt/2 J(x_lin)) dx = b, J by the finite difference of f at x_lin.
extern "C" int tdgl_tgcr_trap(tdgl_ctx* ctx, const double* x_lin, const double* b, double dt,
                              const tdgl_solver_opts* o, double* dx, int* converged,
                              tdgl_step_stats* s) {
  if (!ctx || !x_lin || !b || !o || !dx || !converged || !s) return TDGL_E_INVALID;
  if (o->precision != TDGL_PREC_F64)
    return fail(ctx, TDGL_E_INVALID, "tdgl_tgcr_trap runs the plain FP64 method only");
  memset(s, 0, sizeof(*s));
  int rc;
  if ((rc = ensure_work(ctx)) || (rc = ensure_vec(ctx, 0))) return rc;
  const size_t bytes = sizeof(double) * (size_t)ctx->len;
  HIPCHK(hipMemcpyAsync(ctx->w_x, x_lin, bytes, hipMemcpyDeviceToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(ctx->w_r, b, bytes, hipMemcpyDeviceToDevice, ctx->st));
  HIPCHK(hipMemsetAsync(dx, 0, bytes, ctx->st));  // x = 0 (tgcr.py:114)
  Epi e{};
  e.out0 = ctx->w_f;  // f(x_lin), the base of the finite difference
  launch_stencil(M_RHS, false, ctx->g, ctx->b, ctx->mt, ctx->w_x, nullptr, 0.0, e, ctx->cfg,
                 red_next(ctx), ctx->st);
  s->rhs_evals++;
  red_commit(ctx, launch_dot(ctx->g, ctx->w_x, ctx->w_x, red_next(ctx), ctx->st));
  red_finalize(ctx, 1, 0, RS_XX);
  red_commit(ctx, launch_dot(ctx->g, ctx->w_r, ctx->w_r, red_next(ctx), ctx->st));
  red_finalize(ctx, 1, 0, RS_RR);  // the JVP forms eps from res[RS_RR]
  LAUNCHCHK();
  if ((rc = readback(ctx, RS_XX, 2))) return rc;
  TgcrOp op;
  op.acc_out = dx;
  int ok = 0;
  if ((rc = tgcr(ctx, dt / 2.0, o, ctx->h_res[RS_RR], ctx->h_res[RS_XX], &ok, s, op))) return rc;
  s->tgcr_failed = !ok;
  *converged = ok;
  HIPCHK(hipStreamSynchronize(ctx->st));
  return TDGL_OK;
}

extern "C" int tdgl_profile_enable(tdgl_ctx* ctx, int enable) {
  if (!ctx) return TDGL_E_INVALID;
  HIPCHK(hipDeviceSynchronize());  // the context's own batch may have used its private stream
  prof_harvest(ctx);
  ctx->prof.on = enable != 0;
  ctx->prof.reset();
  return TDGL_OK;
}

extern "C" int tdgl_profile_read(tdgl_ctx* ctx, int cls, double* ms, int64_t* n, double* bytes) {
  if (!ctx || cls < 0 || cls >= TDGL_K_COUNT) return TDGL_E_INVALID;
  HIPCHK(hipDeviceSynchronize());
  prof_harvest(ctx);
  if (ms) *ms = ctx->prof.ms[cls];
  if (n) *n = ctx->prof.n[cls];
  if (bytes) *bytes = ctx->prof.bytes[cls];
  return TDGL_OK;
}

extern "C" {

int tdgl_adjoint_solve(tdgl_ctx* ctx, const double* x1, const double* rhs, double hdt, double tol,
                       int64_t max_iter, int precond, double* out, int* iters) {
  if (!ctx || !x1 || !rhs || !out || !(tol > 0.0) || !(hdt >= 0.0) || precond < -1 ||
      precond > kMaxChebDeg)
    return TDGL_E_INVALID;
  int rc;
  if ((rc = ensure_work(ctx)) || (rc = ensure_vec(ctx, 0)) || (rc = ensure_acell(ctx)) ||
      (rc = ensure_vjpb(ctx)))
    return rc;
  const Geo& g = ctx->g;
  const size_t bytes = sizeof(double) * (size_t)ctx->len;
  HIPCHK(hipMemsetAsync(out, 0, bytes, ctx->st));
  if (iters) *iters = 0;
  double b2;
  if ((rc = tdgl_dot(ctx, rhs, rhs, &b2))) return rc;
  if (!std::isfinite(b2)) return fail(ctx, TDGL_E_NONFINITE, "adjoint right-hand side is not finite");
  if (b2 == 0.0) return TDGL_OK;
  // unit right-hand side: TGCR's breakdown test (||Ap|| < 1e-14) is absolute
  const double bn = sqrt(b2);
  launch_scale(g, ctx->w_r, 1.0 / bn, rhs, ctx->st);
  double rr;
  if ((rc = tdgl_dot(ctx, ctx->w_r, ctx->w_r, &rr))) return rc;
  HIPCHK(hipMemcpyAsync(ctx->d_res + RS_RR, &rr, sizeof(double), hipMemcpyHostToDevice, ctx->st));
  tdgl_solver_opts o{};
  o.tol_gcr = tol;
  o.tgcr_max_iter = max_iter;
  tdgl_step_stats st{};
  TgcrOp op;
  op.vjp = 1;
  op.x1 = x1;
  op.acc_out = out;
  // Chebyshev right preconditioner (the transposed operator has the same spectrum)
  if ((rc = cheb_op(ctx, hdt, precond, op))) return rc;
  int ok = 0;
  if ((rc = tgcr(ctx, hdt, &o, rr, 0.0, &ok, &st, op))) return rc;
  if (iters) *iters = st.tgcr_iters_last;
  if (!ok) return fail(ctx, TDGL_E_NONFINITE, "adjoint TGCR did not reach the tolerance");
  launch_scale(g, out, bn, out, ctx->st);
  LAUNCHCHK();
  return TDGL_OK;
}

}  // extern "C"
