// Private host-side definitions shared by the C-ABI sources: the context of
// one film, the batch of films, error plumbing and event profiling.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/tdgl_hip.h"
#include "tdgl_launch.hpp"

namespace tdgl {

// Per-kernel-class event timing: a HIP event pair around every bracketed
// launch (or run of launches) on the library's stream.  The events carry no
// system-scope fence: with one, every record writes back and invalidates the
// L2, which would be timed with the kernel it brackets (and slow the next
// one).  They are read only after the stream has been synchronised.
struct ProfSink {
  struct Pending {
    hipEvent_t a, b;
    int cls;
    double bytes;
    int launches;  // consecutive launches of the class inside the bracket
    bool kernel;   // the events ride on the bracket's one launch (KernelTiming)
    bool valid;
  };
  bool on = false;
  std::vector<Pending> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  double ms[TDGL_K_COUNT] = {0};
  long long n[TDGL_K_COUNT] = {0};
  double bytes[TDGL_K_COUNT] = {0};

  // kernel: the bracket holds exactly one launch made through launch(), which
  // records the events at the kernel's start and end.
  int begin(hipStream_t st, int cls, double by, int launches = 1, bool kernel = false) {
    if (!on) return -1;
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!pool.empty()) {
      ev = pool.back();
      pool.pop_back();
    } else if (hipEventCreateWithFlags(&ev.first, hipEventDisableSystemFence) != hipSuccess ||
               hipEventCreateWithFlags(&ev.second, hipEventDisableSystemFence) != hipSuccess) {
      return -1;
    }
    if (kernel) {
      KernelTiming& t = kernel_timing();
      t.start = ev.first;
      t.stop = ev.second;
      t.armed = true;
      t.used = false;
    } else if (hipEventRecord(ev.first, st) != hipSuccess) {
      return -1;
    }
    pending.push_back({ev.first, ev.second, cls, by, launches, kernel, true});
    return (int)pending.size() - 1;
  }
  void end(hipStream_t st, int tok) {
    if (tok < 0) return;
    Pending& p = pending[tok];
    if (p.kernel) {
      KernelTiming& t = kernel_timing();
      p.valid = t.used;  // no launch took the events: nothing was timed
      t.armed = t.used = false;
    } else {
      (void)hipEventRecord(p.b, st);
    }
  }
  // Fold finished brackets into the per-class totals (stream already synchronised).
  void harvest() {
    for (auto& p : pending) {
      float t = 0.f;
      if (p.valid && hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
        ms[p.cls] += t;
        n[p.cls] += p.launches;
        bytes[p.cls] += p.bytes;
      }
      pool.push_back({p.a, p.b});
    }
    pending.clear();
  }
  void reset() {
    for (int c = 0; c < TDGL_K_COUNT; ++c) {
      ms[c] = 0.0;
      n[c] = 0;
      bytes[c] = 0.0;
    }
  }
  void destroy() {
    for (auto& e : pool) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    for (auto& e : pending) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    pool.clear();
    pending.clear();
  }
};

}  // namespace tdgl

using tdgl::ProfSink;

struct tdgl_ctx {
  tdgl_grid grid{};
  tdgl::Geo g{};
  tdgl::Bnd b{};
  tdgl::Mat mt{};
  tdgl::StencilCfg cfg{1};
  hipStream_t st = nullptr;
  std::string err;
  int device = 0;

  uint8_t* d_flags = nullptr;
  double* d_k2 = nullptr;
  double* d_sc = nullptr;
  double* d_bfull = nullptr;  // 3 * dim_x
  double* d_pb[2] = {nullptr, nullptr};  // double-buffered reduction partials
  int pb_cur = 0;                         // buffer holding the last published partials
  int pb_count = 0;                       // how many blocks published them
  double* d_res = nullptr;                // scalar slots (kResLen)
  int* d_bad = nullptr;
  double* h_res = nullptr;  // pinned mirror of the scalar slots
  // FP64 TGCR lookahead: per queued iteration (two in flight), the pinned
  // copy of res[RS_BRK .. RS_RR] and the event that completes it
  double* h_la = nullptr;
  hipEvent_t ev_la[2] = {nullptr, nullptr};

  long long len = 0;  // doubles per state buffer
  double* w_x = nullptr;
  double* w_gam = nullptr;
  double* w_f = nullptr;
  double* w_r = nullptr;
  std::vector<double*> P, AP;
  double** d_ptab = nullptr;
  int ptab_cap = 0;
  int ptab_len = 0;  // entries of d_ptab already copied

  // mixed precision: this film as a batch of one (created on first use)
  tdgl_batch* own_batch = nullptr;
  int kz_user = 0;    // tdgl_set_stencil_kz (0: the built-in choice)
  int tile_user = tdgl::kDefaultSweepShape;  // tdgl_set_sweep_tile (mixed-precision sweeps)
  double* d_acell = nullptr;  // adjoint: full grid x 5 cell adjoint
  double* d_vjpb = nullptr;   // adjoint: vjp_planes() x n scatter planes
  double* d_adjt[3] = {nullptr, nullptr, nullptr};  // adjoint solve: Chebyshev iterates
  double kappa_uniform = 0.0; // the kappa of the uniform link coefficients
  double k2max = 0.0;         // largest kappa^2 of any node (preconditioner bounds)
  int pb_blocks = 0;  // capacity of d_pb in blocks

  tdgl::ProfSink prof;  // tdgl_profile_enable
};

namespace tdgl {
// The mixed-precision step of one film and its FP32 JVP (tdgl_batch.hip).
int mixed_step_single(tdgl_ctx* ctx, double* X, double dt, const tdgl_solver_opts* o,
                      tdgl_step_stats* s);
int jvp32_single(tdgl_ctx* ctx, const float* x32, const float* p32, double dt, float* ap32);
// One application of the mixed TGCR's Chebyshev preconditioner (tdgl_batch.hip).
// Upper end b of the Chebyshev interval [1 - hdt, b] of I - hdt J (a
// Gershgorin bound of the stencil; the spectrum of the transpose is the same).
double cheb_upper_bound(const tdgl_ctx* c, double hdt);
int precond_single(tdgl_ctx* ctx, const float* x32, const float* v32, double dt, int degree,
                   float* w32, float* z32, double* info);
// Resolve the no-alloc scratch array's address on `device` (tdgl_noalloc.hip).
int noalloc_init_device(int device, std::string* err);
}  // namespace tdgl
