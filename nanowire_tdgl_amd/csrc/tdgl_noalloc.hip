// Entry points that allocate nothing, for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops).  Their scratch memory
// is one file-scope device array per GPU, resolved to a device address once
// per device when a context is created there.
//
// tdgl_euler_step stages X + dt f(X) in a work buffer the library allocates
// on the first call; tdgl_euler_step_noalloc stages it in `scratch` instead.
// Every call on one GPU shares that GPU's array: calls that overlap on a GPU
// (two streams, two host threads) race.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tdgl_ctx.hpp"

// 32 MiB per GPU: a state of up to 4 Mi doubles (838,860 interior nodes in 3-D).
constexpr long long kScratchDoubles = 4LL << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {nullptr};  // scratch's device address on each GPU

}  // namespace

namespace tdgl {

// Resolve scratch's address on `device` (the current device) once.  Called by
// tdgl_ctx_create, so every device a context lives on has it before any call.
int noalloc_init_device(int device, std::string* err) {
  if (device < 0 || device >= kMaxDevices) {
    if (err) *err = "device index beyond the no-alloc scratch table";
    return TDGL_E_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return TDGL_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    if (err) *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return TDGL_E_HIP;
  }
  g_ptr[device] = static_cast<double*>(p);
  return TDGL_OK;
}

}  // namespace tdgl

using namespace tdgl;

extern "C" {

int64_t tdgl_noalloc_max_len(void) { return kScratchDoubles; }

int tdgl_euler_step_noalloc(tdgl_ctx* ctx, double* X, double dt) {
  if (!ctx || !X) return TDGL_E_INVALID;
  if (ctx->len > kScratchDoubles) {
    ctx->err = "tdgl_euler_step_noalloc: the state (" + std::to_string(ctx->len) +
               " doubles) exceeds the scratch array (" + std::to_string(kScratchDoubles) +
               " doubles); use tdgl_euler_step";
    return TDGL_E_INVALID;
  }
  double* w;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    w = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_ptr[ctx->device] : nullptr;
  }
  if (!w) {
    ctx->err = "tdgl_euler_step_noalloc: scratch address not resolved on this device";
    return TDGL_E_INVALID;
  }
  // X + dt f(X) into scratch, then back into X: the stencil reads X's
  // neighbours, so it cannot write X in place (integrators.py:78-83)
  Epi e{};
  e.out0 = w;
  e.dt = dt;
  Red r{};
  r.pin = ctx->d_pb[ctx->pb_cur];
  r.pcount = ctx->pb_count;
  r.pout = ctx->d_pb[ctx->pb_cur ^ 1];
  r.res = ctx->d_res;
  launch_stencil(M_EULER, false, ctx->g, ctx->b, ctx->mt, X, nullptr, 0.0, e, ctx->cfg, r, ctx->st);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess)
    err = hipMemcpyAsync(X, w, sizeof(double) * (size_t)ctx->len, hipMemcpyDeviceToDevice, ctx->st);
  if (err != hipSuccess) {
    ctx->err = std::string("tdgl_euler_step_noalloc: ") + hipGetErrorString(err);
    return TDGL_E_HIP;
  }
  return TDGL_OK;
}

}  // extern "C"
