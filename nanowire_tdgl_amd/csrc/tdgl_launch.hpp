// Host-side launchers for the TDGL kernels (implemented in tdgl_f64.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdgl_types.hpp"

namespace tdgl {

// Slots of the per-context device scalar array.
enum ResSlot : int {
  RS_NRM = 0,    // ||Ap||_2 after Gram-Schmidt
  RS_INV = 1,    // 1/||Ap||_2
  RS_BRK = 2,    // 1.0 when ||Ap|| < 1e-14 (tgcr.py:141-142)
  RS_XX = 3,     // ||x_lin||_2^2   } the Newton residual kernel's three
  RS_RR = 4,     // ||r||_2^2       } results, in publish order
  RS_GINF = 5,   // ||g||_inf       }
  RS_DXINF = 6,  // ||dx||_inf
  RS_DOT = 7,    // generic dot / norm result
  RS_SELF = 8,   // mixed: ||Ap||^2 before Gram-Schmidt
  RS_NRM1 = 9,   // mixed: ||Ap||^2 after the first classical Gram-Schmidt pass
  RS_NRM2 = 10,  // mixed: ||Ap||^2 after the re-orthogonalisation pass
  RS_BETA = 16,  // beta_j (mixed: first-pass coefficients), j < kMaxKrylov
};
constexpr int kMaxKrylov = 2048;
constexpr int RS_ALPHA = RS_BETA + kMaxKrylov;
constexpr int RS_DBETA = RS_ALPHA + kMaxKrylov;  // mixed: second-pass coefficients
constexpr int kResLen = RS_DBETA + kMaxKrylov;

enum StencilMode : int { M_RHS = 0, M_PRED = 1, M_NEWTON = 2, M_JVP = 3, M_EULER = 4, M_NEWTON32 = 5 };

// Reduction plumbing of one launch: the partials it consumes in its prologue
// (published by the previous launch), where it publishes its own, and the
// scalar slots.
struct Red {
  const double* pin;  // partials to consume, or nullptr
  int pcount;         // number of partial blocks in pin
  double* pout;       // this launch's partials (>= grid * 4 doubles)
  double* res;        // kResLen doubles
};

// Extra inputs/outputs of the stencil epilogue.
struct Epi {
  double* out0;          // f | x_pred | f | Ap | X_new
  double* out1;          // - | gamma  | r | - | -
  const double* gam;     // NEWTON: gamma
  const double* f0;      // JVP: f(x_lin)
  const double* ap0;     // JVP: first Krylov vector (beta_0) or nullptr (norm)
  double dt;             // PRED/EULER: dt
  double hdt;            // dt / 2
  double inv_eps;        // JVP: 1/eps
  float* o32a;           // NEWTON32: x in FP32
  float* o32b;           // NEWTON32: r = -g in FP32
};

struct StencilCfg {
  int kz;                // z planes per block
};

int stencil_blocks(const Geo& g, const StencilCfg& c);
int stream_blocks(const Geo& g);

int launch_stencil(int mode, bool pert, const Geo& g, const Bnd& b, const Mat& mt, const double* x,
                   const double* p, double eps, const Epi& e, const StencilCfg& cfg, const Red& r,
                   hipStream_t st);

// Krylov / vector kernels (state-shaped vectors, internal layout).  Each
// returns the number of partials it publishes.
int launch_mgs_pass(const Geo& g, double* ap, const double* ap_prev, int beta_prev_slot,
                    const double* ap_next, const Red& r, hipStream_t st);
int launch_mgs_last(const Geo& g, double* ap, const double* ap_prev, int beta_slot, const Red& r,
                    hipStream_t st);
int launch_final_p(const Geo& g, const double* rv, const double* const* ptab, int nprev, double* ap,
                   double* pout, const Red& r, hipStream_t st);
int launch_update_r(const Geo& g, double* rv, const double* ap, int alpha_slot, const Red& r,
                    hipStream_t st);
int launch_accum(const Geo& g, double* x, const double* const* ptab, int nvec, const Red& r,
                 hipStream_t st);
int launch_dot(const Geo& g, const double* a, const double* b, const Red& r, hipStream_t st);
int launch_norm_inf(const Geo& g, const double* a, const Red& r, hipStream_t st);
// One block: reduce r.pin into res[slot0 ...] (nsum sums then nmax maxima).
void launch_finalize(int nsum, int nmax, int slot0, const Red& r, hipStream_t st);
// Mixed-precision kernels (tdgl_f32.hip).
int stream_blocks32(const Geo& g);
void launch_jvp32(const Geo& g, const Bnd& b, const Mat& mt, const float* x32, const float* p32,
                  float* ap, double hdt, const StencilCfg& cfg, hipStream_t st);
int launch_mdot32(const Geo& g, const float* ap, const float* const* V, int ncols, int with_self,
                  int second, const Red& r, hipStream_t st);
void launch_colsum(const double* pin, int count, int ncols_total, int n_coef, double* res, int base,
                   int self_slot, int second, hipStream_t st);
int launch_cgs_upd32(const Geo& g, float* ap, const float* const* V, int ncols, int coef_base,
                     int second, const Red& r, hipStream_t st);
int launch_final_p32(const Geo& g, const float* rv, const float* const* P, int nprev, float* ap,
                     float* pout, int nrm_slot, const Red& r, hipStream_t st);
int launch_update_r32(const Geo& g, float* rv, const float* ap, int aslot, const Red& r,
                      hipStream_t st);
int launch_accum32(const Geo& g, double* x, const float* const* P, int nvec, const Red& r,
                   hipStream_t st);
void launch_permute_in(const Geo& g, const void* ref, double* st_out, int* bad, hipStream_t st);
void launch_permute_out(const Geo& g, const double* st_in, void* ref, hipStream_t st);
void launch_observables(const Geo& g, const Bnd& b, const Mat& mt, const double* x, double* bout,
                        double* jout, hipStream_t st);

}  // namespace tdgl
