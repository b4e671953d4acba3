// Host-side launchers for the TDGL kernels (implemented in tdgl_f64.hip).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdgl_types.hpp"

namespace tdgl {

// Kernel timing for the event profiler.  While armed, the next launch made
// through launch() carries the two events of a profiling bracket
// (hipExtLaunchKernelGGL): they then time that kernel alone, as a kernel
// trace does, rather than the stream between two event records.
struct KernelTiming {
  hipEvent_t start = nullptr, stop = nullptr;
  bool armed = false, used = false;
};
KernelTiming& kernel_timing();  // the calling thread's

template <typename F, typename... Args>
inline void launch(F kern, dim3 gr, dim3 bl, unsigned int shmem, hipStream_t st, Args... args) {
  KernelTiming& t = kernel_timing();
  if (t.armed) {
    t.armed = false;
    t.used = true;
    hipExtLaunchKernelGGL(kern, gr, bl, shmem, st, t.start, t.stop, 0, args...);
  } else {
    kern<<<gr, bl, shmem, st>>>(args...);
  }
}

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
  RS_RW1 = 11,   // mixed: <r, Ap> after the first pass
  RS_RW2 = 12,   // mixed: <r, Ap> after the second pass
  RS_R0 = 13,    // ||r_0||_2 of the current TGCR solve (mixed; FP64 when right-preconditioned)
  RS_RK = 14,    // mixed: ||r_k||_2
  RS_GAMMA = 15, // mixed: coefficient of r_0 in the Newton update
  RS_BETA = 16,  // beta_j (mixed: first-pass coefficients; after the solve, the update's), j < kMaxKrylov
};
constexpr int kMaxKrylov = 2048;
// Tile shapes of the FP32 sweeps (k_bjvp32), nodes per plane: 0 = 64 x 4,
// 1 = 32 x 8, 2 = 16 x 16 (x kz planes).  Wide rows coalesce best on large
// films; squarer tiles waste fewer lanes on narrow ones (the batch autotuner
// picks).  Inside a solve, 64 x 4 was the fastest on single films (128x128x16:
// 15.5 us against 19.1 us per Chebyshev sweep for 32 x 8; equal on 256x256x64),
// although 32 x 8 wins when the sweep is timed alone on cache-resident vectors.
// FP32 sweep tile shapes: 0 = 64 x 4, 1 = 32 x 8, 2 = 16 x 16 nodes per plane;
// kSweepShapeFilm = each film a tile of its own (sweep_film_tile)
constexpr int kSweepShapeFilm = 3;
// The sweep piece lists give each XCD a run of near-equal estimated cost
// (padded with empty pieces) rather than an equal count of pieces.
constexpr bool kXcdBalance = true;
constexpr int kSweepShapes = 4;
constexpr int kDefaultSweepShape = 0;
__host__ __device__ constexpr int sweep_tx(int s) { return s == 1 ? 32 : (s == 2 ? 16 : 64); }
__host__ __device__ constexpr int sweep_ty(int s) { return s == 1 ? 8 : (s == 2 ? 16 : 4); }
// Staged cells a sweep block holds per plane: the 64 x 4 tile's (66 x 6).
constexpr int kSweepCells = 66 * 6;
// The tile of a film under kSweepShapeFilm: tw x th nodes per plane, any width
// from 16 to 64 (tw th <= 256 lanes, (tw + 2)(th + 2) staged cells within
// kSweepCells), the one with the least work per plane.  A sweep's time is
// close to proportional to its lanes (idle ones included: 17.8, 18.3 and
// 20.4 ps per lane and plane for 64 x 4, 32 x 8 and 16 x 16 tiles on the
// films of a 64-film batch), a little more per lane for narrow tiles, whose
// waves span more rows: a tile costs 23 + 64 / tw.  On a tie the wider tile.
// A film 67 nodes wide takes tiles 34 wide (2 across, 99 % of the columns
// busy) where 64 x 4 tiles keep 52 % busy.
inline void sweep_film_tile(long long nx, long long ny, int& tw, int& th) {
  double bc = -1.0;
  tw = 64;
  th = 4;
  for (int w = 64; w >= 16; --w)
    for (int h = 1; w * h <= 256; ++h) {
      if ((w + 2) * (h + 2) > kSweepCells) continue;
      const long long t = ((nx + w - 1) / w) * ((ny + h - 1) / h);
      const double c = (double)t * (23.0 + 64.0 / w);
      if (bc < 0.0 || c < bc) {
        bc = c;
        tw = w;
        th = h;
      }
    }
}
// Chebyshev polynomial preconditioner of the mixed TGCR: largest degree and
// the per-film coefficient stride (1/theta, then omega_1 .. omega_d).
constexpr int kMaxChebDeg = 64;
constexpr int kChebStride = kMaxChebDeg + 2;
constexpr int kChebDeg = kChebStride - 1;  // slot of the film's degree (as a double)
constexpr int RS_ALPHA = RS_BETA + kMaxKrylov;
constexpr int RS_DBETA = RS_ALPHA + kMaxKrylov;  // mixed: second-pass coefficients
// FP64 TGCR: 1.0 once the iteration loop has ended (converged, breakdown,
// iteration cap, non-finite residual); the kernels of a later, speculatively
// queued iteration then return at once.
constexpr int RS_STOP = RS_DBETA + kMaxKrylov;
constexpr int kResLen = RS_STOP + 1;

// Column-sum kinds of the mixed TGCR (k_bcolsum).
enum ColSum : int { CS_MD0 = 0, CS_MD1 = 1, CS_MD2 = 2, CS_UP1 = 3, CS_UP2 = 4, CS_DXINF = 5 };
// Per-film TGCR outcome (fstat[4 f]); fstat[4 f + 1] = basis size, [+2] = iterations, [+3] = cap.
enum TgState : int { TG_RUNNING = 0, TG_CONVERGED = 1, TG_BREAK = 2, TG_MAXITER = 3, TG_NONFINITE = 4 };

enum StencilMode : int { M_RHS = 0, M_PRED = 1, M_NEWTON = 2, M_JVP = 3, M_EULER = 4, M_NEWTON32 = 5 };

// Reduction plumbing of one launch: the partials it consumes in its prologue
// (published by the previous launch), where it publishes its own, and the
// scalar slots.
// Blocked modified Gram-Schmidt of the FP64 TGCR: vectors per pass, and the
// sums a pass publishes (kMgsBlock dots, then the Gram entries of its new
// vectors, G_lm at mgs_g(l, m) for l < m).
constexpr int kMgsBlock = 4;
constexpr int kMgsSums = kMgsBlock + kMgsBlock * (kMgsBlock - 1) / 2;
__host__ __device__ constexpr int mgs_g(int l, int m) { return kMgsBlock + m * (m - 1) / 2 + l; }
// Partials per block of the largest publisher (the blocked MGS pass).
constexpr int kPartialStride = kMgsSums;

struct Red {
  const double* pin;  // partials to consume, or nullptr
  int pcount;         // number of partial blocks in pin
  double* pout;       // this launch's partials (>= grid * kPartialStride doubles)
  double* res;        // kResLen doubles
};

// Old vectors (their betas known) and new vectors (dots wanted) of one
// blocked MGS pass.
struct MgsVecs {
  const double* a[kMgsBlock];
  const double* b[kMgsBlock];
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
  // JVP inside TGCR: eps = eps_num / sqrt(res[RS_RR]) on the device (the
  // host's formula, same rounding), and a return when res[RS_STOP] is set
  int tgcr;
  double eps_num;        // eps_mf * (1 + ||x||)
  // JVP with cz: one step of the FP64 TGCR's Chebyshev recurrence in place of
  // storing Ap: cz = com (p + (cr - A p) cthinv) - (com - 1) czpp (czpp = nullptr: 0)
  double* cz;
  const double* cr;
  const double* czpp;
  double com, cthinv;
};

// Kernel-side view of a batch of films (all device pointers).
struct BatchK {
  const FilmDesc* films;
  const int* item_film;  // FP64 stencil block -> film
  const int* sblk_film;  // streaming block -> film
  int* act;              // per film: takes part in the current stage
  double* res;           // nfilm * kResLen scalar slots
  const double* dtv;     // per film: dt, dt/2
  const double* cheb;    // per film, kChebStride doubles: 1/theta, omega_1 .. omega_d
  int nfilm;
  int nitems;            // FP64 stencil blocks of the batch
  int shape;             // FP32 sweep tile shape (sweep_tx / sweep_ty)
  int gen;               // some film has general applied-field vectors (per-cell ghost rules)
  int nsblk;             // streaming blocks of the batch
};

// Work of one FP32 sweep launch: block b runs the piece seg[b] = (film, tile
// column = tile_x + tiles32_x * tile_y, first plane, end plane); the launch
// has nblocks blocks.
struct SweepPart {
  const int4* seg;
  int nblocks;
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
// One blocked MGS pass: Ap -= sum_l beta_l a_l over `no` old vectors; sums of
// <Ap, b_m> and <b_l, b_m> over `nn` new ones, or ||Ap||^2 when nn = 0.
// first: the previous launch is the JVP (one partial: <Ap, a_0>).
int launch_mgs_block(const Geo& g, double* ap, const MgsVecs& v, int no, int nn, int beta_slot,
                     bool first, const Red& r, hipStream_t st);
// End of one FP64 TGCR iteration: ||r||^2 from the previous partials into
// res[RS_RR], and res[RS_STOP] = the loop's exit test for iteration k.
void launch_tgcr_check(double r0, double tol, long long k, long long maxk, const Red& r, hipStream_t st);
int mgs_blocks(long long len);  // blocks of a blocked MGS pass over vectors of len doubles
int launch_final_p(const Geo& g, const double* rv, const double* const* ptab, int nprev, double* ap,
                   double* pout, const Red& r, hipStream_t st, const double* zv = nullptr);
int launch_update_r(const Geo& g, double* rv, const double* ap, int alpha_slot, const Red& r,
                    hipStream_t st);
int launch_accum(const Geo& g, double* x, const double* const* ptab, int nvec, const Red& r,
                 hipStream_t st);
int launch_dot(const Geo& g, const double* a, const double* b, const Red& r, hipStream_t st);
int launch_norm_inf(const Geo& g, const double* a, const Red& r, hipStream_t st);
// One block: reduce r.pin into res[slot0 ...] (nsum sums then nmax maxima).
void launch_finalize(int nsum, int nmax, int slot0, const Red& r, hipStream_t st);
// Batched FP64 stencil (PRED or NEWTON32) over the batch's work items, the
// per-film reduction of its partials into res[slot0 ..], and the per-film
// report used by the host (tdgl_f64.hip).
void launch_bstencil(const BatchK& bk, bool is3d, int mode, const double* x, double* out0,
                     double* out1, const double* gam, float* o32a, float* o32b, double* pout,
                     hipStream_t st);
void launch_bfin(const BatchK& bk, int nsum, int nmax, const double* pin, int slot0, hipStream_t st);
void launch_breport(const BatchK& bk, const int* fstat, double* out, hipStream_t st);
constexpr int kReportLen = 10;  // doubles per film in the report
// Mixed-precision TGCR kernels (tdgl_f32.hip).
void launch_bjvp32(const BatchK& bk, const SweepPart& part, bool is3d, const float* x32, const float* p32,
                   float* const* Q, const int* iter, float* out_direct, hipStream_t st);
// Resident blocks per CU of the FP32 sweep kernel (shape, preconditioner epilogue or not).
int sweep_blocks_per_cu(bool is3d, int shape, bool cheb);
// One step of a Chebyshev recurrence for A = I - dt/2 J(x) (the stencil of
// k_bjvp32 with another epilogue).  With y = s p (s = 1/theta when src_th,
// else 1) and the film's omega_m, 1/theta:
//   res = omega_m (y + (cv v - A y) / theta) - (omega_m - 1) s' prev
//   out = fin ? v - res : res
// (prev == nullptr: no prev term; s' = 1/theta when prev_th).  out may alias
// prev; out == nullptr writes Q[*iter].
// One sweep m of a film's Chebyshev recurrence.  Each film runs its own
// degree d_f (the cheb table's slot kChebDeg): a film with d_f < m sits the
// sweep out; W sweeps (w = 1) write `mid` before the film's last sweep and
// v - rho into `out` (nullptr: Q[*iter]) at m = d_f; z sweeps write `mid`.
struct ChebStep {
  const float* v;
  const float* prev;
  float* out;
  float* mid;
  int m;
  int src_th, prev_th, cv, w;
};
void launch_bcheb32(const BatchK& bk, const SweepPart& part, bool is3d, const float* x32, const float* p32,
                    float* const* Q, const int* iter, const ChebStep& cs, hipStream_t st);
// dst = src on the films with sel[f] != 0 (each film's own elements)
void launch_bcopy_films(const BatchK& bk, const int* sel, const double* src, double* dst, hipStream_t st);
// out = x + ratio (x - xp): the extrapolating predictor (ratio[f] <= 0: film
// f keeps out as it is)
// one film's FP32 vector between the internal layout and the mixed method's
// (each node's link components together; tdgl_stage32.hpp)
void launch_l32_permute(const float* in, float* out, long long n, int nphi, bool to_l32, hipStream_t st);
void launch_bextrap(const BatchK& bk, const double* x, const double* xp, const double* ratio, double* out,
                    hipStream_t st);
void launch_extrap64(const double* x, const double* xp, double ratio, double* out, long long len,
                     hipStream_t st);
// x (FP64) += z_f with z_f the film's Chebyshev iterate: u / theta for degree
// 1, else t[d_f % 2]; partials of ||z_f||_inf.
void launch_baddz(const BatchK& bk, double* x, const float* u, const float* t0, const float* t1, double* pm,
                  hipStream_t st);
void launch_bmdot(const BatchK& bk, int pass, float* const* Q, const float* rv, const int* iter,
                  double* pm, hipStream_t st);
void launch_bupd(const BatchK& bk, int pass, float* const* Q, const float* rv, const int* iter,
                 double* pm, hipStream_t st);
void launch_bfinish(const BatchK& bk, float* const* Q, float* rv, const int* iter, double* H,
                    double* pm, hipStream_t st);
void launch_bcolsum(const BatchK& bk, const int* iter, const double* pm, int kind, int gy,
                    hipStream_t st);
void launch_bstatus(const BatchK& bk, int* fstat, const int* iter, double tol, const double* pm,
                    hipStream_t st);
void launch_bnext(const BatchK& bk, int* iter, int* nact, hipStream_t st);
void launch_btinit(const BatchK& bk, int* fstat, const int* maxk, int* iter, hipStream_t st);
void launch_btrisolve(const BatchK& bk, const int* fstat, const double* H, hipStream_t st);
void launch_baccum(const BatchK& bk, double* x, const float* r0v, float* const* Q, const int* fstat,
                   double* pm, hipStream_t st);
// u = gamma r_0 + sum_m coef_m Q_m in FP32 (the argument of the preconditioner
// in x += M u); u may alias r0v.
void launch_baccum_u(const BatchK& bk, float* u, const float* r0v, float* const* Q, const int* fstat,
                     hipStream_t st);
// Adjoint kernels (tdgl_adjoint.hip).  A: full grid x 5 cell adjoint.  B:
// vjp_planes() x n per-node planes, all written by launch_vjp_cells and turned
// into A by launch_cells_gather.  stop (non-null inside TGCR): res[RS_STOP];
// the kernels of an iteration queued after the loop ended return at once.
int vjp_planes();
void launch_vjp_cells(const Geo& g, const Bnd& b, const Mat& mt, const double* x, const double* w,
                      double* B, double* gsc, double kappa, double* pk, hipStream_t st,
                      const double* stop = nullptr);
void launch_cells_gather(const Geo& g, const double* B, double* A, hipStream_t st,
                         const double* stop = nullptr);
// One step of the adjoint TGCR's Chebyshev recurrence done where J^T zp is
// formed: z = om (zp + (r - (zp - hdt J^T zp)) thinv) - (om - 1) zpp (zpp = nullptr: 0).
struct AdjCheb {
  double* z = nullptr;  // nullptr: write J^T w itself
  const double* zp = nullptr;
  const double* r = nullptr;
  const double* zpp = nullptr;
  double om = 0.0, thinv = 0.0, hdt = 0.0;
};
// TGCR on the adjoint step operator: ap = r - hdt ap, partials of <ap, ap0> (or ||ap||^2).
int launch_adj_combine(const Geo& g, const double* r, double* ap, double hdt, const double* ap0,
                       const Red& rd, hipStream_t st);
void launch_scale(const Geo& g, double* y, double s, const double* x, hipStream_t st);
void launch_bc_gather(const Geo& g, const uint8_t* flags, const double* A, double* out, hipStream_t st,
                      const double* stop = nullptr, const AdjCheb& ch = AdjCheb{});
// launch_cells_gather + launch_bc_gather in one pass over the nodes (the cell
// adjoint is not stored; the same bits).
void launch_planes_gather(const Geo& g, const uint8_t* flags, const double* B, double* out, hipStream_t st,
                          const double* stop = nullptr, const AdjCheb& ch = AdjCheb{});
int launch_field_grad(const Geo& g, const uint8_t* flags, const double* A, double* pout, hipStream_t st);
void launch_bz_cells(const Geo& g, const Bnd& b, const double* gbz, double* A, hipStream_t st);
void launch_permute_in(const Geo& g, const void* ref, double* st_out, int* bad, hipStream_t st);
void launch_permute_out(const Geo& g, const double* st_in, void* ref, hipStream_t st);
void launch_observables(const Geo& g, const Bnd& b, const Mat& mt, const double* x, double* bout,
                        double* jout, hipStream_t st);

}  // namespace tdgl
