// Host driver of the mixed-precision trapezoidal step over a batch of films.
//
// Control flow per film is tdgl3d's (integrators.py:151-172, newton.py:76-120,
// tgcr.py:96-156): predictor, Newton iterations on the FP64 residual, each
// solving (I - dt/2 J) dx = -g with TGCR.  The films advance in lockstep: every
// launch covers all films still taking part, a film that finishes (TGCR
// converged, broke down or hit its cap; Newton converged or failed) drops out
// through its device-side `act` flag, and the host reads back one small
// per-film report per Newton iteration and one flag array per TGCR iteration.
//
// Differences from the FP64 path (the "mixed" method): the Jacobian-vector
// product is the analytic derivative in FP32 instead of the finite difference
// of tgcr.py:129-133, Krylov vectors are FP32 with FP64 accumulation,
// orthogonalisation is classical Gram-Schmidt with a conditional second pass,
// and the TGCR tolerance is max(tol_gcr, mixed_tol_floor).

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

struct tdgl_batch {
  std::vector<tdgl_ctx*> films;
  int nfilm = 0;
  bool is3d = true;
  hipStream_t st = nullptr;
  bool own_stream_from_ctx = false;  // a context's own batch follows the context's stream
  std::string err;

  std::vector<FilmDesc> hf;  // host copy of the descriptors
  std::vector<int> kz;       // z planes per stencil block, per film
  int shape = kDefaultSweepShape;  // FP32 sweep tile shape
  std::vector<int> item_film, sblk_film;
  int nitems = 0, nsblk = 0;
  long long total_len = 0;  // elements of a batched state vector
  long long hlen = 0;       // doubles of the packed TGCR matrices
  std::vector<int> kcap;    // per film: min(n_state, kMaxKrylov)

  FilmDesc* d_films = nullptr;
  int* d_item_film = nullptr;
  int* d_sblk_film = nullptr;
  int* d_act = nullptr;
  int* d_sel = nullptr;     // per film: selected for a batched copy (k_bcopy_films)
  double* d_res = nullptr;
  double* d_dtv = nullptr;
  double* d_ext = nullptr;  // per film: dt / dt_prev of the extrapolating predictor (0: off)
  int* d_fstat = nullptr;
  int* d_maxk = nullptr;
  int* d_iter = nullptr;
  int* d_nact = nullptr;
  double* d_pitem = nullptr;  // stencil partials, 3 per item
  double* d_pm = nullptr;     // column partials, (kMaxKrylov + 2) per streaming block
  double* d_H = nullptr;
  double* d_rep = nullptr;

  double* w_x = nullptr;
  double* w_gam = nullptr;
  float* x32 = nullptr;
  float* r32 = nullptr;
  float* r0_32 = nullptr;
  float* t32[2] = {nullptr, nullptr};  // Chebyshev recurrence vectors
  double* d_cheb = nullptr;            // nfilm * kChebStride: 1/theta, omega_1 .. omega_d
  int cheb_deg = 0;                    // this step's largest preconditioner degree (0: none)
  std::vector<int> cheb_degf;          // per film: its own degree
  std::vector<float*> Q;
  float** d_qtab = nullptr;  // kMaxKrylov entries
  std::vector<void*> retired;  // replaced device buffers, freed with the batch

  double* h_rep = nullptr;  // pinned: nfilm * kReportLen
  int* h_act = nullptr;     // pinned: 4 * nfilm (set_act staging, two readback slots, copy selection)

  // Work runs on a private stream (graph capture needs a non-legacy stream);
  // events order it after, and the caller's stream after it.
  hipStream_t ps = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_act[2] = {nullptr, nullptr};
  bool use_graph = true;
  hipGraphExec_t body = nullptr;  // one TGCR iteration (i >= 1) for the current layout
  double body_tol = -1.0;
  int body_deg = -1;
  int gen = 0;  // some film has general applied-field vectors (set by refresh)
  int body_gen = -1;

  // Work pieces of the FP32 sweeps (SweepPart): part[0] the plain JVP over
  // the films of the step, part[m] Chebyshev sweep m over the films whose
  // degree reaches m; rebuilt when that set or the tuning changes.
  std::vector<SweepPart> part;
  std::vector<int> part_key;
  int4* d_pseg = nullptr;
  size_t pseg_cap = 0;
  int part_gen = 0, body_part = -1;
  int part_lmax = 8;  // deepest sweep piece, in planes (the autotuner's knob)
  int part_lmin = 2;  // shallowest sweep piece when a sweep has little work (the autotuner's knob)
  int cus = 0;        // compute units of the device

  ProfSink own_prof;
  ProfSink* prof = &own_prof;
};

namespace {

int bfail(tdgl_batch* B, int code, const std::string& msg) {
  if (B) B->err = msg;
  return code;
}

#define BCHK(call)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return bfail(B, TDGL_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

// Device buffers are cleared and filled on the batch's own stream, and buffers
// a batch replaces are kept until it is destroyed: no call of a step touches
// the legacy stream or frees memory (which would synchronise the device), so
// batches stepped from several host threads at once do not disturb each
// other's work or graph captures.
int balloc(tdgl_batch* B, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return bfail(B, TDGL_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemsetAsync(*p, 0, bytes, B->ps);
  if (e != hipSuccess) return bfail(B, TDGL_E_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  return TDGL_OK;
}
void retire(tdgl_batch* B, void* p) {
  if (p) B->retired.push_back(p);
}

constexpr int kStencilTX = 64, kStencilTY = 4;

// The FP32 sweeps address a film's vector with 32-bit byte offsets from one
// base (20 bytes per node).
bool fits_sweeps(const tdgl_ctx* c) { return c->g.n * 20LL < (1LL << 32); }
constexpr long long kStreamChunk = 4096;  // elements per streaming block (256 threads x 4 float4)
constexpr int kStreamMaxPerFilm = 2048;

BatchK kview(const tdgl_batch* B) {
  BatchK k;
  k.films = B->d_films;
  k.item_film = B->d_item_film;
  k.sblk_film = B->d_sblk_film;
  k.act = B->d_act;
  k.res = B->d_res;
  k.dtv = B->d_dtv;
  k.cheb = B->d_cheb;
  k.nfilm = B->nfilm;
  k.nitems = B->nitems;
  k.shape = B->shape;
  k.gen = B->gen;
  k.nsblk = B->nsblk;
  return k;
}

// Order the private stream after the caller's stream (entry) and the caller's
// stream after the private one (exit).
int sync_in(tdgl_batch* B) {
  BCHK(hipEventRecord(B->ev_in, B->st));
  BCHK(hipStreamWaitEvent(B->ps, B->ev_in, 0));
  return TDGL_OK;
}
int sync_out(tdgl_batch* B) {
  BCHK(hipEventRecord(B->ev_out, B->ps));
  BCHK(hipStreamWaitEvent(B->st, B->ev_out, 0));
  return TDGL_OK;
}

// z planes per stencil work item: deep enough that the two extra staged planes
// are a small overhead (measured best at 8 on 256x256x64 films and on the
// 64-film batch; 16 and 4 are 10 % slower there inside a solve), shallow
// enough that the work still spreads over the CUs: about 512 work items over
// the whole launch (2 on a 128x128x16 film alone).  A batch's launches cover
// all its films, so the rule applies to the batch's tile-planes together
// (8 on the 64-film batch, 2.0 staged cells per node instead of 2.8 with each
// film sized alone; attempts 1.6 % faster).
long long stencil_tile_planes(const Geo& g) {
  const long long tiles = (long long)((g.nx + kStencilTX - 1) / kStencilTX) * ((g.ny + kStencilTY - 1) / kStencilTY);
  return tiles * (g.is3d ? g.nz : 1);
}
int default_kz(const Geo& g, long long launch_tile_planes) {
  if (!g.is3d) return 1;
  long long kz = launch_tile_planes / 512;
  kz = std::max(2LL, std::min<long long>(kz, 8));
  return (int)std::min<long long>(kz, g.nz);
}

// Work decomposition and buffer offsets.  Called at creation and whenever kz changes.
int build_layout(tdgl_batch* B) {
  const int F = B->nfilm;
  B->hf.assign(F, FilmDesc{});
  B->item_film.clear();
  B->sblk_film.clear();
  long long off = 0, hoff = 0;
  for (int f = 0; f < F; ++f) {
    const tdgl_ctx* c = B->films[f];
    FilmDesc& d = B->hf[f];
    d.g = c->g;
    d.b = c->b;
    d.mt = c->mt;
    d.len = c->len;
    d.off = off;
    off += (c->len + 63) / 64 * 64;
    d.kz = B->kz[f];
    d.tiles_x = (c->g.nx + kStencilTX - 1) / kStencilTX;
    d.tiles_y = (c->g.ny + kStencilTY - 1) / kStencilTY;
    d.tz = c->g.is3d ? (c->g.nz + d.kz - 1) / d.kz : 1;
    d.item0 = (int)B->item_film.size();
    d.nitems = d.tiles_x * d.tiles_y * d.tz;
    B->item_film.insert(B->item_film.end(), d.nitems, f);
    if (B->shape == kSweepShapeFilm) {
      sweep_film_tile(c->g.nx, c->g.ny, d.tw32, d.th32);
    } else {
      d.tw32 = sweep_tx(B->shape);
      d.th32 = sweep_ty(B->shape);
    }
    d.tiles32_x = (c->g.nx + d.tw32 - 1) / d.tw32;
    d.tiles32_y = (c->g.ny + d.th32 - 1) / d.th32;
    d.sb0 = (int)B->sblk_film.size();
    d.nsb = (int)std::min<long long>(kStreamMaxPerFilm, std::max<long long>(1, (c->len + kStreamChunk - 1) / kStreamChunk));
    B->sblk_film.insert(B->sblk_film.end(), d.nsb, f);
    d.hoff = hoff;
    const long long K = B->kcap[f];
    hoff += K * (K + 1) / 2;
  }
  B->total_len = off;
  B->hlen = hoff;
  if (B->body) {
    (void)hipGraphExecDestroy(B->body);
    B->body = nullptr;
  }
  B->nitems = (int)B->item_film.size();
  B->nsblk = (int)B->sblk_film.size();
  B->part_key.clear();  // the sweep tiles may have changed
  // build_layout runs at creation and from the configuration calls (tile
  // shape, z depth), which wait for the batch's stream first: never inside a
  // step, so the old tables are freed at once
  for (void* p : {(void*)B->d_item_film, (void*)B->d_sblk_film, (void*)B->d_pitem, (void*)B->d_pm})
    if (p) BCHK(hipFree(p));
  B->d_item_film = B->d_sblk_film = nullptr;
  B->d_pitem = B->d_pm = nullptr;
  int rc;
  if ((rc = balloc(B, (void**)&B->d_item_film, sizeof(int) * B->nitems))) return rc;
  if ((rc = balloc(B, (void**)&B->d_sblk_film, sizeof(int) * B->nsblk))) return rc;
  if ((rc = balloc(B, (void**)&B->d_pitem, sizeof(double) * 3 * (size_t)B->nitems))) return rc;
  if ((rc = balloc(B, (void**)&B->d_pm, sizeof(double) * (kMaxKrylov + 2) * (size_t)B->nsblk))) return rc;
  BCHK(hipMemcpyAsync(B->d_item_film, B->item_film.data(), sizeof(int) * B->nitems, hipMemcpyHostToDevice,
                      B->ps));
  BCHK(hipMemcpyAsync(B->d_sblk_film, B->sblk_film.data(), sizeof(int) * B->nsblk, hipMemcpyHostToDevice,
                      B->ps));
  BCHK(hipStreamSynchronize(B->ps));  // the host tables may change before the next copy
  return TDGL_OK;
}

// Current operators of the films (fields and materials may have changed).
int refresh(tdgl_batch* B) {
  B->gen = 0;
  for (int f = 0; f < B->nfilm; ++f) {
    B->hf[f].b = B->films[f]->b;
    B->hf[f].mt = B->films[f]->mt;
    if (B->hf[f].b.bz) B->gen = 1;
  }
  BCHK(hipMemcpyAsync(B->d_films, B->hf.data(), sizeof(FilmDesc) * B->nfilm, hipMemcpyHostToDevice, B->ps));
  return TDGL_OK;
}

int ensure_work(tdgl_batch* B) {
  if (B->w_x) return TDGL_OK;
  const size_t n = (size_t)B->total_len;
  int rc;
  if ((rc = balloc(B, (void**)&B->w_x, sizeof(double) * n))) return rc;
  if ((rc = balloc(B, (void**)&B->w_gam, sizeof(double) * n))) return rc;
  if ((rc = balloc(B, (void**)&B->x32, sizeof(float) * n))) return rc;
  if ((rc = balloc(B, (void**)&B->r32, sizeof(float) * n))) return rc;
  if ((rc = balloc(B, (void**)&B->r0_32, sizeof(float) * n))) return rc;
  if ((rc = balloc(B, (void**)&B->d_qtab, sizeof(float*) * kMaxKrylov))) return rc;
  if ((rc = balloc(B, (void**)&B->d_H, sizeof(double) * (size_t)std::max(1LL, B->hlen)))) return rc;
  return TDGL_OK;
}

int ensure_cheb(tdgl_batch* B) {
  const size_t n = (size_t)B->total_len;
  for (float*& t : B->t32) {
    if (t) continue;
    int rc;
    if ((rc = balloc(B, (void**)&t, sizeof(float) * n))) return rc;
  }
  return TDGL_OK;
}

// Krylov vector Q_i (zero padding) and its table entry.
int ensure_q(tdgl_batch* B, int i) {
  if (i >= kMaxKrylov) return bfail(B, TDGL_E_INVALID, "TGCR needs more than 2048 Krylov vectors");
  while ((int)B->Q.size() <= i) {
    float* q = nullptr;
    int rc;
    if ((rc = balloc(B, (void**)&q, sizeof(float) * (size_t)B->total_len))) return rc;
    B->Q.push_back(q);
    const int k = (int)B->Q.size() - 1;
    BCHK(hipMemcpyAsync(B->d_qtab + k, &B->Q[k], sizeof(float*), hipMemcpyHostToDevice, B->ps));
    BCHK(hipStreamSynchronize(B->ps));  // B->Q may move on its next push_back
  }
  return TDGL_OK;
}

int set_act(tdgl_batch* B, const std::vector<int>& act) {
  memcpy(B->h_act, act.data(), sizeof(int) * B->nfilm);
  BCHK(hipMemcpyAsync(B->d_act, B->h_act, sizeof(int) * B->nfilm, hipMemcpyHostToDevice, B->ps));
  return TDGL_OK;
}

// dst[film] = src[film] for the films in `sel`, in one launch (a device copy
// per film cost a launch each).  The selection's pinned staging slot is
// rewritten only after a later report() has synchronised the stream.
int copy_films(tdgl_batch* B, const std::vector<int>& sel, const double* src, double* dst) {
  int* h = B->h_act + 3 * (size_t)B->nfilm;
  memcpy(h, sel.data(), sizeof(int) * B->nfilm);
  BCHK(hipMemcpyAsync(B->d_sel, h, sizeof(int) * B->nfilm, hipMemcpyHostToDevice, B->ps));
  launch_bcopy_films(kview(B), B->d_sel, src, dst, B->ps);
  BCHK(hipGetLastError());
  return TDGL_OK;
}

// One sweep's pieces: every FP32 tile column of the films in `in`, cut along z
// into near-equal pieces of at most Lp planes, one block per piece.  Lp adapts
// to the sweep's work W (tile-planes): W / (the blocks the device holds at
// once), within [lmin, lmax] -- a sweep over every film gets deep pieces (few
// restaged halo planes), a sweep over a few films shallow ones (every CU
// busy).  The pieces are listed film by film (within a film, neighbouring
// tiles of one z piece next to each other: their shared halo cells meet in
// L2), and the sweep kernel gives each XCD one contiguous run of the list
// (xcd_remap); inside each run the deeper pieces come first (by depth class,
// two planes wide), so each XCD's dispatcher starts its longest blocks first.
void part_build(const tdgl_batch* B, const std::vector<char>& in, long long gres, int lmin, int lmax,
                std::vector<int4>& seg, int& nblocks) {
  long long W = 0;
  for (int f = 0; f < B->nfilm; ++f)
    if (in[f]) W += (long long)B->hf[f].tiles32_x * B->hf[f].tiles32_y * (B->is3d ? B->hf[f].g.nz : 1);
  nblocks = 0;
  if (W == 0) return;
  const long long Lp = std::max<long long>(lmin, std::min<long long>(lmax, (W + gres - 1) / gres));
  const size_t seg0 = seg.size();
  for (int f = 0; f < B->nfilm; ++f) {
    if (!in[f]) continue;
    const FilmDesc& d = B->hf[f];
    const int cols = d.tiles32_x * d.tiles32_y;
    if (!B->is3d) {
      for (int t = 0; t < cols; ++t) seg.push_back(make_int4(f, t, 0, 1));
      continue;
    }
    const int nz = d.g.nz, np = (int)((nz + Lp - 1) / Lp);
    for (int q = 0; q < np; ++q) {
      const int k0 = 1 + (int)((long long)nz * q / np), k1 = 1 + (int)((long long)nz * (q + 1) / np);
      for (int t = 0; t < cols; ++t) seg.push_back(make_int4(f, t, k0, k1));
    }
  }
  nblocks = (int)(seg.size() - seg0);
  const int nx = 8;
  if (kXcdBalance && B->is3d && nblocks >= 8 * nx) {
    // Equal block counts per XCD give unequal work (films differ in piece
    // depth and tile width), and the sweep lasts as long as its busiest XCD.
    // So the list is cut into 8 contiguous runs of near-equal estimated cost,
    // a block's set-up counted as ~2.3 planes and a plane's cost as its tile's
    // lanes per node (23 + 64/W, the per-film tile model); each run is padded
    // to the longest with empty pieces (film -1: the block returns at once).
    std::vector<double> cost((size_t)nblocks);
    double total = 0.0;
    for (int q = 0; q < nblocks; ++q) {
      const int4 v = seg[seg0 + q];
      const int tw = B->shape == kSweepShapeFilm ? B->hf[v.x].tw32 : sweep_tx(B->shape);
      cost[q] = ((v.w - v.z) + 2.3) * (23.0 + 64.0 / tw);
      total += cost[q];
    }
    std::vector<int> cut(nx + 1, nblocks);
    cut[0] = 0;
    double acc = 0.0;
    for (int q = 0, x = 1; q < nblocks && x < nx; ++q) {
      acc += cost[q];
      if (acc >= total * x / nx) cut[x++] = q + 1;
    }
    int per = 0;
    for (int x = 0; x < nx; ++x) per = std::max(per, cut[x + 1] - cut[x]);
    std::vector<int4> runs;
    runs.reserve((size_t)per * nx);
    for (int x = 0; x < nx; ++x) {
      for (int q = cut[x]; q < cut[x + 1]; ++q) runs.push_back(seg[seg0 + q]);
      for (int q = cut[x + 1] - cut[x]; q < per; ++q) runs.push_back(make_int4(-1, 0, 0, 0));
    }
    seg.resize(seg0);
    seg.insert(seg.end(), runs.begin(), runs.end());
    nblocks = per * nx;
  }
  // the runs of xcd_remap (tdgl_device.hpp): XCD x holds per or per - 1 blocks
  const int G = nblocks, per = (G + nx - 1) / nx, full = G % nx == 0 ? nx : G % nx;
  for (int x = 0, base = 0; x < nx; ++x) {
    const int len = x < full ? per : per - 1;
    // by depth class (pieces within a plane or so of each other keep their
    // order: a column's z pieces differ by one plane and share halo planes);
    // empty pieces last
    std::stable_sort(seg.begin() + seg0 + base, seg.begin() + seg0 + base + len,
                     [](const int4& a, const int4& b) {
                       const int ca = a.x < 0 ? -1 : (a.w - a.z + 1) / 2, cb = b.x < 0 ? -1 : (b.w - b.z + 1) / 2;
                       return ca > cb;
                     });
    base += len;
  }
}

// The sweep pieces for the films in `act` at the current Chebyshev degrees
// (rebuilt only when the films, degrees, tile shape or tuning change).
int ensure_parts(tdgl_batch* B, const std::vector<int>& act) {
  const int F = B->nfilm, D = B->cheb_deg;
  std::vector<int> key = {B->shape, B->part_lmax, B->part_lmin, D};
  for (int f = 0; f < F; ++f) key.push_back(act[f] ? 1 + (D ? B->cheb_degf[f] : 0) : 0);
  if (key == B->part_key) return TDGL_OK;
  const long long gres = (long long)B->cus * sweep_blocks_per_cu(B->is3d, B->shape, D > 0);
  std::vector<int4> seg;
  std::vector<int> g0(D + 1), nb(D + 1);
  for (int m = 0; m <= D; ++m) {
    std::vector<char> in(F);
    for (int f = 0; f < F; ++f) in[f] = act[f] && (m == 0 || B->cheb_degf[f] >= m);
    g0[m] = (int)seg.size();
    part_build(B, in, gres, B->part_lmin, B->part_lmax, seg, nb[m]);
    // every tile-plane of every film taking part exactly once, every piece
    // inside its film's tiles and planes
    std::vector<long long> cover(F, 0);
    bool ok = (int)seg.size() - g0[m] == nb[m];
    for (int q = g0[m]; ok && q < (int)seg.size(); ++q) {
      const int4 v = seg[q];
      if (v.x == -1 && v.y == 0 && v.z == 0 && v.w == 0) continue;  // an empty piece
      ok = v.x >= 0 && v.x < F && in[v.x];
      if (!ok) break;
      const FilmDesc& d = B->hf[v.x];
      const int k0 = B->is3d ? 1 : 0, k1 = B->is3d ? d.g.Nz : 1;
      ok = v.y >= 0 && v.y < d.tiles32_x * d.tiles32_y && v.z >= k0 && v.z < v.w && v.w <= k1;
      cover[v.x] += v.w - v.z;
    }
    for (int f = 0; ok && f < F; ++f)
      ok = cover[f] == (in[f] ? (long long)B->hf[f].tiles32_x * B->hf[f].tiles32_y *
                                    (B->is3d ? B->hf[f].g.nz : 1) : 0LL);
    if (!ok) return bfail(B, TDGL_E_INVALID, "internal error: inconsistent sweep pieces");
  }
  BCHK(hipStreamSynchronize(B->ps));  // no queued sweep still reads the old tables
  if (seg.size() > B->pseg_cap) {
    retire(B, B->d_pseg);
    B->d_pseg = nullptr;
    int rc;
    if ((rc = balloc(B, (void**)&B->d_pseg, sizeof(int4) * seg.size()))) return rc;
    B->pseg_cap = seg.size();
  }
  if (!seg.empty()) {
    BCHK(hipMemcpyAsync(B->d_pseg, seg.data(), sizeof(int4) * seg.size(), hipMemcpyHostToDevice, B->ps));
    BCHK(hipStreamSynchronize(B->ps));  // seg is a local
  }
  B->part.assign(D + 1, SweepPart{});
  for (int m = 0; m <= D; ++m) B->part[m] = SweepPart{B->d_pseg + g0[m], nb[m]};
  B->part_key = key;
  ++B->part_gen;
  return TDGL_OK;
}

int report(tdgl_batch* B) {
  launch_breport(kview(B), B->d_fstat, B->d_rep, B->ps);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(B->h_rep, B->d_rep, sizeof(double) * kReportLen * B->nfilm, hipMemcpyDeviceToHost, B->ps));
  BCHK(hipStreamSynchronize(B->ps));
  if (!B->prof->pending.empty()) B->prof->harvest();
  return TDGL_OK;
}
const double* rep(const tdgl_batch* B, int f) { return B->h_rep + (size_t)f * kReportLen; }

double s32(const tdgl_batch* B, int f) { return 4.0 * (double)B->hf[f].len; }
double s64(const tdgl_batch* B, int f) { return 8.0 * (double)B->hf[f].len; }
double sum_over(const tdgl_batch* B, const std::vector<int>& act, double (*w)(const tdgl_batch*, int)) {
  double s = 0.0;
  for (int f = 0; f < B->nfilm; ++f)
    if (act[f]) s += w(B, f);
  return s;
}

// Upper end of the spectrum of A = I - dt/2 J: 1 + dt/2 ||J||_inf, the row
// sums of |J| read off k_bjvp32's coefficients (real representation, complex
// rotations counted as sqrt(2)) with |psi| <= psimax.  psi rows: Laplacian
// diagonal and neighbours, the link couplings, (1 - |psi|^2) psi and the tau
// relaxation; link rows: 2 (4 c_a + 2 c_b + 2 c_c) (the doubled along-link
// second difference and the transverse ones) + 4 (c_b + c_c) (curl cross
// terms) = 8 (c_x + c_y + c_z), plus the supercurrent.
double cheb_upper(const tdgl_ctx* c, double hdt) {
  const double psimax = 2.0, r2 = sqrt(2.0);
  const Mat& m = c->mt;
  const bool d3 = c->g.is3d != 0;
  const double ih = m.ihx2 + m.ihy2 + (d3 ? m.ihz2 : 0.0);
  const double k2 = c->k2max;
  const double cs = k2 * (m.ihx2 + m.ihy2 + (d3 ? m.ihz2 : 0.0));
  const double row_psi = 2.0 * ih + 2.0 * r2 * (1.0 + psimax) * ih + 1.0 + 4.0 * psimax * psimax +
                         (m.has_mat ? m.inv_tau : 0.0);
  const double row_phi = 8.0 * cs + 4.0 * r2 * psimax + psimax * psimax;
  return 1.0 + hdt * std::max(row_psi, row_phi);
}

// Chebyshev coefficients of this step for every film: interval [a, b] with
// a = 1 - dt/2 (J's positive eigenvalues stay below 1: the (1 - |psi|^2) psi
// term at psi = 0) and b = cheb_upper; sigma = (b + a)/(b - a),
// omega_1 = 1, omega_2 = 2 sigma^2 / (2 sigma^2 - 1),
// omega_{m+1} = 1 / (1 - omega_m / (4 sigma^2)).  Each film runs its own
// degree d_f (table slot kChebDeg): the request, or the degree that brings
// its polynomial's deviation 1 / T_d(sigma) down to 0.25; the sweeps of a
// step run to the largest, films past their degree sit them out.
int setup_cheb(tdgl_batch* B, int request, const std::vector<double>& dtv,
               const std::vector<int>& act, double* interval0 = nullptr) {
  B->cheb_deg = 0;
  if (request == 0) return TDGL_OK;
  const int F = B->nfilm;
  std::vector<double> lo(F), hi(F);
  std::vector<int> deg(F, 1);
  int dmax = 1;
  for (int f = 0; f < F; ++f) {
    const double hdt = dtv[2 * f + 1];
    lo[f] = 1.0 - hdt;
    hi[f] = cheb_upper(B->films[f], hdt);
    if (act[f] && lo[f] < 0.1) return TDGL_OK;  // dt too large for a positive interval: no preconditioner
    int d = 1;
    if (request > 0) {
      d = request;
    } else if (lo[f] < hi[f]) {
      const double sig = (hi[f] + lo[f]) / (hi[f] - lo[f]);
      d = (int)ceil(acosh(4.0) / acosh(sig));
    }
    deg[f] = std::max(1, std::min(d, kMaxChebDeg));
    if (act[f]) dmax = std::max(dmax, deg[f]);
  }
  std::vector<double> tab((size_t)F * kChebStride, 0.0);
  for (int f = 0; f < F; ++f) {
    double* t = tab.data() + (size_t)f * kChebStride;
    t[kChebDeg] = (double)deg[f];
    if (!(lo[f] < hi[f])) continue;
    const double th = 0.5 * (hi[f] + lo[f]), de = 0.5 * (hi[f] - lo[f]);
    const double sig = th / de, s2 = sig * sig;
    t[0] = 1.0 / th;
    t[1] = 1.0;
    if (deg[f] >= 2) t[2] = 2.0 * s2 / (2.0 * s2 - 1.0);
    for (int mm = 3; mm <= deg[f]; ++mm) t[mm] = 1.0 / (1.0 - t[mm - 1] / (4.0 * s2));
  }
  int rc;
  if ((rc = ensure_cheb(B))) return rc;
  BCHK(hipMemcpyAsync(B->d_cheb, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, B->ps));
  BCHK(hipStreamSynchronize(B->ps));  // tab is a host temporary
  B->cheb_deg = dmax;
  B->cheb_degf = deg;
  if (interval0) {
    interval0[0] = lo[0];
    interval0[1] = hi[0];
  }
  return TDGL_OK;
}

// W = A M v into `out` (nullptr: Q[*iter]) at the linearisation point x32:
// the plain JVP (no preconditioner), or the d sweeps of the residual
// recurrence rho_m = P_m(A) v, P_m(t) = T_m((theta - t)/delta) / T_m(sigma):
// rho_1 = v - A v / theta, rho_m = omega_m (rho_{m-1} - A rho_{m-1} / theta)
// - (omega_m - 1) rho_{m-2}, W = v - rho_d (= A z_d, z_d the Chebyshev
// iterate for A z = v; M = (I - P_d(A)) A^{-1}).  rho_m lives in
// t32[(m - 1) % 2]; each sweep overwrites its rho_{m-2} in place.
void enqueue_w(tdgl_batch* B, bool prof, const std::vector<int>& act, const float* x32,
               const float* v, float* out) {
  const BatchK bk = kview(B);
  const int d = B->cheb_deg;
  if (d == 0) {
    const double S = prof ? sum_over(B, act, s32) : 0.0;
    const int tok = prof ? B->prof->begin(B->ps, TDGL_K_JVP, 3.0 * S, 1, true) : -1;
    launch_bjvp32(bk, B->part[0], B->is3d, x32, v, B->d_qtab, B->d_iter, out, B->ps);
    if (prof) B->prof->end(B->ps, tok);
    return;
  }
  for (int m = 1; m <= d; ++m) {
    ChebStep cs{};
    cs.m = m;
    cs.w = 1;
    cs.v = v;
    cs.out = out;
    cs.mid = B->t32[(m - 1) % 2];
    cs.prev = m == 1 ? nullptr : (m == 2 ? v : B->t32[(m - 1) % 2]);
    const float* src = m == 1 ? v : B->t32[(m - 2) % 2];
    // bytes per film still sweeping: x, source, output; prev (m >= 2); v at
    // its last sweep unless it is prev (m >= 3)
    double by = 0.0;
    if (prof)
      for (int f = 0; f < B->nfilm; ++f)
        if (act[f] && B->cheb_degf[f] >= m)
          by += s32(B, f) * (3.0 + (m >= 2 ? 1.0 : 0.0) + (m == B->cheb_degf[f] && m >= 3 ? 1.0 : 0.0));
    const int tok = prof ? B->prof->begin(B->ps, TDGL_K_CHEB, by, 1, true) : -1;
    launch_bcheb32(bk, B->part[m], B->is3d, x32, src, B->d_qtab, B->d_iter, cs, B->ps);
    if (prof) B->prof->end(B->ps, tok);
  }
}

// z = M u: the Chebyshev iterate z_{d_f} of A z = u (z_1 = u / theta,
// z_m = omega_m (z_{m-1} + (u - A z_{m-1}) / theta) - (omega_m - 1) z_{m-2}).
// z_m lives in t32[m % 2], so film f's result is in t32[d_f % 2] (d_f >= 2) or
// is u / theta (d_f = 1, applied by the consumer).
void enqueue_z(tdgl_batch* B, bool prof, const std::vector<int>& act, const float* x32, const float* u) {
  const int d = B->cheb_deg;
  if (d < 2) return;
  const BatchK bk = kview(B);
  for (int m = 2; m <= d; ++m) {
    ChebStep cs{};
    cs.m = m;
    cs.cv = 1;
    cs.v = u;
    const float* src = m == 2 ? u : B->t32[(m - 1) % 2];
    cs.src_th = m == 2;
    cs.prev = m == 2 ? nullptr : (m == 3 ? u : B->t32[m % 2]);
    cs.prev_th = m == 3;
    cs.mid = B->t32[m % 2];
    // bytes per film still sweeping: x, source, output; u unless it is the
    // source (m >= 3); prev unless it is u (m >= 4)
    double by = 0.0;
    if (prof)
      for (int f = 0; f < B->nfilm; ++f)
        if (act[f] && B->cheb_degf[f] >= m) by += s32(B, f) * (3.0 + (m >= 3 ? 1.0 : 0.0) + (m >= 4 ? 1.0 : 0.0));
    const int tok = prof ? B->prof->begin(B->ps, TDGL_K_CHEB, by, 1, true) : -1;
    launch_bcheb32(bk, B->part[m], B->is3d, x32, src, B->d_qtab, B->d_iter, cs, B->ps);
    if (prof) B->prof->end(B->ps, tok);
  }
}

// x += M u for the films in `ok`, u = R y in r0_32, with ||M u||_inf partials.
void enqueue_precond_update(tdgl_batch* B, bool prof, const std::vector<int>& ok) {
  const BatchK bk = kview(B);
  enqueue_z(B, prof, ok, B->x32, B->r0_32);
  const double S = prof ? sum_over(B, ok, s32) : 0.0;
  const int tok = prof ? B->prof->begin(B->ps, TDGL_K_ACCUM, S + 2.0 * sum_over(B, ok, s64), 2) : -1;
  launch_baddz(bk, B->w_x, B->r0_32, B->t32[0], B->t32[1], B->d_pm, B->ps);
  launch_bcolsum(bk, B->d_iter, B->d_pm, CS_DXINF, 1, B->ps);
  if (prof) B->prof->end(B->ps, tok);
}

// Newton residual in FP64 at w_x for the active films: x32, r32 = -g, and
// res[RS_XX, RS_RR, RS_GINF].
int newton_eval(tdgl_batch* B, const std::vector<int>& act, std::vector<tdgl_step_stats>& S) {
  const double by = 2.0 * sum_over(B, act, s64) + 2.0 * sum_over(B, act, s32);
  const int tok = B->prof->begin(B->ps, TDGL_K_NEWTON, by);
  launch_bstencil(kview(B), B->is3d, M_NEWTON32, B->w_x, nullptr, nullptr, B->w_gam, B->x32, B->r32,
                  B->d_pitem, B->ps);
  launch_bfin(kview(B), 2, 1, B->d_pitem, RS_XX, B->ps);
  B->prof->end(B->ps, tok);
  BCHK(hipGetLastError());
  for (int f = 0; f < B->nfilm; ++f)
    if (act[f]) {
      S[f].rhs_evals++;
      S[f].bytes_moved += 2.0 * s64(B, f) + 2.0 * s32(B, f);
    }
  return TDGL_OK;
}

constexpr int kColsumY = 32;  // column blocks per film in the coefficient sums

// The kernels of TGCR iteration i >= 1 (every argument is iteration-independent:
// the kernels read i from the device counter).  With `prof`, each pass is
// bracketed with its algorithmic bytes for iteration i and the host's view
// `act` of the films taking part.
void enqueue_body(tdgl_batch* B, double tol, bool prof, int i, const std::vector<int>& act) {
  const BatchK bk = kview(B);
  const double S3 = prof ? sum_over(B, act, s32) : 0.0;
  enqueue_w(B, prof, act, B->x32, B->r32, nullptr);
  int tok = prof ? B->prof->begin(B->ps, TDGL_K_MGS, (i + 1.0) * S3, 2) : -1;  // read Q_0..Q_{i-1}, W
  launch_bmdot(bk, 1, B->d_qtab, B->r32, B->d_iter, B->d_pm, B->ps);
  launch_bcolsum(bk, B->d_iter, B->d_pm, CS_MD1, kColsumY, B->ps);
  if (prof) B->prof->end(B->ps, tok);
  tok = prof ? B->prof->begin(B->ps, TDGL_K_MGS_LAST, (i + 3.0) * S3, 2) : -1;  // Q_j, r; W r/w
  launch_bupd(bk, 1, B->d_qtab, B->r32, B->d_iter, B->d_pm, B->ps);
  launch_bcolsum(bk, B->d_iter, B->d_pm, CS_UP1, 1, B->ps);
  if (prof) B->prof->end(B->ps, tok);
  // second pass: the kernels return at once for films that do not need it
  tok = prof ? B->prof->begin(B->ps, TDGL_K_CGS2, 0.0, 4) : -1;
  launch_bmdot(bk, 2, B->d_qtab, B->r32, B->d_iter, B->d_pm, B->ps);
  launch_bcolsum(bk, B->d_iter, B->d_pm, CS_MD2, kColsumY, B->ps);
  launch_bupd(bk, 2, B->d_qtab, B->r32, B->d_iter, B->d_pm, B->ps);
  launch_bcolsum(bk, B->d_iter, B->d_pm, CS_UP2, 1, B->ps);
  if (prof) B->prof->end(B->ps, tok);
  tok = prof ? B->prof->begin(B->ps, TDGL_K_FINAL_P, 4.0 * S3) : -1;
  launch_bfinish(bk, B->d_qtab, B->r32, B->d_iter, B->d_H, B->d_pm, B->ps);
  if (prof) B->prof->end(B->ps, tok);
  tok = prof ? B->prof->begin(B->ps, TDGL_K_UPDATE_R, 0.0, 2) : -1;
  launch_bstatus(bk, B->d_fstat, B->d_iter, tol, B->d_pm, B->ps);
  launch_bnext(bk, B->d_iter, B->d_nact, B->ps);
  if (prof) B->prof->end(B->ps, tok);
}

// Iteration i >= 1 without profiling: replay the captured graph of the body.
int launch_body(tdgl_batch* B, double tol) {
  if (!B->use_graph) {
    enqueue_body(B, tol, false, 1, {});
    BCHK(hipGetLastError());
    return TDGL_OK;
  }
  if (B->body && (B->body_tol != tol || B->body_deg != B->cheb_deg || B->body_gen != B->gen ||
                  B->body_part != B->part_gen)) {
    BCHK(hipGraphExecDestroy(B->body));
    B->body = nullptr;
  }
  if (!B->body) {
    hipGraph_t g = nullptr;
    BCHK(hipStreamBeginCapture(B->ps, hipStreamCaptureModeThreadLocal));
    enqueue_body(B, tol, false, 1, {});
    const hipError_t e = hipGetLastError();
    BCHK(hipStreamEndCapture(B->ps, &g));
    BCHK(e);
    const hipError_t ei = hipGraphInstantiate(&B->body, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BCHK(ei);
    B->body_tol = tol;
    B->body_deg = B->cheb_deg;
    B->body_gen = B->gen;
    B->body_part = B->part_gen;
  }
  BCHK(hipGraphLaunch(B->body, B->ps));
  return TDGL_OK;
}

// Queue the copy of the films' flags into readback slot `slot`.
int post_flags(tdgl_batch* B, int slot) {
  BCHK(hipMemcpyAsync(B->h_act + (size_t)B->nfilm * (1 + slot), B->d_act, sizeof(int) * B->nfilm,
                      hipMemcpyDeviceToHost, B->ps));
  BCHK(hipEventRecord(B->ev_act[slot], B->ps));
  return TDGL_OK;
}
// Wait for slot `slot` and count the films still iterating.
int read_flags(tdgl_batch* B, int slot, std::vector<int>& act, int* live) {
  BCHK(hipEventSynchronize(B->ev_act[slot]));
  const int* h = B->h_act + (size_t)B->nfilm * (1 + slot);
  *live = 0;
  for (int f = 0; f < B->nfilm; ++f) {
    act[f] = h[f];
    *live += act[f];
  }
  return TDGL_OK;
}

// TGCR for the active films on (I - dt/2 J) dx = r32 at x32; on return the
// report holds each film's outcome.  Without profiling the host keeps one
// iteration queued ahead of the flags it reads (a film that stopped makes the
// extra iteration a no-op for it), and iterations i >= 1 replay a hipGraph.
int tgcr(tdgl_batch* B, const std::vector<int>& act0, double tol, std::vector<tdgl_step_stats>& S) {
  const int F = B->nfilm;
  const bool prof = B->prof->on;
  int rc;
  if ((rc = set_act(B, act0))) return rc;
  launch_btinit(kview(B), B->d_fstat, B->d_maxk, B->d_iter, B->ps);
  BCHK(hipMemcpyAsync(B->r0_32, B->r32, sizeof(float) * (size_t)B->total_len, hipMemcpyDeviceToDevice, B->ps));
  std::vector<int> act = act0;
  const BatchK bk = kview(B);
  if ((rc = ensure_q(B, 1))) return rc;
  // iteration 0
  {
    const double S3 = prof ? sum_over(B, act, s32) : 0.0;
    enqueue_w(B, prof, act, B->x32, B->r32, nullptr);
    int tok = B->prof->begin(B->ps, TDGL_K_MGS, 2.0 * S3, 2);
    launch_bmdot(bk, 0, B->d_qtab, B->r32, B->d_iter, B->d_pm, B->ps);
    launch_bcolsum(bk, B->d_iter, B->d_pm, CS_MD0, 1, B->ps);
    B->prof->end(B->ps, tok);
    tok = B->prof->begin(B->ps, TDGL_K_FINAL_P, 4.0 * S3);
    launch_bfinish(bk, B->d_qtab, B->r32, B->d_iter, B->d_H, B->d_pm, B->ps);
    B->prof->end(B->ps, tok);
    tok = B->prof->begin(B->ps, TDGL_K_UPDATE_R, 0.0, 2);
    launch_bstatus(bk, B->d_fstat, B->d_iter, tol, B->d_pm, B->ps);
    launch_bnext(bk, B->d_iter, B->d_nact, B->ps);
    B->prof->end(B->ps, tok);
    BCHK(hipGetLastError());
    if ((rc = post_flags(B, 0))) return rc;
  }
  int live = 0;
  for (int i = 0;; ++i) {
    // iteration i is queued; flags of iteration i are in slot i % 2
    if (prof) {
      if ((rc = read_flags(B, i % 2, act, &live))) return rc;
      B->prof->harvest();
      if (!live) break;
      if ((rc = ensure_q(B, i + 1))) return rc;
      enqueue_body(B, tol, true, i + 1, act);
      BCHK(hipGetLastError());
      if ((rc = post_flags(B, (i + 1) % 2))) return rc;
      continue;
    }
    const bool ahead = i + 1 < kMaxKrylov;
    if (ahead) {
      if ((rc = ensure_q(B, std::min(kMaxKrylov - 1, ((i + 1) / 16 + 1) * 16)))) return rc;
      if ((rc = launch_body(B, tol))) return rc;
      if ((rc = post_flags(B, (i + 1) % 2))) return rc;
    }
    if ((rc = read_flags(B, i % 2, act, &live))) return rc;
    if (!live) break;
    if (!ahead) return bfail(B, TDGL_E_INVALID, "TGCR needs more than 2048 Krylov vectors");
  }
  if ((rc = report(B))) return rc;
  for (int f = 0; f < F; ++f) {
    if (!act0[f]) continue;
    const double kit = rep(B, f)[8];
    const int d = B->cheb_deg ? B->cheb_degf[f] : 0;
    S[f].jvp_evals += (int)kit * std::max(d, 1);
    // iteration i: W (3 without a preconditioner), dots (i + 1), update (i + 3), finish 4
    const double wb = d == 0 ? 3.0 : 3.0 * d + (d >= 2 ? d - 1.0 : 0.0) + (d >= 3 ? 1.0 : 0.0);
    if (kit >= 1) S[f].bytes_moved += (wb + 6.0 + (wb + 8.0) * (kit - 1) + (kit - 1) * kit) * s32(B, f);
  }
  return TDGL_OK;
}

int batch_step_impl(tdgl_batch* B, double* X, const double* dt, const int* active,
                    const tdgl_solver_opts* o, tdgl_step_stats* stats);

int batch_step(tdgl_batch* B, double* X, const double* dt, const int* active,
               const tdgl_solver_opts* o, tdgl_step_stats* stats) {
  int rc;
  if ((rc = sync_in(B))) return rc;
  rc = batch_step_impl(B, X, dt, active, o, stats);
  const int rc2 = sync_out(B);
  return rc ? rc : rc2;
}

int batch_step_impl(tdgl_batch* B, double* X, const double* dt, const int* active,
                    const tdgl_solver_opts* o, tdgl_step_stats* stats) {
  const int F = B->nfilm;
  int rc;
  if ((rc = ensure_work(B))) return rc;
  if ((rc = refresh(B))) return rc;
  std::vector<double> dtv(2 * F);
  std::vector<int> maxk(F), req(F);
  for (int f = 0; f < F; ++f) {
    dtv[2 * f] = dt[f];
    dtv[2 * f + 1] = dt[f] / 2.0;
    const long long N = (long long)(1 + B->hf[f].g.nphi) * B->hf[f].g.n;
    // tgcr.py:110-112 default cap: the state length
    long long r = o->tgcr_max_iter > 0 ? o->tgcr_max_iter : std::max<long long>(N, llround(0.2 * (double)N));
    req[f] = (int)std::min<long long>(r, 1LL << 30);
    maxk[f] = std::min(req[f], B->kcap[f]);
  }
  BCHK(hipMemcpyAsync(B->d_dtv, dtv.data(), sizeof(double) * 2 * F, hipMemcpyHostToDevice, B->ps));
  BCHK(hipMemcpyAsync(B->d_maxk, maxk.data(), sizeof(int) * F, hipMemcpyHostToDevice, B->ps));
  const double floor_ = o->mixed_tol_floor > 0 ? o->mixed_tol_floor : 1e-6;
  const double tol = std::max(o->tol_gcr, floor_);

  std::vector<tdgl_step_stats> S(F);
  memset(S.data(), 0, sizeof(tdgl_step_stats) * F);
  std::vector<int> act(F);
  for (int f = 0; f < F; ++f) act[f] = active ? (active[f] != 0) : 1;
  if ((rc = setup_cheb(B, o->mixed_precond, dtv, act))) return rc;
  if ((rc = ensure_parts(B, act))) return rc;

  // predictor + gamma (integrators.py:156-158)
  if ((rc = set_act(B, act))) return rc;
  {
    const int tok = B->prof->begin(B->ps, TDGL_K_PRED, 3.0 * sum_over(B, act, s64));
    launch_bstencil(kview(B), B->is3d, M_PRED, X, B->w_x, B->w_gam, nullptr, nullptr, nullptr,
                    B->d_pitem, B->ps);
    B->prof->end(B->ps, tok);
    BCHK(hipGetLastError());
    for (int f = 0; f < F; ++f)
      if (act[f]) {
        S[f].rhs_evals++;
        S[f].bytes_moved += 3.0 * s64(B, f);
      }
  }
  if (o->x_prev) {  // Newton starts from the extrapolated state where a film has a last step
    std::vector<double> ratio(F, 0.0);
    double by = 0.0;
    for (int f = 0; f < F; ++f)
      if (act[f] && o->dt_prev[f] > 0.0) {
        ratio[f] = dt[f] / o->dt_prev[f];
        by += 3.0 * s64(B, f);
        S[f].bytes_moved += 3.0 * s64(B, f);
      }
    if (by > 0.0) {
      BCHK(hipMemcpyAsync(B->d_ext, ratio.data(), sizeof(double) * F, hipMemcpyHostToDevice, B->ps));
      const int tok = B->prof->begin(B->ps, TDGL_K_PRED, by);
      launch_bextrap(kview(B), X, o->x_prev, B->d_ext, B->w_x, B->ps);
      B->prof->end(B->ps, tok);
      BCHK(hipGetLastError());
    }
  }
  if ((rc = newton_eval(B, act, S))) return rc;
  if ((rc = report(B))) return rc;
  for (int f = 0; f < F; ++f)
    if (act[f]) S[f].g_inf = rep(B, f)[2];

  std::vector<int> newton = act;
  for (int it = 1; it <= o->newton_max_iter; ++it) {
    if ((rc = tgcr(B, newton, tol, S))) return rc;
    std::vector<int> ok(F, 0);
    int nok = 0;
    for (int f = 0; f < F; ++f) {
      if (!newton[f]) continue;
      const double* r = rep(B, f);
      const int state = (int)r[6], kit = (int)r[8];
      S[f].tgcr_iters_last = kit;
      S[f].tgcr_iters_total += kit;
      S[f].tgcr_r0 = r[4];
      S[f].tgcr_rk = r[5];
      if (state == TG_NONFINITE) return bfail(B, TDGL_E_NONFINITE, "TGCR residual is not finite");
      if (state == TG_MAXITER && maxk[f] < req[f])
        return bfail(B, TDGL_E_INVALID, "TGCR needs more than 2048 Krylov vectors");
      if (state == TG_CONVERGED) {
        ok[f] = 1;
        ++nok;
      } else {  // newton.py:102-106: TGCR failed, the step attempt fails
        S[f].tgcr_failed = 1;
        S[f].newton_iters = it;
        S[f].converged = 0;
        newton[f] = 0;
      }
    }
    if (nok) {
      if ((rc = set_act(B, ok))) return rc;
      const int d = B->cheb_deg;
      if (d == 0) {
        double by = 2.0 * sum_over(B, ok, s64);
        for (int f = 0; f < F; ++f)
          if (ok[f]) by += (rep(B, f)[7] + 1.0) * s32(B, f);
        const int tok = B->prof->begin(B->ps, TDGL_K_ACCUM, by, 3);
        launch_btrisolve(kview(B), B->d_fstat, B->d_H, B->ps);
        launch_baccum(kview(B), B->w_x, B->r0_32, B->d_qtab, B->d_fstat, B->d_pm, B->ps);
        launch_bcolsum(kview(B), B->d_iter, B->d_pm, CS_DXINF, 1, B->ps);
        B->prof->end(B->ps, tok);
      } else {  // u = R y into r0_32, then x += M u
        double by = 0.0;
        for (int f = 0; f < F; ++f)
          if (ok[f]) by += (rep(B, f)[7] + 2.0) * s32(B, f);
        const int tok = B->prof->begin(B->ps, TDGL_K_ACCUM, by, 2);
        launch_btrisolve(kview(B), B->d_fstat, B->d_H, B->ps);
        launch_baccum_u(kview(B), B->r0_32, B->r0_32, B->d_qtab, B->d_fstat, B->ps);
        B->prof->end(B->ps, tok);
        enqueue_precond_update(B, B->prof->on, ok);
      }
      BCHK(hipGetLastError());
      for (int f = 0; f < F; ++f) {
        if (!ok[f]) continue;
        const double kk = rep(B, f)[7];
        if (d == 0) {
          S[f].bytes_moved += (kk + 1.0) * s32(B, f) + 2.0 * s64(B, f);
        } else {
          const int df = B->cheb_degf[f];
          const double zb = df >= 2 ? 3.0 * (df - 1) + (df >= 3 ? 1.0 : 0.0) + (df >= 4 ? 2.0 * (df - 3) : 0.0) : 0.0;
          S[f].bytes_moved += (kk + 2.0 + zb + 1.0) * s32(B, f) + 2.0 * s64(B, f);
          S[f].jvp_evals += std::max(df - 1, 0);
        }
      }
      if ((rc = newton_eval(B, ok, S))) return rc;  // newton.py:109-113
      if ((rc = report(B))) return rc;
      std::vector<int> conv(F, 0);
      int nconv = 0;
      for (int f = 0; f < F; ++f) {
        if (!ok[f]) continue;
        const double* r = rep(B, f);
        S[f].g_inf = r[2];
        S[f].dx_inf = r[3];
        if (!std::isfinite(S[f].g_inf) || !std::isfinite(S[f].dx_inf))
          return bfail(B, TDGL_E_NONFINITE, "Newton residual is not finite");
        if (S[f].g_inf <= o->newton_tol_f && S[f].dx_inf <= o->newton_tol_dx) {  // newton.py:118
          conv[f] = 1;
          ++nconv;
          S[f].bytes_moved += 2.0 * s64(B, f);
          S[f].converged = 1;
          S[f].newton_iters = it;
          newton[f] = 0;
        }
      }
      // the films that converged take their iterate: one launch for all of them
      if (nconv && (rc = copy_films(B, conv, B->w_x, X))) return rc;
    }
    bool any = false;
    for (int f = 0; f < F; ++f) any = any || newton[f];
    if (!any) break;
  }
  for (int f = 0; f < F; ++f)
    if (newton[f]) {
      S[f].converged = 0;
      S[f].newton_iters = o->newton_max_iter;
    }
  BCHK(hipStreamSynchronize(B->ps));
  memcpy(stats, S.data(), sizeof(tdgl_step_stats) * F);
  return TDGL_OK;
}

// A context's own batch of one follows the context's tuning.
bool nfilm_is_own(const tdgl_batch* B) { return B->nfilm == 1 && B->own_stream_from_ctx; }

int batch_init(tdgl_batch* B) {
  const int F = B->nfilm;
  int rc;
  BCHK(hipStreamCreateWithFlags(&B->ps, hipStreamNonBlocking));
  if ((rc = balloc(B, (void**)&B->d_films, sizeof(FilmDesc) * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_act, sizeof(int) * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_sel, sizeof(int) * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_res, sizeof(double) * kResLen * (size_t)F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_dtv, sizeof(double) * 2 * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_ext, sizeof(double) * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_fstat, sizeof(int) * 4 * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_maxk, sizeof(int) * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_iter, sizeof(int)))) return rc;
  if ((rc = balloc(B, (void**)&B->d_nact, sizeof(int)))) return rc;
  if ((rc = balloc(B, (void**)&B->d_rep, sizeof(double) * kReportLen * F))) return rc;
  if ((rc = balloc(B, (void**)&B->d_cheb, sizeof(double) * kChebStride * F))) return rc;
  BCHK(hipHostMalloc((void**)&B->h_rep, sizeof(double) * kReportLen * F, hipHostMallocDefault));
  BCHK(hipHostMalloc((void**)&B->h_act, sizeof(int) * 4 * F, hipHostMallocDefault));
  BCHK(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
  BCHK(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming));
  BCHK(hipEventCreateWithFlags(&B->ev_act[0], hipEventDisableTiming));
  BCHK(hipEventCreateWithFlags(&B->ev_act[1], hipEventDisableTiming));
  {
    int dev = 0;
    BCHK(hipGetDevice(&dev));
    BCHK(hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  B->kz.resize(F);
  B->kcap.resize(F);
  if (nfilm_is_own(B)) B->shape = B->films[0]->tile_user;
  long long planes = 0;
  for (int f = 0; f < F; ++f) planes += stencil_tile_planes(B->films[f]->g);
  for (int f = 0; f < F; ++f) {
    const Geo& g = B->films[f]->g;
    B->kz[f] = B->films[f]->kz_user > 0 && nfilm_is_own(B) ? B->films[f]->kz_user : default_kz(g, planes);
    B->kcap[f] = (int)std::min<long long>((long long)(1 + g.nphi) * g.n, kMaxKrylov);
  }
  return build_layout(B);
}

}  // namespace

namespace tdgl {
double cheb_upper_bound(const tdgl_ctx* c, double hdt) { return cheb_upper(c, hdt); }
}  // namespace tdgl


namespace tdgl {

// The context's own batch of one: its profile and tuning are the context's.
int own_batch(tdgl_ctx* ctx) {
  if (ctx->own_batch) return TDGL_OK;
  if (!fits_sweeps(ctx)) {
    ctx->err = "film too large for the mixed-precision kernels (20 n_interior >= 2^32)";
    return TDGL_E_INVALID;
  }
  tdgl_batch* B = new tdgl_batch();
  B->films.assign(1, ctx);
  B->nfilm = 1;
  B->is3d = ctx->g.is3d != 0;
  B->st = ctx->st;
  B->own_stream_from_ctx = true;
  B->prof = &ctx->prof;
  const int rc = batch_init(B);
  if (rc) {
    ctx->err = B->err;
    tdgl_batch_destroy(B);
    return rc;
  }
  ctx->own_batch = B;
  return TDGL_OK;
}

int mixed_step_single(tdgl_ctx* ctx, double* X, double dt, const tdgl_solver_opts* o,
                      tdgl_step_stats* s) {
  int rc0;
  if ((rc0 = own_batch(ctx))) return rc0;
  tdgl_batch* B = ctx->own_batch;
  B->st = ctx->st;
  const int rc = batch_step(B, X, &dt, nullptr, o, s);
  if (rc) ctx->err = B->err;
  return rc;
}

// The single-film FP32 entry points take and return vectors in the internal
// layout; the sweeps work in the mixed method's (each node's links together),
// so the vectors pass through the batch's work buffers on the way.
int jvp32_single(tdgl_ctx* ctx, const float* x32, const float* p32, double dt, float* ap32) {
  int rc;
  if ((rc = own_batch(ctx))) return rc;
  tdgl_batch* B = ctx->own_batch;
  B->st = ctx->st;
  if ((rc = ensure_work(B)) || (rc = sync_in(B)) || (rc = refresh(B))) {
    ctx->err = B->err;
    return rc;
  }
  const long long n = B->hf[0].g.n;
  const int nphi = B->hf[0].g.nphi;
  const double dtv[2] = {dt, dt / 2.0};
  std::vector<int> one(1, 1);
  if ((rc = set_act(B, one))) return rc;
  if (hipMemcpyAsync(B->d_dtv, dtv, sizeof(dtv), hipMemcpyHostToDevice, B->ps) != hipSuccess)
    return TDGL_E_HIP;
  B->cheb_deg = 0;
  if ((rc = ensure_parts(B, one))) {
    ctx->err = B->err;
    return rc;
  }
  launch_l32_permute(x32, B->x32, n, nphi, true, B->ps);
  launch_l32_permute(p32, B->r32, n, nphi, true, B->ps);
  launch_bjvp32(kview(B), B->part[0], B->is3d, B->x32, B->r32, nullptr, nullptr, B->r0_32, B->ps);
  launch_l32_permute(B->r0_32, ap32, n, nphi, false, B->ps);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(B->ps) != hipSuccess) return TDGL_E_HIP;
  return sync_out(B);
}

int precond_single(tdgl_ctx* ctx, const float* x32, const float* v32, double dt, int degree,
                   float* w32, float* z32, double* info) {
  int rc;
  if ((rc = own_batch(ctx))) return rc;
  tdgl_batch* B = ctx->own_batch;
  B->st = ctx->st;
  auto done = [&](int code) {
    if (code) ctx->err = B->err;
    return code;
  };
  if ((rc = ensure_work(B)) || (rc = sync_in(B)) || (rc = refresh(B))) return done(rc);
  const std::vector<double> dtv = {dt, dt / 2.0};
  std::vector<int> one(1, 1);
  if ((rc = set_act(B, one))) return done(rc);
  if (hipMemcpyAsync(B->d_dtv, dtv.data(), sizeof(double) * 2, hipMemcpyHostToDevice, B->ps) != hipSuccess)
    return done(bfail(B, TDGL_E_HIP, "hipMemcpyAsync"));
  double iv[2] = {0.0, 0.0};
  if ((rc = setup_cheb(B, degree, dtv, one, iv))) return done(rc);
  if (B->cheb_deg == 0) return done(bfail(B, TDGL_E_INVALID, "no preconditioner for this degree and dt"));
  if ((rc = ensure_parts(B, one))) return done(rc);
  const long long n = B->hf[0].g.n;
  const int nphi = B->hf[0].g.nphi;
  launch_l32_permute(x32, B->x32, n, nphi, true, B->ps);
  launch_l32_permute(v32, B->r0_32, n, nphi, true, B->ps);
  enqueue_w(B, false, one, B->x32, B->r0_32, B->r32);
  enqueue_z(B, false, one, B->x32, B->r0_32);
  launch_l32_permute(B->r32, w32, n, nphi, false, B->ps);
  const float* z = B->cheb_deg >= 2 ? B->t32[B->cheb_deg % 2] : nullptr;  // one film: its own degree
  if (z) launch_l32_permute(z, z32, n, nphi, false, B->ps);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(B->ps) != hipSuccess)
    return done(bfail(B, TDGL_E_HIP, "preconditioner kernels failed"));
  // degree 1: z = v / theta
  info[0] = iv[0];
  info[1] = iv[1];
  info[2] = (double)B->cheb_deg;
  info[3] = z ? 0.0 : 1.0;
  return done(sync_out(B));
}

}  // namespace tdgl

extern "C" {

int tdgl_batch_create(tdgl_ctx* const* films, int nfilm, tdgl_batch** out) {
  if (!films || nfilm <= 0 || !out) return TDGL_E_INVALID;
  *out = nullptr;
  const bool is3d = films[0] && films[0]->g.is3d;
  for (int f = 0; f < nfilm; ++f)
    if (!films[f] || (films[f]->g.is3d != 0) != is3d || !fits_sweeps(films[f])) return TDGL_E_INVALID;
  tdgl_batch* B = new tdgl_batch();
  B->films.assign(films, films + nfilm);
  B->nfilm = nfilm;
  B->is3d = is3d;
  B->st = films[0]->st;
  const int rc = batch_init(B);
  if (rc) {
    tdgl_batch_destroy(B);
    return rc;
  }
  *out = B;
  return TDGL_OK;
}

int tdgl_batch_destroy(tdgl_batch* B) {
  if (!B) return TDGL_OK;
  if (B->ps) (void)hipStreamSynchronize(B->ps);
  else (void)hipDeviceSynchronize();
  if (B->body) (void)hipGraphExecDestroy(B->body);
  for (hipEvent_t e : {B->ev_in, B->ev_out, B->ev_act[0], B->ev_act[1]})
    if (e) (void)hipEventDestroy(e);
  if (B->ps) (void)hipStreamDestroy(B->ps);
  for (void* p : {(void*)B->d_films, (void*)B->d_item_film, (void*)B->d_pseg,
                  (void*)B->d_sblk_film,
                  (void*)B->d_act, (void*)B->d_sel,
                  (void*)B->d_res, (void*)B->d_dtv, (void*)B->d_ext, (void*)B->d_fstat, (void*)B->d_maxk,
                  (void*)B->d_iter, (void*)B->d_nact, (void*)B->d_pitem, (void*)B->d_pm,
                  (void*)B->d_H, (void*)B->d_rep, (void*)B->w_x, (void*)B->w_gam, (void*)B->x32,
                  (void*)B->r32, (void*)B->r0_32, (void*)B->d_qtab, (void*)B->d_cheb,
                  (void*)B->t32[0], (void*)B->t32[1]})
    if (p) (void)hipFree(p);
  for (float* q : B->Q) (void)hipFree(q);
  for (void* p : B->retired) (void)hipFree(p);
  if (B->h_rep) (void)hipHostFree(B->h_rep);
  if (B->h_act) (void)hipHostFree(B->h_act);
  B->own_prof.destroy();
  delete B;
  return TDGL_OK;
}

int tdgl_batch_set_stream(tdgl_batch* B, void* s) {
  if (!B) return TDGL_E_INVALID;
  B->st = (hipStream_t)s;
  return TDGL_OK;
}

// HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues and
// shares a queue between streams once they run out; batches stepped at once on
// one queue run one after the other.  A stream made with a CU mask always gets
// a queue of its own, so the batch's work moves to one with every CU enabled.
// The new stream is a blocking one (it orders after the legacy stream's work).
int tdgl_batch_own_queue(tdgl_batch* B) {
  if (!B) return TDGL_E_INVALID;
  int dev = 0;
  hipDeviceProp_t prop;
  BCHK(hipGetDevice(&dev));
  BCHK(hipGetDeviceProperties(&prop, dev));
  const int ncu = prop.multiProcessorCount;
  std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
  for (int i = 0; i < ncu; ++i) mask[(size_t)i / 32] |= 1u << (i % 32);
  hipStream_t s = nullptr;
  BCHK(hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()));
  if (const hipError_t e = hipStreamSynchronize(B->ps); e != hipSuccess) {
    (void)hipStreamDestroy(s);
    return bfail(B, TDGL_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  }
  if (B->body) {
    (void)hipGraphExecDestroy(B->body);
    B->body = nullptr;
  }
  BCHK(hipStreamDestroy(B->ps));
  B->ps = s;
  return TDGL_OK;
}

int tdgl_batch_work_stream(const tdgl_batch* B, void** s) {
  if (!B || !s) return TDGL_E_INVALID;
  *s = (void*)B->ps;
  return TDGL_OK;
}

const char* tdgl_batch_last_error(const tdgl_batch* B) { return B ? B->err.c_str() : "null batch"; }

int64_t tdgl_batch_total_len(const tdgl_batch* B) { return B ? B->total_len : -1; }

int64_t tdgl_batch_offset(const tdgl_batch* B, int f) {
  if (!B || f < 0 || f >= B->nfilm) return -1;
  return B->hf[f].off;
}

int tdgl_batch_set_kz(tdgl_batch* B, int f, int kz) {
  if (!B || f < 0 || f >= B->nfilm || kz < 1) return TDGL_E_INVALID;
  const Geo& g = B->films[f]->g;
  B->kz[f] = g.is3d ? std::min(kz, g.nz) : 1;
  (void)hipStreamSynchronize(B->ps);
  return build_layout(B);
}

int tdgl_batch_get_kz(const tdgl_batch* B, int f) {
  if (!B || f < 0 || f >= B->nfilm) return -1;
  return B->kz[f];
}

int tdgl_batch_set_sweep_split(tdgl_batch* B, int max_planes, int min_planes) {
  if (!B || min_planes < 1 || max_planes < min_planes || max_planes > 4096) return TDGL_E_INVALID;
  B->part_lmax = max_planes;
  B->part_lmin = min_planes;
  B->part_key.clear();
  return TDGL_OK;
}
int tdgl_batch_get_sweep_split(const tdgl_batch* B, int* max_planes, int* min_planes) {
  if (!B) return TDGL_E_INVALID;
  if (max_planes) *max_planes = B->part_lmax;
  if (min_planes) *min_planes = B->part_lmin;
  return TDGL_OK;
}

int tdgl_batch_set_sweep_tile(tdgl_batch* B, int shape) {
  if (!B || shape < 0 || shape >= kSweepShapes) return TDGL_E_INVALID;
  B->shape = shape;
  (void)hipStreamSynchronize(B->ps);
  return build_layout(B);
}

int tdgl_batch_get_sweep_tile(const tdgl_batch* B) { return B ? B->shape : -1; }

int tdgl_batch_time_stencil(tdgl_batch* B, int reps, double* ms) {
  if (!B || reps < 1 || !ms) return TDGL_E_INVALID;
  int rc;
  if ((rc = ensure_work(B))) return rc;
  if ((rc = ensure_q(B, 0))) return rc;
  if ((rc = sync_in(B)) || (rc = refresh(B))) return rc;
  std::vector<int> all(B->nfilm, 1);
  std::vector<double> dtv(2 * B->nfilm, 0.1);
  if ((rc = set_act(B, all))) return rc;
  BCHK(hipMemcpyAsync(B->d_dtv, dtv.data(), sizeof(double) * dtv.size(), hipMemcpyHostToDevice, B->ps));
  B->cheb_deg = 0;  // the plain JVP sweep over every film (the next step sets the degrees again)
  if ((rc = ensure_parts(B, all))) return rc;
  hipEvent_t a, b;
  BCHK(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
  BCHK(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
  launch_bjvp32(kview(B), B->part[0], B->is3d, B->x32, B->r32, nullptr, nullptr, B->Q[0], B->ps);  // warm-up
  BCHK(hipEventRecord(a, B->ps));
  for (int r = 0; r < reps; ++r)
    launch_bjvp32(kview(B), B->part[0], B->is3d, B->x32, B->r32, nullptr, nullptr, B->Q[0], B->ps);
  BCHK(hipEventRecord(b, B->ps));
  BCHK(hipEventSynchronize(b));
  float t = 0.f;
  BCHK(hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = (double)t / reps;
  return sync_out(B);
}

int tdgl_batch_set_graphs(tdgl_batch* B, int enable) {
  if (!B) return TDGL_E_INVALID;
  B->use_graph = enable != 0;
  return TDGL_OK;
}

int tdgl_batch_trap_step(tdgl_batch* B, double* X, const double* dt, const int* active,
                         const tdgl_solver_opts* o, tdgl_step_stats* stats) {
  if (!B || !X || !dt || !o || !stats) return TDGL_E_INVALID;
  if (o->precision != TDGL_PREC_MIXED)
    return bfail(B, TDGL_E_INVALID, "batches run the mixed-precision method (TDGL_PREC_MIXED)");
  if (o->x_prev && !o->dt_prev) return bfail(B, TDGL_E_INVALID, "x_prev needs dt_prev");
  for (int f = 0; f < B->nfilm; ++f)
    if (!(dt[f] > 0.0) && (!active || active[f])) return bfail(B, TDGL_E_INVALID, "dt must be > 0");
  return batch_step(B, X, dt, active, o, stats);
}

int tdgl_batch_profile_enable(tdgl_batch* B, int enable) {
  if (!B) return TDGL_E_INVALID;
  if (hipStreamSynchronize(B->ps) != hipSuccess) return TDGL_E_HIP;
  B->prof->harvest();
  B->prof->on = enable != 0;
  B->prof->reset();
  return TDGL_OK;
}

int tdgl_batch_profile_read(tdgl_batch* B, int cls, double* ms, int64_t* n, double* bytes) {
  if (!B || cls < 0 || cls >= TDGL_K_COUNT) return TDGL_E_INVALID;
  if (hipStreamSynchronize(B->ps) != hipSuccess) return TDGL_E_HIP;
  B->prof->harvest();
  if (ms) *ms = B->prof->ms[cls];
  if (n) *n = B->prof->n[cls];
  if (bytes) *bytes = B->prof->bytes[cls];
  return TDGL_OK;
}

}  // extern "C"
