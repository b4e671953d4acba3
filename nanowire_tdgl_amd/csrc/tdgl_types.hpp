// Plain structs passed by value from the host driver to the kernels.
#pragma once

#include <stdint.h>

namespace tdgl {

// Per-node flag bits (full-grid numbering m = i + mj*j + mk*k).
constexpr uint8_t kHoleX = 1;      // x link at m is a hole-boundary link
constexpr uint8_t kHoleY = 2;
constexpr uint8_t kHoleZ = 4;
constexpr uint8_t kInsulator = 8;  // interior_sc_mask == 0 at m (binary masks only)

struct Geo {
  int Nx, Ny, Nz;       // cells (SimulationParameters.Nx ...)
  int nx, ny, nz;       // interior nodes per axis (nz = 1 in 2-D)
  int is3d;
  int per_x, per_y, per_z;
  int mj;               // Nx + 1
  long long mk;         // (Nx + 1)(Ny + 1)
  long long n;          // interior nodes
  int nphi;             // 3 (3-D) or 2 (2-D)
};

// Applied field seen by the ghost rules.
struct Bnd {
  double cx, cy, cz;    // (Bx*hy)*hz, (By*hz)*hx, (Bz*hx)*hy for uniform fields
  const double* bx;     // full-grid BoundaryVectors (general mode) or nullptr
  const double* by;
  const double* bz;
  double hx, hy, hz;
};

// Material / operator coefficients.
struct Mat {
  const uint8_t* flags; // full-grid flag bits or nullptr
  const double* k2;     // per-node kappa^2 (internal order) or nullptr (uniform)
  const double* sc;     // per-node interior_sc_mask (internal order) or nullptr
  int has_mat;          // FPSI uses the sc/insulator form
  int sc_bits;          // sc is binary and stored as kInsulator bits
  // uniform-kappa coefficients kappa^2/h^2 and the summed LPHI diagonals
  double cx, cy, cz, dgx, dgy, dgz;
  // 1/h^2 (the reference divides the CSR operators by h^2 through 1./h**2)
  double ihx2, ihy2, ihz2;
  double dgpsi;         // ((-2/hx^2) + (-2/hy^2)) + (-2/hz^2)
  double hx2, hy2, hz2; // h^2 for per-node kappa^2/h^2
  double inv_tau;       // fl(1/0.1): numpy's complex/real division multiplies by the reciprocal
};

// One film of a batch: its operator, and where its work and data sit in the
// batch's launches and buffers.
struct FilmDesc {
  Geo g;
  Bnd b;
  Mat mt;
  long long off;                 // first element of the film in batched vectors (multiple of 64)
  long long len;                 // elements of one state vector: (2 + nphi) n
  int tiles_x, tiles_y, kz, tz;  // FP64 stencil decomposition: tiles_x*tiles_y*tz work items
  int item0, nitems;             // the film's FP64 stencil blocks
  int tw32, th32;                // FP32 sweep tile of the film: nodes per row, rows
  int tiles32_x, tiles32_y;      // FP32 sweep tiles per plane
  int sb0, nsb;                  // the film's streaming blocks
  long long hoff;                // the film's packed triangular TGCR matrix in the batch's H buffer
};

}  // namespace tdgl
