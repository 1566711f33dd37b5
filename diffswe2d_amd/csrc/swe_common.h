// Shared device plumbing: launch geometry, ghost-cell fetch, control block.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swe_math.h"

namespace swe {

constexpr int kWave = 64;              // CDNA wavefront
constexpr int kHalo = 2;               // ghost / stencil halo (ng = 2)
constexpr int kOut = kWave - 2 * kHalo;  // output columns per wavefront strip
constexpr int kSlots = 64;             // wave-speed accumulator slots (spread atomics)
constexpr int kWavesPerBlock = 4;

enum Bc : int { BC_TRANSMISSIVE = 0, BC_PERIODIC = 1, BC_CONSTANT = 2, BC_WATER_LEVEL = 3, BC_WALL = 4,
                BC_ABSORBING = 5 };
enum Side4 : int { LEFT = 0, RIGHT = 1, TOP = 2, BOTTOM = 3 };

// Device-resident control block at the head of the workspace.
struct Control {
  double elapsed[2];   // time advanced so far, indexed by step parity
  double duration;     // requested duration of the current run
  double tol;          // max(1e-12, 1e-12 * duration)
  double cap;          // max_dt_cap, < 0 when absent
  double bad_dt;       // offending dt when status != 0
  int32_t done;        // a step found remaining <= tol (or failed)
  int32_t steps;       // steps completed in the current run
  int32_t status;      // 0 ok, 1 non-finite dt, 2 non-positive dt
  int32_t pad;
  double last_dt;      // dt of the most recent step (host-driven steps)
  // chunked runs: the duration is one chunk; at each chunk end the clock kernel
  // samples the gauges, moves to the next chunk's boundary level and restarts
  // the chunk's elapsed time, until the last chunk ends
  int32_t chunked;
  int32_t chunk;       // current chunk
  int32_t n_chunks;
  int32_t pad2;
  double level_f64;    // the current chunk's boundary level, read by the stage kernels
  float level_f32;
  float pad3;
  const void* levels;  // per-chunk levels (device, the run's dtype)
};

// Grid geometry and field pointers of one model (one batch of members).
template <class T> struct Fields {
  int B, ny, nx;
  int64_t plane;         // ny * nx
  int64_t bed_bstride;   // 0 when the bed is shared by all members
  int64_t rough_bstride;
  int bc[4];
  const T* bed;
  const T* rough;
  const T* level;        // boundary level per member (stride level_stride)
  int level_stride;
};

// Where a (possibly ghost) padded cell takes its values from.
struct GhostMap {
  int r, c;        // edge / source cell of h and tangential momentum and bed
  int mr, mc;      // mirror cell of normal momentum (wall)
  int kind;        // -1 interior, else the BC code of the side
  int side;
};

SWE_HD int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Map padded coordinates (r, c) -- interior rows/cols are [0, ny) x [0, nx) -- to
// their source cells.  Sides act on their own span only: left/right ghosts of
// interior rows, top/bottom ghosts of interior columns (boundary.py:301-411);
// corner ghosts never reach interior cells and are filled by replication.
SWE_HD GhostMap ghost_map(int r, int c, int ny, int nx, const int* bc) {
  GhostMap g;
  bool in_r = r >= 0 && r < ny, in_c = c >= 0 && c < nx;
  g.kind = -1;
  g.side = 0;
  if (in_r && in_c) {
    g.r = g.mr = r;
    g.c = g.mc = c;
    return g;
  }
  if (in_r && !in_c) {
    int side = c < 0 ? LEFT : RIGHT;
    int edge = c < 0 ? 0 : nx - 1;
    int kind = bc[side];
    g.side = side;
    g.kind = kind;
    g.r = g.mr = r;
    if (kind == BC_PERIODIC) {
      g.c = g.mc = clampi(c < 0 ? c + nx : c - nx, 0, nx - 1);
    } else {
      g.c = edge;
      g.mc = clampi(c < 0 ? -c - 1 : 2 * nx - 1 - c, 0, nx - 1);
    }
    return g;
  }
  if (in_c && !in_r) {
    int side = r < 0 ? TOP : BOTTOM;
    int edge = r < 0 ? 0 : ny - 1;
    int kind = bc[side];
    g.side = side;
    g.kind = kind;
    g.c = g.mc = c;
    if (kind == BC_PERIODIC) {
      g.r = g.mr = clampi(r < 0 ? r + ny : r - ny, 0, ny - 1);
    } else {
      g.r = edge;
      g.mr = clampi(r < 0 ? -r - 1 : 2 * ny - 1 - r, 0, ny - 1);
    }
    return g;
  }
  // corner (unused): replicate the nearest interior cell
  g.r = g.mr = clampi(r, 0, ny - 1);
  g.c = g.mc = clampi(c, 0, nx - 1);
  return g;
}

// Apply a side's ghost rule (boundary.py:279-328 and the mirrored blocks for the
// other sides) to the state fetched at the edge cell (h, hu, hv) and the
// normal momentum fetched at the mirror cell.
template <class T>
SWE_HD void apply_ghost(const GhostMap& g, T& h, T& hu, T& hv, T mirror_normal, T z_ghost, T level,
                                   const SweConsts<T>& k) {
  if (g.kind < 0) return;
  bool x_side = g.side == LEFT || g.side == RIGHT;
  switch (g.kind) {
    case BC_CONSTANT:
      h = hu = hv = k.cval;
      break;
    case BC_WATER_LEVEL:
      h = clamp_lo(level - z_ghost, T(0));
      break;
    case BC_WALL:
      if (x_side) hu = -mirror_normal;
      else hv = -mirror_normal;
      break;
    case BC_ABSORBING: {
      T h_ref = clamp_lo(level - z_ghost, T(0));
      T h_c = h_ref > T(1e-4) ? h_ref : h;
      T c = sqrt_rn(T(9.81) * clamp_lo(h_c, T(1e-6)));
      bool outward_pos = g.side == RIGHT || g.side == BOTTOM;
      T flow = outward_pos ? c * (h - h_ref) : (-c) * (h - h_ref);
      if (x_side) hu = flow;
      else hv = flow;
      break;
    }
    default:  // transmissive, periodic: values as fetched
      break;
  }
}

}  // namespace swe

namespace swe {

// Adjoint of apply_ghost: the adjoint (gh, ghu, ghv) of a ghost cell's raw
// state goes back to the edge cell (ge_*), the mirror cell's normal momentum
// (g_mirror) and the boundary level (g_level).  h_edge is the raw depth of the
// edge cell (the replicated value the absorbing rule reads).
template <class T>
SWE_HD void ghost_bwd(const GhostMap& g, T h_edge, T z_ghost, T level, T gh, T ghu, T ghv, const SweConsts<T>& k,
                      T& ge_h, T& ge_hu, T& ge_hv, T& g_mirror, T& g_level) {
  ge_h = ge_hu = ge_hv = g_mirror = g_level = T(0);
  bool x_side = g.side == LEFT || g.side == RIGHT;
  switch (g.kind) {
    case BC_CONSTANT:
      return;
    case BC_WATER_LEVEL:
      if (level - z_ghost >= T(0)) g_level = gh;
      ge_hu = ghu;
      ge_hv = ghv;
      return;
    case BC_WALL:
      ge_h = gh;
      if (x_side) {
        g_mirror = -ghu;
        ge_hv = ghv;
      } else {
        g_mirror = -ghv;
        ge_hu = ghu;
      }
      return;
    case BC_ABSORBING: {
      T gflow = x_side ? ghu : ghv;
      if (x_side) ge_hv = ghv;
      else ge_hu = ghu;
      ge_h = gh;
      T d = level - z_ghost;
      T h_ref = clamp_lo(d, T(0));
      bool use_ref = h_ref > T(1e-4);
      T h_c = use_ref ? h_ref : h_edge;
      T cc = clamp_lo(h_c, T(1e-6));
      T c = sqrt_rn(T(9.81) * cc);
      T diff = h_edge - h_ref;
      bool outward_pos = g.side == RIGHT || g.side == BOTTOM;
      T sgn = outward_pos ? T(1) : T(-1);
      T gc = sgn * gflow * diff;
      T gdiff = sgn * gflow * c;
      ge_h += gdiff;
      T g_href = -gdiff;
      T ghc = h_c >= T(1e-6) ? gc * T(9.81) / (T(2) * c) : T(0);
      if (use_ref) g_href += ghc;
      else ge_h += ghc;
      if (d >= T(0)) g_level = g_href;
      return;
    }
    default:  // transmissive, periodic
      ge_h = gh;
      ge_hu = ghu;
      ge_hv = ghv;
      return;
  }
}

}  // namespace swe
