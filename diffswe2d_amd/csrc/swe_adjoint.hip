// Hand-written adjoint of one time step on gfx950.
//
// The step U -> U' (solver.py:572-663) with dt held fixed (the reference
// detaches dt, :625-626) is reversed by
//   adj_L<2>       friction and the corrector clean-up reversed on each row as
//                  the sweep fetches it (-> lam_k2, and lam_acc on the
//                  wavefront's own cells), then L^T at the predictor
//                  P = clean(U, U + dt/2 k1)                 -> lam_P (+ ghosts)
//   adj_ghost      ghost-cell adjoints folded back onto the edge cells (four
//                  per-side launches) and the boundary-level reduction
//   adj_L<1>       the predictor clean-up reversed on each row as the sweep
//                  fetches it (-> lam_k1 and its direct adjoint of U), then
//                  L^T at U                                  -> lam_U = lam_acc + ...
//   adj_ghost      as above
// On grids that fill the GPU many times over, folding the corrector reverse
// into stage-2 L^T (recomputed on the halo cells, ~24 % more of it) saves a
// streaming pass over U, k2, the output adjoint and two stored adjoints: -2 %
// per reverse step on the 8192^2 basin.  Where a launch holds few wavefronts
// the sweep's time is its rows' dependent chain, which the fold lengthens
// (+6.5 % per epoch of the 68 x 42 Monai inversion, +3 % for the 544 x 344
// run-up), and adj_corrector, an elementwise launch, reverses it first.
// L^T mirrors the forward stage's decomposition: one wavefront streams down a
// strip of 64 columns, now with a 3-cell halo (58 outputs): a cell's adjoint
// depends on the four faces within two cells of it and those faces on the
// cells within three.  Forward values are recomputed from the stored state, so
// every branch (minmod selection, HLL region, wet/dry tests) is taken exactly
// as in the forward.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "swe_adjoint_math.h"
#include "swe_common.h"
#include "swe_host.h"
#include "swe_fetch.h"
#include "../../include/swe2d_hip.h"

namespace swe {

constexpr int kHaloA = 3;
constexpr int kOutA = kWave - 2 * kHaloA;  // 58
// L^T wavefronts per launch (times members) from which stage-2 L^T reverses the
// corrector itself: 8 per SIMD of the 1,024 (above that the launch is bound by
// issue, not by one wavefront's chain)
constexpr long long kFoldCorrectorWaves = 8192;

template <class T> struct AdjArgs {
  SweConsts<T> k;
  Fields<T> f;
  int strips, segs, S;   // tiling of the padded plane [-2, ny+2) x [-2, nx+2)
  int wsr, positivity;
  const T* U;            // state at step start
  const T* k1;           // first-stage tendency (stage 2 rebuilds the predictor from it)
  T half_dt;
  const T* lam;          // adjoint of the stage tendency [B,3,ny,nx] (unused when lam_P is set, and by
                         // stage 2 when it forms it from lam_out, k2 and U)
  const T* base;         // added to the interior output (nullable; required when lam_P is set)
  T* out;                // adjoint of the stage input state [B,3,ny,nx]
  T* ghost;              // [B][4 sides][2 layers][lmax][3]
  int lmax;
  // elementwise reverses
  const T* k2;           // second-stage tendency
  const T* lam_out;      // adjoint of the step output
  const T* lam_P;        // adjoint of the predictor state: when set, stage-1 L^T reverses the
                         // predictor clean-up itself (lam of each row from lam_P, k1 and U; the
                         // clean-up's direct adjoint of U added to its output)
  T dt;
  int friction, friction_model;
  T* acc;                // adjoint reaching U directly past the corrector (written by its reverse)
  T* part;               // per-block level-gradient partials of the ghost fold
  int part_blocks;       // blocks per side in the ghost fold
  T* bedgrad;            // adjoint of the bed, [B, ny, nx], accumulated (nullable)
};

template <class T> struct Lam3 {
  T h, u, v;
};

template <class T> __device__ inline Lam3<T> load_lam(const AdjArgs<T>& a, int b, int r, int c) {
  if (r < 0 || r >= a.f.ny || c < 0 || c >= a.f.nx) return Lam3<T>{T(0), T(0), T(0)};
  int64_t i = (int64_t)b * 3 * a.f.plane + (int64_t)r * a.f.nx + c;
  return Lam3<T>{a.lam[i], a.lam[i + a.f.plane], a.lam[i + 2 * a.f.plane]};
}

template <class T> __device__ inline SideGrad<T> shupg(const SideGrad<T>& g) {
  return SideGrad<T>{shup(g.hlin), shup(g.eta), shup(g.un), shup(g.ut), shup(g.hcell), shup(g.z), shup(g.zcell)};
}

// values per ghost slot: the adjoints of the ghost cell's h, hu, hv and bed
constexpr int kGhostVals = 4;

// ghost buffer slot of padded cell (r, c), or -1 for interior / corner / far cells
__device__ inline int64_t ghost_slot(int b, int r, int c, int ny, int nx, int lmax) {
  bool in_r = r >= 0 && r < ny, in_c = c >= 0 && c < nx;
  int side, layer, i;
  if (in_r && !in_c) {
    if (c == -1 || c == -2) {
      side = LEFT;
      layer = -c - 1;
    } else if (c == nx || c == nx + 1) {
      side = RIGHT;
      layer = c - nx;
    } else {
      return -1;
    }
    i = r;
  } else if (in_c && !in_r) {
    if (r == -1 || r == -2) {
      side = TOP;
      layer = -r - 1;
    } else if (r == ny || r == ny + 1) {
      side = BOTTOM;
      layer = r - ny;
    } else {
      return -1;
    }
    i = c;
  } else {
    return -1;
  }
  return ((((int64_t)b * 4 + side) * 2 + layer) * lmax + i) * kGhostVals;
}

// loads of the corrector reverse at padded cell (r, c): k2, the step-output adjoint
// and the roughness, issued ahead of their use
template <class T> struct CorrLoads {
  T kh, ku, kv, gh, gu, gv, coef;
  bool in;  // interior cell
};
// at interior cell `cell` of member b (state index i)
template <class T>
__device__ inline CorrLoads<T> corrector_loads_at(const AdjArgs<T>& a, int b, int64_t cell, int64_t i) {
  const int64_t P = a.f.plane;
  return CorrLoads<T>{a.k2[i], a.k2[i + P], a.k2[i + 2 * P], a.lam_out[i], a.lam_out[i + P], a.lam_out[i + 2 * P],
                      a.friction ? a.f.rough[b * a.f.rough_bstride + cell] : T(0), true};
}
template <class T> __device__ inline CorrLoads<T> corrector_loads(const AdjArgs<T>& a, int b, int r, int c) {
  if (!(r >= 0 && r < a.f.ny && c >= 0 && c < a.f.nx))
    return CorrLoads<T>{T(0), T(0), T(0), T(0), T(0), T(0), T(0), false};
  const int64_t cell = (int64_t)r * a.f.nx + c;
  return corrector_loads_at(a, b, cell, (int64_t)b * 3 * a.f.plane + cell);
}

// friction and the corrector clean-up reversed at one interior cell (solver.py:646-661)
// from its step-start state o* and the loads q: lam = adjoint of k2, acc = adjoint
// reaching U directly
template <class T>
__device__ inline void corrector_core(const AdjArgs<T>& a, T oh, T ou, T ov, const CorrLoads<T>& q, Lam3<T>& lam,
                                      Lam3<T>& acc) {
  const SweConsts<T>& k = a.k;
  const T dt = a.dt;
  T ch = oh + dt * q.kh, cu = ou + dt * q.ku, cv = ov + dt * q.kv;
  T gh = q.gh, gu = q.gu, gv = q.gv;
  if (a.friction) {
    T vh = ch, vu = cu, vv = cv;
    clean_and_limit(oh, ou, ov, vh, vu, vv, k, a.positivity != 0);
    friction_bwd(vh, vu, vv, q.coef, dt, k, a.friction_model, gh, gu, gv, gh, gu, gv);
  }
  T go_h = 0, go_u = 0, go_v = 0, gc_h, gc_u, gc_v;
  clean_and_limit_bwd(oh, ou, ov, ch, cu, cv, k, a.positivity != 0, gh, gu, gv, go_h, go_u, go_v, gc_h, gc_u, gc_v);
  acc = Lam3<T>{go_h + gc_h, go_u + gc_u, go_v + gc_v};
  lam = Lam3<T>{dt * gc_h, dt * gc_u, dt * gc_v};
}

// predictor clean-up reversed at one interior cell (solver.py:639) from its state o*,
// stage-1 tendency k* and the predictor-state adjoint g*
template <class T>
__device__ inline void predictor_core(const AdjArgs<T>& a, T oh, T ou, T ov, T kh, T ku, T kv, T gh, T gu, T gv,
                                      Lam3<T>& lam, Lam3<T>& acc) {
  const T half = a.half_dt;
  T ch = oh + half * kh, cu = ou + half * ku, cv = ov + half * kv;
  T go_h = 0, go_u = 0, go_v = 0, gc_h, gc_u, gc_v;
  clean_and_limit_bwd(oh, ou, ov, ch, cu, cv, a.k, a.positivity != 0, gh, gu, gv, go_h, go_u, go_v, gc_h, gc_u, gc_v);
  acc = Lam3<T>{go_h + gc_h, go_u + gc_u, go_v + gc_v};
  lam = Lam3<T>{half * gc_h, half * gc_u, half * gc_v};
}

// loads of the predictor reverse at padded cell (r, c), issued ahead of their use
template <class T> struct PredLoads {
  T kh, ku, kv, gh, gu, gv;
  bool in;  // interior cell (the stage tendency exists on interior cells only)
};
template <class T> __device__ inline PredLoads<T> predictor_loads(const AdjArgs<T>& a, int b, int r, int c) {
  PredLoads<T> p{T(0), T(0), T(0), T(0), T(0), T(0), r >= 0 && r < a.f.ny && c >= 0 && c < a.f.nx};
  if (p.in) {
    const int64_t P = a.f.plane, i = (int64_t)b * 3 * P + (int64_t)r * a.f.nx + c;
    p.kh = a.k1[i];
    p.ku = a.k1[i + P];
    p.kv = a.k1[i + 2 * P];
    p.gh = a.lam_P[i];
    p.gu = a.lam_P[i + P];
    p.gv = a.lam_P[i + 2 * P];
  }
  return p;
}

// one thread per interior cell (blockIdx.y: the member): lam_k2 = adjoint of k2,
// acc = adjoint reaching U
template <class T>
__global__ __launch_bounds__(256) void adj_corrector_kernel(AdjArgs<T> a, T* lam_k2) {
  const int64_t cell = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (cell >= a.f.plane) return;
  const int b = blockIdx.y;
  const int64_t i = (int64_t)b * 3 * a.f.plane + cell, P = a.f.plane;
  const CorrLoads<T> q = corrector_loads_at(a, b, cell, i);
  Lam3<T> lam, acc;
  corrector_core(a, a.U[i], a.U[i + P], a.U[i + 2 * P], q, lam, acc);
  a.acc[i] = acc.h;
  a.acc[i + P] = acc.u;
  a.acc[i + 2 * P] = acc.v;
  lam_k2[i] = lam.h;
  lam_k2[i + P] = lam.u;
  lam_k2[i + 2 * P] = lam.v;
}

// x sweep of one row (all lanes take part): adjoint of the primitives of each
// lane's cell from the x-faces of the row.
template <class T>
__device__ inline PrimGrad<T> x_sweep_bwd(const AdjArgs<T>& a, const Prim<T>& P, const Lam3<T>& L) {
  const SweConsts<T>& k = a.k;
  Cell<T> cx = P.x();
  Cell<T> cm = shup(cx), cp = shdn(cx);
  bool reset;
  Incr<T> xi = increments_r(cm, cx, cp, k, a.wsr != 0, reset);
  Side<T> plus = side_plus(cx, xi, k);
  Side<T> minus_next = shdn(side_minus(cx, xi, k));
  T lh1 = shdn(L.h), lu1 = shdn(L.u), lv1 = shdn(L.v);
  // face (c | c+1): normal momentum hu, tangential hv
  T gfh = gdiv_by(lh1 - L.h, k.ddx);
  T gfn = gdiv_by(lu1 - L.u, k.ddx);
  T gft = gdiv_by(lv1 - L.v, k.ddx);
  T gsL = gdiv_by(L.u, k.ddx);
  T gsR = gdiv_by(-lu1, k.ddx);
  SideGrad<T> gL, gR;
  face_flux_bwd(plus, cx.h, cx.z, minus_next, cp.h, cp.z, k, gfh, gfn, gft, gsL, gsR, gL, gR);
  SideGrad<T> gm = shupg(gR);  // this cell as the right state of face (c-1 | c)
  CellGrad<T> cg = sides_bwd(cx, xi, reset, gL, gm, k);
  PrimGrad<T> own{cg.h, cg.eta, cg.un, cg.ut, cg.z};
  PrimGrad<T> to_m{}, to_p{};
  increments_bwd(cm, cx, cp, cg, k, to_m, own, to_p);
  own.h += shdn(to_m.h) + shup(to_p.h);
  own.eta += shdn(to_m.eta) + shup(to_p.eta);
  own.un += shdn(to_m.un) + shup(to_p.un);
  own.ut += shdn(to_m.ut) + shup(to_p.ut);
  own.z += shdn(to_m.z) + shup(to_p.z);
  return own;
}

// FP32: 3 waves per SIMD (at most 168 VGPRs) instead of the 2 the unconstrained
// allocation (174) leaves; the few values this spills cost far less than the
// latency the third wave hides.  FP64 keeps its natural allocation.
template <class T> constexpr int adj_waves_per_eu() { return sizeof(T) == 4 ? 3 : 1; }

// BED: also accumulate the adjoint of the bed into a.bedgrad (interior cells) and
// the ghost slots' fourth value (ghost cells).
template <class T, int STAGE, bool BED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(adj_waves_per_eu<T>())))
void adj_L_kernel(AdjArgs<T> a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = blockIdx.x * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int b = blockIdx.y;
  const int strip = wid % a.strips;
  const int seg = wid / a.strips;
  if (seg >= a.segs) return;  // uniform per wavefront
  const Fields<T>& f = a.f;
  const SweConsts<T>& k = a.k;
  const int x0 = -2 + strip * kOutA;
  const int col = x0 - kHaloA + lane;
  const int y0 = -2 + seg * a.S;
  const int rows = min(a.S, f.ny + 2 - y0);
  const bool lane_out = lane >= kHaloA && lane < kWave - kHaloA && col < f.nx + 2;
  const T half = a.half_dt;
  // Stage 1 with the predictor reverse fused: the clean-up's direct adjoint of U
  // of the row fetched now waits in LDS (per lane, a ring four rows deep) until
  // the row is emitted three iterations later.
  __shared__ T pacc_ring[STAGE == 1 ? kWavesPerBlock : 1][4][3][kWave];
  T(*ring)[3][kWave] = pacc_ring[STAGE == 1 ? (threadIdx.x >> 6) : 0];

  Prim<T> P3{}, P2{}, P1{};          // rows n-3, n-2, n-1
  Lam3<T> L3{}, L2{}, L1{};          // adjoint of the tendency, same rows
  Incr<T> yi_prev{};                 // y increments of row n-2
  bool reset_prev = false;
  SideGrad<T> gm_pend{};             // row n-1 as the lower state of face (n-2 | n-1)
  SideGrad<T> gm_prev{};             // row n-2 as the lower state of face (n-3 | n-2)
  PrimGrad<T> A3{}, A2{}, A1{};      // y-sweep adjoints of primitives, rows n-3..n-1
  // Rows whose 64 cells all have depth exactly 0 and a zero adjoint of both
  // momentum tendencies: a face between two such rows' cells is dry on both
  // sides (its fluxes are the constant 0) and its topographic source terms meet a
  // zero adjoint, so it passes no gradient; the x sweep of such a row and such
  // y-faces are skipped (their adjoints are zeros).  dry1..dry3: rows n-1..n-3.
  bool dry1 = false, dry2 = false, dry3 = false;

  for (int it = 0; it < rows + 6; ++it) {
    const int n = y0 - kHaloA + it;
    // row n's loads go out first; the x sweep of row n-3 does not need them
    CellLoads<T> raw = fetch_loads<T, STAGE, AdjArgs<T>>(a, b, n, col);
    const bool fused_pred = STAGE == 1 && a.lam_P != nullptr;
    PredLoads<T> pl{};
    if (fused_pred) pl = predictor_loads(a, b, n, col);
    // stage 2 with lam_out set forms lam of each row itself: friction and the
    // corrector clean-up reversed from the step-output adjoint, k2 and U (on
    // interior cells the stage-2 fetch loaded the cell's own step-start state)
    const bool fused_corr = STAGE == 2 && a.lam_out != nullptr;
    PrimGrad<T> gx{};
    if (it >= 6 && n - 3 >= 0 && n - 3 < f.ny && !dry3) gx = x_sweep_bwd(a, P3, L3);  // x-faces: interior rows
    Prim<T> Pn = fetch_finish<T, STAGE, AdjArgs<T>>(a, b, raw, half);
    Lam3<T> Ln{T(0), T(0), T(0)};
    if (fused_pred) {
      // lam of row n: the predictor clean-up reversed (on interior cells the stage-1
      // fetch loaded the cell's own state)
      Lam3<T> pacc{T(0), T(0), T(0)};
      if (pl.in) predictor_core(a, raw.h, raw.hu, raw.hv, pl.kh, pl.ku, pl.kv, pl.gh, pl.gu, pl.gv, Ln, pacc);
      if constexpr (STAGE == 1) {
        const int slot = it & 3;
        ring[slot][0][lane] = pacc.h;
        ring[slot][1][lane] = pacc.u;
        ring[slot][2][lane] = pacc.v;
      }
    } else if (fused_corr) {
      const CorrLoads<T> cl = corrector_loads(a, b, n, col);
      if (cl.in) {
        Lam3<T> cacc;
        corrector_core(a, raw.h, raw.hu, raw.hv, cl, Ln, cacc);
        // the corrector's direct adjoint of U, once per cell: on the wavefront's own
        // rows and columns
        if (lane_out && n >= y0 && n < y0 + rows) {
          const int64_t i = (int64_t)b * 3 * f.plane + (int64_t)n * f.nx + col;
          a.acc[i] = cacc.h;
          a.acc[i + f.plane] = cacc.u;
          a.acc[i + 2 * f.plane] = cacc.v;
        }
      }
    } else {
      Ln = load_lam(a, b, n, col);
    }
    const bool dry0 = __all(Pn.h == T(0) && Ln.u == T(0) && Ln.v == T(0));
    Incr<T> yi_cur{};
    bool reset_cur = false;
    PrimGrad<T> An{};
    if (it >= 2) yi_cur = increments_r(P2.y(), P1.y(), Pn.y(), k, a.wsr != 0, reset_cur);  // row n-1
    if (it >= 3) {
      // face (n-2 | n-1): normal momentum hv, tangential hu
      Side<T> up = side_plus(P2.y(), yi_prev, k);
      Side<T> dn = side_minus(P1.y(), yi_cur, k);
      T gfh = gdiv_by(L1.h - L2.h, k.ddy);
      T gfn = gdiv_by(L1.v - L2.v, k.ddy);
      T gft = gdiv_by(L1.u - L2.u, k.ddy);
      T gsL = gdiv_by(L2.v, k.ddy);
      T gsR = gdiv_by(-L1.v, k.ddy);
      SideGrad<T> gL{}, gR{};
      if (!(dry3 && dry2 && dry1 && dry0))
        face_flux_bwd(up, P2.h, P2.z, dn, P1.h, P1.z, k, gfh, gfn, gft, gsL, gsR, gL, gR);
      // row n-2 has both face adjoints now
      CellGrad<T> cg = sides_bwd(P2.y(), yi_prev, reset_prev, gL, gm_prev, k);
      A2.h += cg.h;
      A2.eta += cg.eta;
      A2.un += cg.un;
      A2.ut += cg.ut;
      A2.z += cg.z;
      increments_bwd(P3.y(), P2.y(), P1.y(), cg, k, A3, A2, A1);
      gm_pend = gR;
    }
    if (it >= 6) {
      // row n-3 is complete in y; add its x sweep and emit
      const int r = n - 3;
      T gh = A3.h + gx.h, geta = A3.eta + gx.eta;
      T gu = A3.ut + gx.un;  // y sweep: un = v, ut = u
      T gv = A3.un + gx.ut;
      T lh, lu, lv;
      primitive_bwd(P3.hr, P3.qu, P3.qv, k, gh, gu, gv, geta, lh, lu, lv);
      const T gz = (A3.z + gx.z) + geta;  // eta = h + z
      if (lane_out) {
        if (r >= 0 && r < f.ny && col >= 0 && col < f.nx) {
          int64_t i = (int64_t)b * 3 * f.plane + (int64_t)r * f.nx + col;
          if (a.base != nullptr) {
            T bh = a.base[i], bu = a.base[i + f.plane], bv = a.base[i + 2 * f.plane];
            if constexpr (STAGE == 1) {
              if (a.lam_P != nullptr) {
                // the predictor clean-up's direct adjoint of U (formed when the row
                // was fetched), after the corrector's
                const int slot = (it + 1) & 3;
                bh += ring[slot][0][lane];
                bu += ring[slot][1][lane];
                bv += ring[slot][2][lane];
              }
            }
            lh += bh;
            lu += bu;
            lv += bv;
          }
          a.out[i] = lh;
          a.out[i + f.plane] = lu;
          a.out[i + 2 * f.plane] = lv;
          if (BED) a.bedgrad[(int64_t)b * f.plane + (int64_t)r * f.nx + col] += gz;
        } else {
          int64_t gs = ghost_slot(b, r, col, f.ny, f.nx, a.lmax);
          if (gs >= 0) {
            a.ghost[gs] = lh;
            a.ghost[gs + 1] = lu;
            a.ghost[gs + 2] = lv;
            a.ghost[gs + 3] = BED ? gz : T(0);
          }
        }
      }
    }
    // shift the pipeline
    P3 = P2;
    P2 = P1;
    P1 = Pn;
    L3 = L2;
    L2 = L1;
    L1 = Ln;
    A3 = A2;
    A2 = A1;
    A1 = An;
    yi_prev = yi_cur;
    reset_prev = reset_cur;
    gm_prev = gm_pend;
    dry3 = dry2;
    dry2 = dry1;
    dry1 = dry0;
  }
}

// Fold the ghost-cell adjoints of one side back onto the cells they were
// copied from and form per-block partials of the boundary-level gradient.
// One thread per position along the side (both ghost layers, in order); the
// four sides run as four launches in turn, so a cell reached from two sides
// (near a corner) is updated in a fixed order and no two threads of a launch
// touch the same cell.
// Folds the two ghost layers at position i along `side` of member b; returns
// their boundary-level adjoint.  The edge state is read only where the ghost
// rule depends on it (absorbing), the bed only where the level enters (water
// level, absorbing); both layers' inputs are read before the edge cells are
// updated, layer 0 first.
// The inputs of one position's fold: both layers' ghost adjoints and maps, and
// what their rules read (edge state for absorbing, bed and level where the level
// enters) -- none of which the fold itself changes.
template <class T> struct FoldIn {
  GhostMap g[2];
  T gh[2], gu[2], gv[2], gz[2], eh[2], z[2];
  T level;
};

template <class T>
__device__ inline FoldIn<T> ghost_fold_load(const AdjArgs<T>& a, int stage, int side, int b, int i) {
  const Fields<T>& f = a.f;
  const int64_t base = (int64_t)b * 3 * f.plane;
  FoldIn<T> in;
  for (int layer = 0; layer < 2; ++layer) {
    int r, c;
    if (side == LEFT) { r = i; c = -1 - layer; }
    else if (side == RIGHT) { r = i; c = f.nx + layer; }
    else if (side == TOP) { r = -1 - layer; c = i; }
    else { r = f.ny + layer; c = i; }
    in.g[layer] = ghost_map(r, c, f.ny, f.nx, f.bc);
    int64_t gs = ((((int64_t)b * 4 + side) * 2 + layer) * a.lmax + i) * kGhostVals;
    in.gh[layer] = a.ghost[gs];
    in.gu[layer] = a.ghost[gs + 1];
    in.gv[layer] = a.ghost[gs + 2];
    in.gz[layer] = a.ghost[gs + 3];
  }
  const int kind = in.g[0].kind;  // both layers belong to the same side
  const bool uses_level = kind == BC_WATER_LEVEL || kind == BC_ABSORBING;
  in.level = uses_level ? f.level[b * f.level_stride] : T(0);
  for (int layer = 0; layer < 2; ++layer) {
    const GhostMap& m = in.g[layer];
    T eh = T(0), eu, ev;
    if (kind == BC_ABSORBING) {
      if (stage == 2) load_state<T, 2, AdjArgs<T>>(a, base, m.r, m.c, a.half_dt, eh, eu, ev);
      else load_state<T, 1, AdjArgs<T>>(a, base, m.r, m.c, a.half_dt, eh, eu, ev);
    }
    in.eh[layer] = eh;
    in.z[layer] = uses_level ? f.bed[(int64_t)b * f.bed_bstride + (int64_t)m.r * f.nx + m.c] : T(0);
  }
  return in;
}

// Apply one position's fold (layer 0 first); returns its boundary-level adjoint.
template <class T>
__device__ inline T ghost_fold_apply(const AdjArgs<T>& a, int side, int b, const FoldIn<T>& in) {
  const Fields<T>& f = a.f;
  const int64_t base = (int64_t)b * 3 * f.plane;
  T glev_sum = T(0);
  for (int layer = 0; layer < 2; ++layer) {
    const GhostMap& m = in.g[layer];
    T ge_h, ge_u, ge_v, gmir, glev;
    ghost_bwd(m, in.eh[layer], in.z[layer], in.level, in.gh[layer], in.gu[layer], in.gv[layer], a.k, ge_h, ge_u,
              ge_v, gmir, glev);
    int64_t e = base + (int64_t)m.r * f.nx + m.c;
    a.out[e] += ge_h;
    a.out[e + f.plane] += ge_u;
    a.out[e + 2 * f.plane] += ge_v;
    if (gmir != T(0)) {
      int64_t mm = base + (int64_t)m.mr * f.nx + m.mc;
      if (side == LEFT || side == RIGHT) a.out[mm + f.plane] += gmir;
      else a.out[mm + 2 * f.plane] += gmir;
    }
    if (a.bedgrad != nullptr) {
      // a ghost cell's bed is its source cell's; the water-level and absorbing
      // rules read it through level - z, so it takes minus the level adjoint too
      a.bedgrad[(int64_t)b * f.plane + (int64_t)m.r * f.nx + m.c] += in.gz[layer] - glev;
    }
    glev_sum += glev;
  }
  return glev_sum;
}

// Folds the two ghost layers at position i along `side` of member b.
template <class T>
__device__ inline T ghost_fold_at(const AdjArgs<T>& a, int stage, int side, int b, int i) {
  return ghost_fold_apply(a, side, b, ghost_fold_load(a, stage, side, b, i));
}

template <class T> __device__ inline T block_sum(T v, T* part) {
  part[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  return part[0];
}

template <class T>
__global__ __launch_bounds__(256) void adj_ghost_side_kernel(AdjArgs<T> a, int stage, int side) {
  __shared__ T part[256];
  const int b = blockIdx.y;
  const int len = (side == LEFT || side == RIGHT) ? a.f.ny : a.f.nx;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  T glev = i < len ? ghost_fold_at(a, stage, side, b, i) : T(0);
  T sum = block_sum(glev, part);
  if (threadIdx.x == 0) a.part[((int64_t)b * 4 + side) * a.part_blocks + blockIdx.x] = sum;
}

// Small grids: all four sides in one block per member (sides in turn, separated
// by barriers), level adjoint summed in the block -- one launch instead of five.
// PREFETCH > 0 (sides of at most PREFETCH cells, blocks of at most PREFETCH
// threads): one position per thread on every side, all four sides' inputs loaded
// at once -- one memory round trip instead of four -- then the sides applied in
// turn; the thread bound leaves the loaded inputs their registers.
template <class T, int PREFETCH>
__global__ __launch_bounds__(PREFETCH > 0 ? PREFETCH : 1024) void adj_ghost_all_kernel(AdjArgs<T> a, int stage,
                                                                                    T* lam_level) {
  __shared__ T part[PREFETCH > 0 ? PREFETCH : 1024];
  const int b = blockIdx.x;
  T glev = T(0);
  if constexpr (PREFETCH > 0) {
    FoldIn<T> in[4];
    for (int side = 0; side < 4; ++side) {
      const int len = (side == LEFT || side == RIGHT) ? a.f.ny : a.f.nx;
      if ((int)threadIdx.x < len) in[side] = ghost_fold_load(a, stage, side, b, (int)threadIdx.x);
    }
    for (int side = 0; side < 4; ++side) {
      const int len = (side == LEFT || side == RIGHT) ? a.f.ny : a.f.nx;
      if ((int)threadIdx.x < len) glev += ghost_fold_apply(a, side, b, in[side]);
      __syncthreads();
    }
  } else {
    for (int side = 0; side < 4; ++side) {
      const int len = (side == LEFT || side == RIGHT) ? a.f.ny : a.f.nx;
      for (int i = threadIdx.x; i < len; i += blockDim.x) glev += ghost_fold_at(a, stage, side, b, i);
      __syncthreads();
    }
  }
  T sum = block_sum(glev, part);
  if (threadIdx.x == 0 && lam_level != nullptr) lam_level[b] += sum;
}

// lam_level[b] += sum of the member's ghost-fold partials, in a fixed order
template <class T>
__global__ __launch_bounds__(256) void adj_level_reduce_kernel(const T* part, int n_part, T* lam_level) {
  __shared__ T acc[256];
  const int b = blockIdx.x;
  T s = T(0);
  for (int j = threadIdx.x; j < n_part; j += blockDim.x) s += part[(int64_t)b * n_part + j];
  acc[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) acc[threadIdx.x] += acc[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lam_level[b] += acc[0];
}

// adjoint of a gauge observation h + z (member 0): lam_h += g at each gauge cell,
// gauges in order (two gauges may share a cell)
template <class T>
__global__ void inject_gauges_kernel(T* lam, int nx, const int32_t* gauges, int n_gauges, const T* g) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < n_gauges; ++i) lam[(int64_t)gauges[2 * i] * nx + gauges[2 * i + 1]] += g[i];
}

}  // namespace swe

// ===================================================================== host side
using namespace swe;

namespace {

thread_local char g_msg[256] = {0};

int afail(int code, const char* msg) {
  strncpy(g_msg, msg, sizeof(g_msg) - 1);
  return code;
}

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

constexpr int kGhostThreads = 256;
constexpr int kGhostOneBlock = 2048;  // longest side folded by one block per member

template <class T> struct AdjWs {
  T *lam_acc, *lam_P, *lam_k, *ghost, *part;  // lam_k: lam_k2 when the corrector runs alone
  T* bedgrad;                                 // [B, ny, nx], accumulated over calls
  int lmax, part_blocks;
};

inline int ghost_blocks(const swe_grid* g) {
  int lmax = g->ny > g->nx ? g->ny : g->nx;
  return (lmax + kGhostThreads - 1) / kGhostThreads;
}

// [lam_acc | lam_P | lam_k | ghost adjoints | level partials | bed adjoint],
// each 256-byte aligned
template <class T> size_t adj_bed_offset(const swe_grid* g) {
  size_t state = (size_t)g->batch * 3 * (size_t)g->ny * g->nx * sizeof(T);
  int lmax = g->ny > g->nx ? g->ny : g->nx;
  size_t ghost = (size_t)g->batch * 4 * 2 * lmax * kGhostVals * sizeof(T);
  size_t part = (size_t)g->batch * 4 * ghost_blocks(g) * sizeof(T);
  return 3 * al256(state) + al256(ghost) + al256(part);
}

template <class T> size_t adj_ws_bytes(const swe_grid* g) {
  return adj_bed_offset<T>(g) + al256((size_t)g->batch * (size_t)g->ny * g->nx * sizeof(T));
}

template <class T> AdjWs<T> adj_ws(void* ws, const swe_grid* g) {
  size_t state = al256((size_t)g->batch * 3 * (size_t)g->ny * g->nx * sizeof(T));
  int lmax = g->ny > g->nx ? g->ny : g->nx;
  size_t ghost = al256((size_t)g->batch * 4 * 2 * lmax * kGhostVals * sizeof(T));
  char* p = (char*)ws;
  AdjWs<T> w;
  w.bedgrad = (T*)(p + adj_bed_offset<T>(g));
  w.lam_acc = (T*)p;
  w.lam_P = (T*)(p + state);
  w.lam_k = (T*)(p + 2 * state);
  w.ghost = (T*)(p + 3 * state);
  w.part = (T*)(p + 3 * state + ghost);
  w.lmax = lmax;
  w.part_blocks = ghost_blocks(g);
  return w;
}

// L^T of one stage, with the bed adjoint when the call accumulates it
template <class T, int STAGE> void launch_adj_L(const AdjArgs<T>& a, dim3 grid, dim3 block, hipStream_t s) {
  if (a.bedgrad != nullptr)
    hipLaunchKernelGGL((adj_L_kernel<T, STAGE, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((adj_L_kernel<T, STAGE, false>), grid, block, 0, s, a);
}

// the ghost fold of one L^T sweep: four side launches, then the level reduction
template <class T>
void ghost_fold(const AdjArgs<T>& a, int stage, int batch, T* lam_level, hipStream_t s) {
  if (a.lmax <= kGhostOneBlock) {
    // as many threads as the longest side needs (a power of two for the block sum):
    // the fold's four phases and the sum then synchronise as few wavefronts as possible
    int threads = 64;
    while (threads < a.lmax && threads < 1024) threads <<= 1;
    // FP64 inputs of four sides take twice the registers: prefetch up to 512 positions
    constexpr int kPrefetchMax = sizeof(T) == 4 ? 1024 : 512;
    const dim3 grid(batch), block(threads);
    if (a.lmax <= 256)
      hipLaunchKernelGGL((adj_ghost_all_kernel<T, 256>), grid, block, 0, s, a, stage, lam_level);
    else if (a.lmax <= 512)
      hipLaunchKernelGGL((adj_ghost_all_kernel<T, 512>), grid, block, 0, s, a, stage, lam_level);
    else if (a.lmax <= kPrefetchMax)
      hipLaunchKernelGGL((adj_ghost_all_kernel<T, kPrefetchMax>), grid, block, 0, s, a, stage, lam_level);
    else
      hipLaunchKernelGGL((adj_ghost_all_kernel<T, 0>), grid, block, 0, s, a, stage, lam_level);
    return;
  }
  dim3 grid((unsigned)a.part_blocks, (unsigned)batch);
  for (int side = 0; side < 4; ++side)
    hipLaunchKernelGGL((adj_ghost_side_kernel<T>), grid, dim3(kGhostThreads), 0, s, a, stage, side);
  if (lam_level != nullptr)
    hipLaunchKernelGGL((adj_level_reduce_kernel<T>), dim3(batch), dim3(256), 0, s, a.part, 4 * a.part_blocks,
                       lam_level);
}


int adj_rows(const swe_grid* g, int hint) {
  if (hint > 0) return hint;
  int strips = (g->nx + 4 + kOutA - 1) / kOutA;
  long long per = (long long)strips * g->batch;
  int S = (int)((long long)(g->ny + 4) * per / 4096);
  // small grids are latency-bound: shorter strips, one row when that still gives
  // at most one wavefront per SIMD (1,024 on the chip)
  if (S < 2) S = (long long)(g->ny + 4) * per <= 1024 ? 1 : 2;
  if (S > 64) S = 64;
  return S;
}

template <class T>
int step_adjoint_impl(const swe_grid* g, const swe_config* c, const T* U, const T* k1, const T* k2, const T* bed,
                      const T* rough, const T* level, int level_stride, double dt_in, const T* lam_out, T* lam_U,
                      T* lam_level, void* ws, int rows_hint, void* stream) {
  if (!g || !c || !U || !k1 || !k2 || !bed || !lam_out || !lam_U || !ws) return afail(SWE_ERR_ARG, "null argument");
  if (g->batch <= 0 || g->ny < 2 || g->nx < 2) return afail(SWE_ERR_SHAPE, "bad grid");
  if (c->apply_friction && !rough) return afail(SWE_ERR_ARG, "roughness required");
  for (int s = 0; s < 4; ++s)
    if ((c->boundary[s] == BC_WATER_LEVEL || c->boundary[s] == BC_ABSORBING) && !level)
      return afail(SWE_ERR_ARG, "boundary level pointer is required");
  hipStream_t s = (hipStream_t)stream;
  AdjWs<T> w = adj_ws<T>(ws, g);
  AdjArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.k = host_consts<T>(c, g->dx, g->dy);
  a.f.B = g->batch;
  a.f.ny = g->ny;
  a.f.nx = g->nx;
  a.f.plane = (int64_t)g->ny * g->nx;
  a.f.bed_bstride = g->bed_batch == 1 ? 0 : a.f.plane;
  a.f.rough_bstride = g->rough_batch == 1 ? 0 : a.f.plane;
  for (int q = 0; q < 4; ++q) a.f.bc[q] = c->boundary[q];
  a.f.bed = bed;
  a.f.rough = rough;
  a.f.level = level;
  a.f.level_stride = level_stride;
  a.bedgrad = c->grad_bed ? w.bedgrad : nullptr;
  a.S = adj_rows(g, rows_hint);
  a.strips = (g->nx + 4 + kOutA - 1) / kOutA;
  a.segs = (g->ny + 4 + a.S - 1) / a.S;
  a.wsr = c->water_slope_reset;
  a.positivity = c->positivity_limiter;
  a.U = U;
  a.k1 = k1;
  T dt = (T)dt_in;
  T half = T(0.5) * dt;
  a.half_dt = half;
  a.ghost = w.ghost;
  a.lmax = w.lmax;
  a.part = w.part;
  a.part_blocks = w.part_blocks;
  long long waves = (long long)a.strips * a.segs;
  dim3 lg((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock), (unsigned)g->batch);
  dim3 lb(kWave * kWavesPerBlock);

  // stage 2: corrector + friction reversed, then L^T at the predictor
  AdjArgs<T> a2 = a;
  a2.k2 = k2;
  a2.lam_out = lam_out;
  a2.dt = dt;
  a2.friction = c->apply_friction;
  a2.friction_model = c->friction_model;
  a2.acc = w.lam_acc;
  if (waves * g->batch >= kFoldCorrectorWaves) {
    a2.lam = nullptr;  // stage-2 L^T forms it from lam_out, k2 and U
  } else {
    dim3 eb(256), eg((unsigned)((a.f.plane + 255) / 256), (unsigned)g->batch);
    hipLaunchKernelGGL((adj_corrector_kernel<T>), eg, eb, 0, s, a2, w.lam_k);
    a2.lam = w.lam_k;
    a2.lam_out = nullptr;  // stage-2 L^T reads lam instead
  }
  a2.base = nullptr;
  a2.out = w.lam_P;
  launch_adj_L<T, 2>(a2, lg, lb, s);
  ghost_fold(a2, 2, g->batch, lam_level, s);
  // stage 1: predictor clean-up reversed inside L^T at U (rows of lam formed from
  // lam_P, k1 and U as the sweep reaches them), plus the corrector's direct part
  AdjArgs<T> a1 = a;
  a1.lam_P = w.lam_P;
  a1.lam = nullptr;
  a1.base = w.lam_acc;
  a1.out = lam_U;
  launch_adj_L<T, 1>(a1, lg, lb, s);
  ghost_fold(a1, 1, g->batch, lam_level, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return afail(SWE_ERR_HIP, hipGetErrorString(e));
  return 0;
}

template <class T>
int operator_adjoint_impl(const swe_grid* g, const swe_config* c, const T* U, const T* bed, const T* level,
                          int level_stride, const T* lam_rhs, T* lam_U, T* lam_level, void* ws, int rows_hint,
                          void* stream) {
  if (!g || !c || !U || !bed || !lam_rhs || !lam_U || !ws) return afail(SWE_ERR_ARG, "null argument");
  if (g->batch <= 0 || g->ny < 2 || g->nx < 2) return afail(SWE_ERR_SHAPE, "bad grid");
  hipStream_t s = (hipStream_t)stream;
  AdjWs<T> w = adj_ws<T>(ws, g);
  AdjArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.k = host_consts<T>(c, g->dx, g->dy);
  a.f.B = g->batch;
  a.f.ny = g->ny;
  a.f.nx = g->nx;
  a.f.plane = (int64_t)g->ny * g->nx;
  a.f.bed_bstride = g->bed_batch == 1 ? 0 : a.f.plane;
  a.f.rough_bstride = 0;
  for (int q = 0; q < 4; ++q) a.f.bc[q] = c->boundary[q];
  a.f.bed = bed;
  a.f.rough = bed;
  a.f.level = level;
  a.f.level_stride = level_stride;
  a.bedgrad = c->grad_bed ? w.bedgrad : nullptr;
  a.S = adj_rows(g, rows_hint);
  a.strips = (g->nx + 4 + kOutA - 1) / kOutA;
  a.segs = (g->ny + 4 + a.S - 1) / a.S;
  a.wsr = c->water_slope_reset;
  a.positivity = c->positivity_limiter;
  a.U = U;
  a.k1 = U;
  a.lam = lam_rhs;
  a.out = lam_U;
  a.ghost = w.ghost;
  a.lmax = w.lmax;
  a.part = w.part;
  a.part_blocks = w.part_blocks;
  long long waves = (long long)a.strips * a.segs;
  dim3 lg((unsigned)((waves + kWavesPerBlock - 1) / kWavesPerBlock), (unsigned)g->batch);
  launch_adj_L<T, 1>(a, lg, dim3(kWave * kWavesPerBlock), s);
  ghost_fold(a, 1, g->batch, lam_level, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return afail(SWE_ERR_HIP, hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" int swe_step_record_f32(const swe_grid*, const swe_config*, const float*, const float*, const float*,
                                   const float*, int, double, int, float*, float*, float*, void*, int, void*);
extern "C" const char* swe_last_error_message(void);
extern "C" int swe_step_record_f64(const swe_grid*, const swe_config*, const double*, const double*, const double*,
                                   const double*, int, double, int, double*, double*, double*, void*, int, void*);

namespace {

inline int record_step_t(const swe_grid* g, const swe_config* c, const float* U, const float* bed, const float* rough,
                         const float* level, int ls, double dt, int parity, float* out, float* k1, float* k2, void* ws,
                         int rows, void* stream) {
  return swe_step_record_f32(g, c, U, bed, rough, level, ls, dt, parity, out, k1, k2, ws, rows, stream);
}
inline int record_step_t(const swe_grid* g, const swe_config* c, const double* U, const double* bed,
                         const double* rough, const double* level, int ls, double dt, int parity, double* out,
                         double* k1, double* k2, void* ws, int rows, void* stream) {
  return swe_step_record_f64(g, c, U, bed, rough, level, ls, dt, parity, out, k1, k2, ws, rows, stream);
}

// Reverse sweep over a recorded segment: Ustates holds the n_steps states the
// steps start from, K1 / K2 their stage tendencies (each B*3*ny*nx values).
template <class T>
int segment_reverse_impl(const swe_grid* g, const swe_config* c, const T* Ustates, const T* K1, const T* K2,
                         const T* bed, const T* rough, const T* level, int ls, const double* dts, int n_steps, T* lam,
                         T* lam_tmp, T* lam_level, void* adj_ws, int rows, void* stream) {
  if (!g || !c || !Ustates || !K1 || !K2 || !dts || !lam || !lam_tmp || !adj_ws || n_steps <= 0)
    return afail(SWE_ERR_ARG, "bad segment arguments");
  const size_t S = (size_t)g->batch * 3 * (size_t)g->ny * g->nx;
  T* cur = lam;
  T* nxt = lam_tmp;
  for (int k = n_steps - 1; k >= 0; --k) {
    int rc = step_adjoint_impl<T>(g, c, Ustates + k * S, K1 + k * S, K2 + k * S, bed, rough, level, ls, dts[k], cur,
                                  nxt, lam_level, adj_ws, rows, stream);
    if (rc) return rc;
    T* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != lam) {
    hipError_t e = hipMemcpyAsync(lam, cur, S * sizeof(T), hipMemcpyDeviceToDevice, (hipStream_t)stream);
    if (e != hipSuccess) return afail(SWE_ERR_HIP, hipGetErrorString(e));
  }
  return 0;
}

// Chunked runs (per-chunk boundary level, gauges observed at chunk ends): step k
// of the segment used levels[chunk[k]]; where ends_chunk[k] is set, the step's
// output was observed, so the observation adjoint g_obs[chunk[k]] enters lam
// before the step is reversed.  The level adjoint of step k accumulates into
// level_grad[chunk[k]] (member 0; batch 1).
template <class T>
int segment_reverse_chunked_impl(const swe_grid* g, const swe_config* c, const T* Ustates, const T* K1, const T* K2,
                                 const T* bed, const T* rough, const T* levels, const double* dts,
                                 const int32_t* chunk, const int32_t* ends_chunk, int n_steps, const int32_t* gauges,
                                 int n_gauges, const T* g_obs, T* lam, T* lam_tmp, T* level_grad, void* adj_ws,
                                 int rows, void* stream) {
  if (!g || !c || !Ustates || !K1 || !K2 || !levels || !dts || !chunk || !ends_chunk || !lam || !lam_tmp ||
      !level_grad || !adj_ws || n_steps <= 0 || (n_gauges > 0 && (!gauges || !g_obs)))
    return afail(SWE_ERR_ARG, "bad chunked segment arguments");
  if (g->batch != 1) return afail(SWE_ERR_SHAPE, "chunked runs have one member");
  const size_t S = (size_t)3 * (size_t)g->ny * g->nx;
  hipStream_t s = (hipStream_t)stream;
  T* cur = lam;
  T* nxt = lam_tmp;
  for (int k = n_steps - 1; k >= 0; --k) {
    if (ends_chunk[k] && n_gauges > 0) {
      hipLaunchKernelGGL((inject_gauges_kernel<T>), dim3(1), dim3(64), 0, s, cur, g->nx, gauges, n_gauges,
                         g_obs + (size_t)chunk[k] * n_gauges);
    }
    int rc = step_adjoint_impl<T>(g, c, Ustates + k * S, K1 + k * S, K2 + k * S, bed, rough, levels + chunk[k], 0,
                                  dts[k], cur, nxt, level_grad + chunk[k], adj_ws, rows, stream);
    if (rc) return rc;
    T* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != lam) {
    hipError_t e = hipMemcpyAsync(lam, cur, S * sizeof(T), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return afail(SWE_ERR_HIP, hipGetErrorString(e));
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : afail(SWE_ERR_HIP, hipGetErrorString(e));
}

template <class T>
int segment_adjoint_chunked_impl(const swe_grid* g, const swe_config* c, const T* U_ckpt, const T* bed,
                                 const T* rough, const T* levels, const double* dts, const int32_t* chunk,
                                 const int32_t* ends_chunk, int n_steps, int parity, const int32_t* gauges,
                                 int n_gauges, const T* g_obs, T* lam, T* lam_tmp, T* level_grad, T* tape,
                                 void* fwd_ws, void* adj_ws, int rows, void* stream) {
  if (!g || !c || !U_ckpt || !levels || !dts || !chunk || !tape || !fwd_ws || n_steps <= 0)
    return afail(SWE_ERR_ARG, "bad chunked segment arguments");
  const size_t S = (size_t)g->batch * 3 * (size_t)g->ny * g->nx;
  T* Ustates = tape;
  T* K1 = tape + (size_t)(n_steps + 1) * S;
  T* K2 = K1 + (size_t)n_steps * S;
  hipError_t e = hipMemcpyAsync(Ustates, U_ckpt, S * sizeof(T), hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e != hipSuccess) return afail(SWE_ERR_HIP, hipGetErrorString(e));
  for (int k = 0; k < n_steps; ++k) {
    int rc = record_step_t(g, c, Ustates + k * S, bed, rough, levels + chunk[k], 0, dts[k], (parity + k) & 1,
                           Ustates + (k + 1) * S, K1 + k * S, K2 + k * S, fwd_ws, rows, stream);
    if (rc) return afail(rc, swe_last_error_message());
  }
  return segment_reverse_chunked_impl<T>(g, c, Ustates, K1, K2, bed, rough, levels, dts, chunk, ends_chunk, n_steps,
                                         gauges, n_gauges, g_obs, lam, lam_tmp, level_grad, adj_ws, rows, stream);
}

// Replay one checkpoint segment onto the tape (U_k, k1_k, k2_k), then sweep it in reverse.
template <class T>
int segment_adjoint_impl(const swe_grid* g, const swe_config* c, const T* U_ckpt, const T* bed, const T* rough,
                         const T* level, int ls, const double* dts, int n_steps, int parity, T* lam, T* lam_tmp,
                         T* lam_level, T* tape, void* fwd_ws, void* adj_ws, int rows, void* stream) {
  if (!g || !c || !U_ckpt || !dts || !lam || !lam_tmp || !tape || !fwd_ws || !adj_ws || n_steps <= 0)
    return afail(SWE_ERR_ARG, "bad segment arguments");
  const size_t S = (size_t)g->batch * 3 * (size_t)g->ny * g->nx;
  T* Ustates = tape;
  T* K1 = tape + (size_t)(n_steps + 1) * S;
  T* K2 = K1 + (size_t)n_steps * S;
  hipError_t e = hipMemcpyAsync(Ustates, U_ckpt, S * sizeof(T), hipMemcpyDeviceToDevice, (hipStream_t)stream);
  if (e != hipSuccess) return afail(SWE_ERR_HIP, hipGetErrorString(e));
  for (int k = 0; k < n_steps; ++k) {
    int rc = record_step_t(g, c, Ustates + k * S, bed, rough, level, ls, dts[k], (parity + k) & 1,
                           Ustates + (k + 1) * S, K1 + k * S, K2 + k * S, fwd_ws, rows, stream);
    if (rc) return afail(rc, swe_last_error_message());
  }
  return segment_reverse_impl<T>(g, c, Ustates, K1, K2, bed, rough, level, ls, dts, n_steps, lam, lam_tmp, lam_level,
                                 adj_ws, rows, stream);
}

}  // namespace

extern "C" {

size_t swe_segment_tape_bytes(const swe_grid* g, int n_steps, int dtype_bytes) {
  if (!g || n_steps <= 0) return 0;
  return (size_t)(3 * n_steps + 1) * (size_t)g->batch * 3 * (size_t)g->ny * g->nx * (size_t)dtype_bytes;
}

int swe_segment_adjoint_f32(const swe_grid* g, const swe_config* c, const float* U_ckpt, const float* bed,
                            const float* rough, const float* level, int level_stride, const double* dts, int n_steps,
                            int parity, float* lam, float* lam_tmp, float* lam_level, float* tape, void* fwd_ws,
                            void* adj_ws, int rows_hint, void* stream) {
  return segment_adjoint_impl<float>(g, c, U_ckpt, bed, rough, level, level_stride, dts, n_steps, parity, lam,
                                     lam_tmp, lam_level, tape, fwd_ws, adj_ws, rows_hint, stream);
}

int swe_segment_adjoint_f64(const swe_grid* g, const swe_config* c, const double* U_ckpt, const double* bed,
                            const double* rough, const double* level, int level_stride, const double* dts,
                            int n_steps, int parity, double* lam, double* lam_tmp, double* lam_level, double* tape,
                            void* fwd_ws, void* adj_ws, int rows_hint, void* stream) {
  return segment_adjoint_impl<double>(g, c, U_ckpt, bed, rough, level, level_stride, dts, n_steps, parity, lam,
                                      lam_tmp, lam_level, tape, fwd_ws, adj_ws, rows_hint, stream);
}

size_t swe_adjoint_workspace_bytes(const swe_grid* g, int dtype_bytes) {
  if (!g || g->batch <= 0 || g->ny < 2 || g->nx < 2) return 0;
  return dtype_bytes == 8 ? adj_ws_bytes<double>(g) : adj_ws_bytes<float>(g);
}

const char* swe_adjoint_last_error_message(void) { return g_msg; }

size_t swe_adjoint_bed_grad_offset(const swe_grid* g, int dtype_bytes) {
  if (!g || g->batch <= 0 || g->ny < 2 || g->nx < 2) return 0;
  return dtype_bytes == 8 ? adj_bed_offset<double>(g) : adj_bed_offset<float>(g);
}

int swe_operator_adjoint_f32(const swe_grid* g, const swe_config* c, const float* U, const float* bed,
                             const float* level, int level_stride, const float* lam_rhs, float* lam_U,
                             float* lam_level, void* ws, int rows_hint, void* stream) {
  return operator_adjoint_impl<float>(g, c, U, bed, level, level_stride, lam_rhs, lam_U, lam_level, ws, rows_hint,
                                      stream);
}

int swe_operator_adjoint_f64(const swe_grid* g, const swe_config* c, const double* U, const double* bed,
                             const double* level, int level_stride, const double* lam_rhs, double* lam_U,
                             double* lam_level, void* ws, int rows_hint, void* stream) {
  return operator_adjoint_impl<double>(g, c, U, bed, level, level_stride, lam_rhs, lam_U, lam_level, ws, rows_hint,
                                       stream);
}

int swe_step_adjoint_f32(const swe_grid* g, const swe_config* c, const float* U, const float* k1, const float* k2,
                         const float* bed, const float* rough, const float* level, int level_stride, double dt,
                         const float* lam_out, float* lam_U, float* lam_level, void* ws, int rows_hint,
                         void* stream) {
  return step_adjoint_impl<float>(g, c, U, k1, k2, bed, rough, level, level_stride, dt, lam_out, lam_U, lam_level,
                                  ws, rows_hint, stream);
}

int swe_step_adjoint_f64(const swe_grid* g, const swe_config* c, const double* U, const double* k1,
                         const double* k2, const double* bed, const double* rough, const double* level,
                         int level_stride, double dt, const double* lam_out, double* lam_U, double* lam_level,
                         void* ws, int rows_hint, void* stream) {
  return step_adjoint_impl<double>(g, c, U, k1, k2, bed, rough, level, level_stride, dt, lam_out, lam_U, lam_level,
                                   ws, rows_hint, stream);
}

int swe_segment_reverse_f32(const swe_grid* g, const swe_config* c, const float* tape_U, const float* tape_k1,
                            const float* tape_k2, const float* bed, const float* rough, const float* level,
                            int level_stride, const double* dts, int n_steps, float* lam, float* lam_tmp,
                            float* lam_level, void* adj_ws, int rows_hint, void* stream) {
  return segment_reverse_impl<float>(g, c, tape_U, tape_k1, tape_k2, bed, rough, level, level_stride, dts, n_steps,
                                     lam, lam_tmp, lam_level, adj_ws, rows_hint, stream);
}

int swe_segment_reverse_f64(const swe_grid* g, const swe_config* c, const double* tape_U, const double* tape_k1,
                            const double* tape_k2, const double* bed, const double* rough, const double* level,
                            int level_stride, const double* dts, int n_steps, double* lam, double* lam_tmp,
                            double* lam_level, void* adj_ws, int rows_hint, void* stream) {
  return segment_reverse_impl<double>(g, c, tape_U, tape_k1, tape_k2, bed, rough, level, level_stride, dts, n_steps,
                                      lam, lam_tmp, lam_level, adj_ws, rows_hint, stream);
}

int swe_segment_reverse_chunked_f32(const swe_grid* g, const swe_config* c, const float* tape_U,
                                    const float* tape_k1, const float* tape_k2, const float* bed, const float* rough,
                                    const float* levels, const double* dts, const int32_t* chunk_of_step,
                                    const int32_t* ends_chunk, int n_steps, const int32_t* gauges, int n_gauges,
                                    const float* g_obs, float* lam, float* lam_tmp, float* level_grad, void* adj_ws,
                                    int rows_hint, void* stream) {
  return segment_reverse_chunked_impl<float>(g, c, tape_U, tape_k1, tape_k2, bed, rough, levels, dts, chunk_of_step,
                                             ends_chunk, n_steps, gauges, n_gauges, g_obs, lam, lam_tmp, level_grad,
                                             adj_ws, rows_hint, stream);
}

int swe_segment_reverse_chunked_f64(const swe_grid* g, const swe_config* c, const double* tape_U,
                                    const double* tape_k1, const double* tape_k2, const double* bed,
                                    const double* rough, const double* levels, const double* dts,
                                    const int32_t* chunk_of_step, const int32_t* ends_chunk, int n_steps,
                                    const int32_t* gauges, int n_gauges, const double* g_obs, double* lam,
                                    double* lam_tmp, double* level_grad, void* adj_ws, int rows_hint, void* stream) {
  return segment_reverse_chunked_impl<double>(g, c, tape_U, tape_k1, tape_k2, bed, rough, levels, dts,
                                              chunk_of_step, ends_chunk, n_steps, gauges, n_gauges, g_obs, lam,
                                              lam_tmp, level_grad, adj_ws, rows_hint, stream);
}

int swe_segment_adjoint_chunked_f32(const swe_grid* g, const swe_config* c, const float* U_ckpt, const float* bed,
                                    const float* rough, const float* levels, const double* dts,
                                    const int32_t* chunk_of_step, const int32_t* ends_chunk, int n_steps, int parity,
                                    const int32_t* gauges, int n_gauges, const float* g_obs, float* lam,
                                    float* lam_tmp, float* level_grad, float* tape, void* fwd_ws, void* adj_ws,
                                    int rows_hint, void* stream) {
  return segment_adjoint_chunked_impl<float>(g, c, U_ckpt, bed, rough, levels, dts, chunk_of_step, ends_chunk,
                                             n_steps, parity, gauges, n_gauges, g_obs, lam, lam_tmp, level_grad, tape,
                                             fwd_ws, adj_ws, rows_hint, stream);
}

int swe_segment_adjoint_chunked_f64(const swe_grid* g, const swe_config* c, const double* U_ckpt,
                                    const double* bed, const double* rough, const double* levels, const double* dts,
                                    const int32_t* chunk_of_step, const int32_t* ends_chunk, int n_steps, int parity,
                                    const int32_t* gauges, int n_gauges, const double* g_obs, double* lam,
                                    double* lam_tmp, double* level_grad, double* tape, void* fwd_ws, void* adj_ws,
                                    int rows_hint, void* stream) {
  return segment_adjoint_chunked_impl<double>(g, c, U_ckpt, bed, rough, levels, dts, chunk_of_step, ends_chunk,
                                              n_steps, parity, gauges, n_gauges, g_obs, lam, lam_tmp, level_grad,
                                              tape, fwd_ws, adj_ws, rows_hint, stream);
}

}  // extern "C"
