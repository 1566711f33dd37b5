// Ghost-aware fetch of a stencil cell, shared by the forward and adjoint kernels.
#pragma once

#include "swe_common.h"

namespace swe {

// One-lane shifts across the whole 64-lane wavefront.  DPP = true: single DPP
// moves in the VALU (wave_shr:1 = control 0x138: lane i reads lane i - 1;
// wave_shl:1 = 0x130: lane i reads lane i + 1); the lane with no source (0 for
// shup, 63 for shdn) reads 0 (bound_ctrl).  DPP = false: __shfl_up / __shfl_down
// by 1, an LDS-routed ds_bpermute that keeps the VALU free and leaves the
// sourceless lane its own value.  Each kernel takes the form measured faster for
// it.  The kernels' halo lanes absorb the sourceless lane's value: no stored
// result and no cross-lane reduction depends on a value shifted into lane 0 or
// lane 63.  64-bit values move as two 32-bit halves.
template <int CTRL> __device__ inline int dpp_shift(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xf, 0xf, true);
}
template <int CTRL, class T> __device__ inline T dpp_shift_value(T v) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, dpp_shift<CTRL>(__builtin_bit_cast(int, v)));
  } else {
    static_assert(sizeof(T) == 8, "32- or 64-bit lane values");
    long long w = __builtin_bit_cast(long long, v);
    int lo = dpp_shift<CTRL>((int)(w & 0xffffffffLL));
    int hi = dpp_shift<CTRL>((int)(w >> 32));
    return __builtin_bit_cast(T, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
  }
}
template <bool DPP = true, class T> __device__ inline T shup(T v) {
  if constexpr (DPP) return dpp_shift_value<0x138>(v);
  else return __shfl_up(v, 1, kWave);
}
template <bool DPP = true, class T> __device__ inline T shdn(T v) {
  if constexpr (DPP) return dpp_shift_value<0x130>(v);
  else return __shfl_down(v, 1, kWave);
}

template <bool DPP = true, class T> __device__ inline Cell<T> shup(const Cell<T>& c) {
  return Cell<T>{shup<DPP>(c.h), shup<DPP>(c.un), shup<DPP>(c.ut), shup<DPP>(c.eta), shup<DPP>(c.z)};
}
template <bool DPP = true, class T> __device__ inline Cell<T> shdn(const Cell<T>& c) {
  return Cell<T>{shdn<DPP>(c.h), shdn<DPP>(c.un), shdn<DPP>(c.ut), shdn<DPP>(c.eta), shdn<DPP>(c.z)};
}
template <bool DPP = true, class T> __device__ inline Side<T> shdn(const Side<T>& s) {
  return Side<T>{shdn<DPP>(s.hlin), shdn<DPP>(s.eta), shdn<DPP>(s.z), shdn<DPP>(s.un), shdn<DPP>(s.ut)};
}

// eta = h + z is formed where a Cell is built (the value primitive_cell gives),
// so the row pipelines carry one register less per row.
template <class T> struct Prim {
  T h, u, v, z;
  T hr, qu, qv;  // raw (ghost-applied) state the primitives came from
  __device__ Cell<T> x() const { return Cell<T>{h, u, v, h + z, z}; }
  __device__ Cell<T> y() const { return Cell<T>{h, v, u, h + z, z}; }
};

// state of interior cell (r, c) as stage STAGE sees it
template <class T, int STAGE, class Args, class A = RefRound>
__device__ inline void load_state(const Args& a, int64_t base, int r, int c, T half_dt, T& h, T& hu,
                                  T& hv) {
  int64_t i = base + (int64_t)r * a.f.nx + c;
  int64_t P = a.f.plane;
  h = a.U[i];
  hu = a.U[i + P];
  hv = a.U[i + 2 * P];
  if (STAGE == 2) {
    T ch = h + half_dt * a.k1[i];
    T cu = hu + half_dt * a.k1[i + P];
    T cv = hv + half_dt * a.k1[i + 2 * P];
    clean_and_limit<T, A>(h, hu, hv, ch, cu, cv, a.k, a.positivity != 0);
    h = ch;
    hu = cu;
    hv = cv;
  }
}

// A stencil cell's loads (its source cell's state, stage-1 tendency and bed),
// issued apart from the arithmetic on them so that a kernel can do unrelated
// work while they are in flight.
template <class T> struct CellLoads {
  GhostMap g;
  T h, hu, hv, k1h, k1u, k1v, z;
};

// Raw buffer access for interior rows.  A buffer resource (base, size) sits in
// scalar registers; a load takes the lane's byte offset in a vector register
// (fixed for the whole strip) and the row's byte offset in a scalar register,
// so moving to the next row is one scalar add instead of a 64-bit vector
// address per load and lane.  Word 3 of the resource is the gfx9 raw-dword
// format.  Callers use these only for rows and columns inside the arrays.
__device__ inline __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, unsigned long long bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)(unsigned)bytes, 0x00020000);
}
template <class T> __device__ inline T buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
  }
}
template <class T> __device__ inline void buf_store(T v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
  } else {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, (int)voff, (int)soff, 0);
  }
}

// One member's arrays as buffers, for wavefronts whose strip has no ghost column.
// usable: every array of the member fits the 32-bit buffer range.
template <class T> struct RowBufs {
  __amdgpu_buffer_rsrc_t U, k1, bed;
  unsigned plane_bytes, row_bytes;
  bool usable;
};
template <class T, class Args> __device__ inline RowBufs<T> row_bufs(const Args& a, int b) {
  const Fields<T>& f = a.f;
  RowBufs<T> q;
  const unsigned long long pb = (unsigned long long)f.plane * sizeof(T);
  q.usable = 3 * pb <= 0xffffffffull;
  q.plane_bytes = (unsigned)pb;
  q.row_bytes = (unsigned)f.nx * (unsigned)sizeof(T);
  q.U = buffer_rsrc(a.U + (int64_t)b * 3 * f.plane, 3 * pb);
  q.k1 = buffer_rsrc(a.k1 + (int64_t)b * 3 * f.plane, 3 * pb);
  q.bed = buffer_rsrc(f.bed + (int64_t)b * f.bed_bstride, pb);
  return q;
}

// strip_in: every lane's column of this wavefront is interior and the member's
// arrays fit buffers (wave-uniform); then an interior row needs no ghost rule
// and loads through the buffers q at lane byte offset voff.
template <class T, int STAGE, class Args>
__device__ inline CellLoads<T> fetch_loads(const Args& a, int b, int r, int c, bool strip_in = false,
                                           const RowBufs<T>* q = nullptr, unsigned voff = 0) {
  const Fields<T>& f = a.f;
  CellLoads<T> w;
  if (strip_in && r >= 0 && r < f.ny) {
    w.g.r = w.g.mr = r;
    w.g.c = w.g.mc = c;
    w.g.kind = -1;
    w.g.side = 0;
    const unsigned so = (unsigned)r * q->row_bytes;
    w.h = buf_load<T>(q->U, voff, so);
    w.hu = buf_load<T>(q->U, voff, so + q->plane_bytes);
    w.hv = buf_load<T>(q->U, voff, so + 2 * q->plane_bytes);
    if (STAGE == 2) {
      w.k1h = buf_load<T>(q->k1, voff, so);
      w.k1u = buf_load<T>(q->k1, voff, so + q->plane_bytes);
      w.k1v = buf_load<T>(q->k1, voff, so + 2 * q->plane_bytes);
    }
    w.z = buf_load<T>(q->bed, voff, so);
    return w;
  }
  w.g = ghost_map(r, c, f.ny, f.nx, f.bc);
  const int64_t i = (int64_t)b * 3 * f.plane + (int64_t)w.g.r * f.nx + w.g.c;
  const int64_t P = f.plane;
  w.h = a.U[i];
  w.hu = a.U[i + P];
  w.hv = a.U[i + 2 * P];
  if (STAGE == 2) {
    w.k1h = a.k1[i];
    w.k1u = a.k1[i + P];
    w.k1v = a.k1[i + 2 * P];
  }
  w.z = f.bed[(int64_t)b * f.bed_bstride + (int64_t)w.g.r * f.nx + w.g.c];
  return w;
}

// the cell from its loads: the predictor clean-up (stage 2), the ghost rule,
// the primitives -- the arithmetic of load_state and fetch, in their order
template <class T, int STAGE, class Args, class A = RefRound>
__device__ inline Prim<T> fetch_finish(const Args& a, int b, const CellLoads<T>& w, T half_dt) {
  const Fields<T>& f = a.f;
  const GhostMap& g = w.g;
  T h = w.h, hu = w.hu, hv = w.hv;
  if (STAGE == 2) {
    T ch = h + half_dt * w.k1h;
    T cu = hu + half_dt * w.k1u;
    T cv = hv + half_dt * w.k1v;
    clean_and_limit<T, A>(h, hu, hv, ch, cu, cv, a.k, a.positivity != 0);
    h = ch;
    hu = cu;
    hv = cv;
  }
  const T z = w.z;
  if (g.kind >= 0) {
    T mirror = T(0);
    if (g.kind == BC_WALL) {
      T mh, mu, mv;
      load_state<T, STAGE, Args, A>(a, (int64_t)b * 3 * f.plane, g.mr, g.mc, half_dt, mh, mu, mv);
      mirror = (g.side == LEFT || g.side == RIGHT) ? mu : mv;
    }
    T level = (g.kind == BC_WATER_LEVEL || g.kind == BC_ABSORBING) ? f.level[b * f.level_stride] : T(0);
    apply_ghost(g, h, hu, hv, mirror, z, level, a.k);
  }
  Prim<T> p;
  T eta;
  primitive_cell<T, A>(h, hu, hv, z, a.k, p.h, p.u, p.v, eta);
  p.z = z;
  p.hr = h;
  p.qu = hu;
  p.qv = hv;
  return p;
}

template <class T, int STAGE, class Args>
__device__ inline Prim<T> fetch(const Args& a, int b, int r, int c, T half_dt) {
  return fetch_finish<T, STAGE, Args>(a, b, fetch_loads<T, STAGE, Args>(a, b, r, c), half_dt);
}

}  // namespace swe
