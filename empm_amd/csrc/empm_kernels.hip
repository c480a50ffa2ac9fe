// libempm: the MLS-MPM substep of rai-opensource/empm and its adjoint as
// hand-written HIP kernels for gfx950 (MI355X), behind the C-ABI of
// include/empm.h.
//
// Per substep the forward runs:
//   k_bin_*  counting sort of the particles by stencil cell (two launches
//            after the previous G2P did the counting);
//   k_p2g    per particle block: APIC scatter of momentum, mass and the
//            stress impulse into block-owned LDS tiles;
//   k_grid   per node: gather of the covering tiles, normalise, gravity,
//            damping and every collider in registration order (a table
//            passed by value), writes v_out;
//   k_g2p    per particle: 27-node gather, x/v/C/F_trial update, then the
//            next substep's return map + stress (one SVD), so the scatter
//            carries no SVD (k_stress does it for a call's first substep).
// The adjoint of one substep uses the recorded grid, Fe and binning when
// the forward kept them (else recomputes them), then runs
//   k_g2p_adj (per particle) and k_g2p_adj_scatter (dL/dv_out into tiles),
//   k_grid_adj (per node), k_p2g_adj (gather of dL/d(mv, m), stress and
//   return-map adjoints).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "empm.h"
#include "empm_math.h"

using namespace empm;

namespace {

thread_local char g_err[512] = "";

int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return set_err(EMPM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr int kBlock = 64;  // per-particle launches: one wave per workgroup spreads the waves evenly over the SIMDs
constexpr int kWaveSize = 64;

// One simulation of a batch (a single rollout is a batch of one).  Its
// particles, blocks and grid nodes are the ranges [p0, p0 + N), [b0, b0 +
// nb^3) and [g0, g0 + n^3) of the batch's arenas; the per-particle arrays
// point at the whole particle arena and are indexed by the global particle.
struct DevModel {
  int N, n, material, nb;
  int p0, b0;
  int w0, pad0;  // first wave of this simulation in per-particle launches
  long long g0;
  float grid_lim, dx, inv_dx, rpic, damp_scale;
  float g[3];
  MatParams mp;
  float* mu;
  float* lam;
  float* ys;
  const float* mass;
  const float* vol;
  const int* sel;
};

// The simulations of a launch: a device table ordered by p0 (and b0, g0, w0).
struct Batch {
  const DevModel* cand;
  int ncand;
  int N;       // particles in the arena
  int nblk;    // blocks in the arena
  int nwaves;  // waves of a per-particle launch: sum over simulations of ceil(N / 64)
  // Per-material launches cover only that material's waves: entries
  // moff[m] .. moff[m + 1] - 1 of mlist are (simulation, its first wave in
  // material m's launch) for the non-empty simulations of material m, in
  // table order; mwaves[m] is that launch's wave count.
  const int2* mlist;
  int moff[9];
  int mwaves[8];
};

// The simulation owning global particle p / global block bg (binary search
// on the table; a batch of one returns at once).
EMPM_DEV const DevModel& cand_of_particle(const Batch& Bt, int p) {
  int lo = 0, hi = Bt.ncand - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (Bt.cand[mid].p0 <= p) lo = mid;
    else hi = mid - 1;
  }
  return Bt.cand[lo];
}
EMPM_DEV const DevModel& cand_of_block(const Batch& Bt, int bg) {
  int lo = 0, hi = Bt.ncand - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (Bt.cand[mid].b0 <= bg) lo = mid;
    else hi = mid - 1;
  }
  return Bt.cand[lo];
}

// Per-particle launches give each simulation its own whole waves (wave w of
// the launch belongs to the simulation with w0 <= w < w0 + ceil(N / 64)), so
// the simulation of a wave is uniform: its table entry is read with scalar
// loads.  With mat >= 0 the launch covers the waves of material `mat` only
// (Batch::mlist; launch with material_grid), else every simulation's waves
// (particle_grid).  body(M, p) runs for the thread's particle p.
template <typename F>
EMPM_DEV void per_particle(const Batch& Bt, int mat, F body) {
  const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) / kWaveSize));
  int ci, wl;
  if (mat >= 0) {
    if (w >= Bt.mwaves[mat]) return;
    int lo = Bt.moff[mat];
    for (int hi = Bt.moff[mat + 1] - 1; lo < hi;) {
      const int mid = (lo + hi + 1) >> 1;
      if (Bt.mlist[mid].y <= w) lo = mid;
      else hi = mid - 1;
    }
    const int2 e = Bt.mlist[lo];
    ci = e.x;
    wl = w - e.y;
  } else {
    if (w >= Bt.nwaves) return;
    int lo = 0;
    for (int hi = Bt.ncand - 1; lo < hi;) {
      const int mid = (lo + hi + 1) >> 1;
      if (Bt.cand[mid].w0 <= w) lo = mid;
      else hi = mid - 1;
    }
    ci = lo;
    wl = w - Bt.cand[lo].w0;
  }
  const DevModel& M = Bt.cand[__builtin_amdgcn_readfirstlane(ci)];
  const int k = __builtin_amdgcn_readfirstlane(wl) * kWaveSize + (int)(threadIdx.x % kWaveSize);
  if (k >= M.N) return;
  body(M, M.p0 + k);
}

// One substep's schedule entry: its time and collider table.  The node
// passes read it from device memory (the context's schedule buffer, uploaded
// once per call), so a captured launch sequence replays with new times and
// collider positions.
struct ColTable {
  float time;
  int n;
  int pad0, pad1;
  empm_collider c[EMPM_MAX_COLLIDERS];
};
static_assert(sizeof(ColTable) % 16 == 0, "ColTable is copied in 16-byte pieces");

// The node passes copy the substep's table into LDS once per workgroup of
// kBlock threads, its loads in flight together with the pass's first list
// loads, before their node loop: evaluating the colliders straight from
// device memory chains one scalar load per collider field per node.
constexpr int kTablePieces = (int)(sizeof(ColTable) / 16);
constexpr int kTableRounds = (kTablePieces + 63) / 64;
struct TableCopy {
  int4 v[kTableRounds];
};
EMPM_DEV TableCopy table_load(const ColTable* __restrict__ T) {
  TableCopy t;
  const int4* src = reinterpret_cast<const int4*>(T);
#pragma unroll
  for (int r = 0; r < kTableRounds; ++r) {
    const int i = (int)threadIdx.x + 64 * r;
    t.v[r] = i < kTablePieces ? src[i] : make_int4(0, 0, 0, 0);
  }
  return t;
}
EMPM_DEV void table_store(const TableCopy& t, ColTable& Ts) {  // every thread of the workgroup calls this
  int4* dst = reinterpret_cast<int4*>(&Ts);
#pragma unroll
  for (int r = 0; r < kTableRounds; ++r) {
    const int i = (int)threadIdx.x + 64 * r;
    if (i < kTablePieces) dst[i] = t.v[r];
  }
  __syncthreads();
}

// Quadratic B-spline stencil of one particle (mpm_utils.py:365-383).  The
// reference's w = wp.mat33(vec, vec, vec) takes the vectors as columns, so
// w[d][i] = N_i(f_d); base is wp.int(), i.e. truncation toward zero.
struct Stencil {
  int base[3];
  float f[3];
  float w[3][3];
  float dw[3][3];
};

EMPM_DEV void make_stencil(const float* xp, float inv_dx, Stencil& s) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float gp = xp[d] * inv_dx;
    s.base[d] = (int)(gp - 0.5f);
    const float f = gp - (float)s.base[d];
    s.f[d] = f;
    const float wa = 1.5f - f, wb = f - 1.f, wc = f - 0.5f;
    s.w[d][0] = wa * wa * 0.5f;
    s.w[d][1] = 0.75f - wb * wb;
    s.w[d][2] = wc * wc * 0.5f;
    s.dw[d][0] = f - 1.5f;
    s.dw[d][1] = -2.f * (f - 1.f);
    s.dw[d][2] = f - 0.5f;
  }
}

// Whether the 3x3x3 stencils of all active lanes lie inside the n^3 grid
// (almost always: positions are clamped two cells in from the bounds), so
// that their node loads need no bounds tests: node (a, b, c) is then at
// node_id(base) + a n^2 + b n + c.
EMPM_DEV bool wave_stencils_inside(const Stencil& s, int n) {
  const unsigned hi = (unsigned)(n - 3);
  const bool in = (unsigned)s.base[0] <= hi && (unsigned)s.base[1] <= hi && (unsigned)s.base[2] <= hi;
  return __ballot(!in) == 0;
}

// Weight lookups by a possibly run-time node index through selects, so the
// stencil stays in registers whether or not the node loops are unrolled.
EMPM_DEV float pick3(int i, const float* a) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }
EMPM_DEV float sw(const Stencil& s, int d, int i) { return pick3(i, s.w[d]); }
EMPM_DEV float sdw(const Stencil& s, int d, int i) { return pick3(i, s.dw[d]); }

EMPM_DEV bool node_in(int ix, int iy, int iz, int n) {
  return (unsigned)ix < (unsigned)n && (unsigned)iy < (unsigned)n && (unsigned)iz < (unsigned)n;
}
EMPM_DEV size_t node_id(int ix, int iy, int iz, int n) { return ((size_t)ix * n + iy) * n + iz; }

EMPM_DEV void load9(const float* __restrict__ a, size_t p, float* o) {
#pragma unroll
  for (int i = 0; i < 9; ++i) o[i] = a[9 * p + i];
}
EMPM_DEV void store9(float* __restrict__ a, size_t p, const float* o) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a[9 * p + i] = o[i];
}

// APIC affine matrix with the RPIC blend (mpm_utils.py:397-411)
EMPM_DEV void rpic_blend(const float* Cp, float r, float* Ct) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Ct[i * 3 + j] = (1.f - r) * Cp[i * 3 + j] + r / 2.f * (Cp[i * 3 + j] - Cp[j * 3 + i]);
  if (r < -0.001f) m_zero(Ct);
}

// ---------------------------------------------------------------------------
// Colliders, one node.  v is updated as the reference kernels do; fac tracks
// dv_out/dv_in per component (every collider op is a per-component overwrite
// or scale), which is exactly what the adjoint needs.
// ---------------------------------------------------------------------------
EMPM_DEV void v_zero(float* v, float* fac) {
  v[0] = v[1] = v[2] = 0.f;
  fac[0] = fac[1] = fac[2] = 0.f;
}
EMPM_DEV void v_cut(float* v, float* fac, int iz, float dx) {
  const float z = (float)iz * dx;
  if (z < 0.4f || z > 0.53f) {
    v_zero(v, fac);
  } else {
    v[0] *= 0.3f;
    v[1] = 0.f;
    v[2] *= 0.3f;
    fac[0] *= 0.3f;
    fac[1] = 0.f;
    fac[2] *= 0.3f;
  }
}
EMPM_DEV bool geo_inside(const empm_collider& c, const float* off) {
  if (c.geo_type == 1)
    return fabsf(off[0]) < c.size[0] && fabsf(off[1]) < c.size[1] && fabsf(off[2]) < c.size[2];
  if (c.geo_type == 2) return sqrtf(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]) < c.radius;
  if (c.geo_type == 3) {
    const float pr = off[0] * c.normal[0] + off[1] * c.normal[1] + off[2] * c.normal[2];
    const float r0 = off[0] - pr * c.normal[0], r1 = off[1] - pr * c.normal[1], r2 = off[2] - pr * c.normal[2];
    return fabsf(pr) < c.half_height && sqrtf(r0 * r0 + r1 * r1 + r2 * r2) < c.radius;
  }
  return false;
}

EMPM_DEV void collide(const empm_collider& c, int ix, int iy, int iz, int n, float time, float dt, float dx, float* v,
                      float* fac) {
  const bool win = time >= c.start_time && time < c.end_time;
  float off[3] = {(float)ix * dx - c.point[0], (float)iy * dx - c.point[1], (float)iz * dx - c.point[2]};
  switch (c.kind) {
    case 0: {  // add_surface_collider (mpm_solver_diff.py:473-528): sticky,
               // slip and separate all end in v = 0; type 11 is "cut"
      if (!win) return;
      const float d = off[0] * c.normal[0] + off[1] * c.normal[1] + off[2] * c.normal[2];
      if (d < 0.f) {
        if (c.surface_type == 11) v_cut(v, fac, iz, dx);
        else v_zero(v, fac);
      }
      return;
    }
    case 1: {  // set_velocity_on_cuboid (mpm_solver_diff.py:835-858)
      if (win) {
        if (fabsf(off[0]) < c.size[0] && fabsf(off[1]) < c.size[1] && fabsf(off[2]) < c.size[2]) {
          v[0] = c.velocity[0];
          v[1] = c.velocity[1];
          v[2] = c.velocity[2];
          fac[0] = fac[1] = fac[2] = 0.f;
        }
      } else if (c.reset == 1) {
        if (time < c.end_time + 15.f * dt) v_zero(v, fac);
      }
      return;
    }
    case 2: {  // add_bounding_box (mpm_solver_diff.py:878-935), padding 3
      if (!win) return;
      const int idx[3] = {ix, iy, iz};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (idx[d] < 3 && v[d] < 0.f) {
          v[d] = 0.f;
          fac[d] = 0.f;
        }
        if (idx[d] >= n - 3 && v[d] > 0.f) {
          v[d] = 0.f;
          fac[d] = 0.f;
        }
      }
      return;
    }
    case 3: {  // add_geometry_collider (mpm_solver_diff.py:612-672)
      if (!(win && c.active)) return;
      if (c.geo_type == 4) {
        const float d = off[0] * c.normal[0] + off[1] * c.normal[1] + off[2] * c.normal[2];
        if (d < 0.f) {
          if (c.surface_type == 11) v_cut(v, fac, iz, dx);
          else v_zero(v, fac);
        }
        return;
      }
      if (geo_inside(c, off)) v_zero(v, fac);
      return;
    }
    case 4: {  // set_grid_velocity_on_geometry (mpm_solver_diff.py:756-793)
      if (win && c.active) {
        const float av = c.angular_velocity;
        const float cr0 = c.rot_axis[1] * off[2] - c.rot_axis[2] * off[1];
        const float cr1 = c.rot_axis[2] * off[0] - c.rot_axis[0] * off[2];
        const float cr2 = c.rot_axis[0] * off[1] - c.rot_axis[1] * off[0];
        bool inside;
        if (c.geo_type == 4) inside = (off[0] * c.normal[0] + off[1] * c.normal[1] + off[2] * c.normal[2]) < 0.f;
        else inside = geo_inside(c, off);
        if (inside) {
          v[0] = c.velocity[0] + av * cr0;
          v[1] = c.velocity[1] + av * cr1;
          v[2] = c.velocity[2] + av * cr2;
          fac[0] = fac[1] = fac[2] = 0.f;
        }
      } else if (c.reset == 1) {
        if (time < c.end_time + 15.f * dt) v_zero(v, fac);
      }
      return;
    }
    case 5:  // enforce_grid_velocity_by_mask (mpm_solver_diff.py:1275-1286)
      if (c.mask[node_id(ix, iy, iz, n)] >= 1) v_zero(v, fac);
      return;
    default:
      return;
  }
}

// grid_normalization_and_gravity (mpm_utils.py:427-438) + add_damping_via_grid
// (mpm_utils.py:722-734) + colliders, for one node.
EMPM_DEV void grid_node(const DevModel& M, const ColTable& T, float dt, int ix, int iy, int iz, const float4 g,
                        float* v, float* fac) {
  const float time = T.time;
  v[0] = v[1] = v[2] = 0.f;
  fac[0] = fac[1] = fac[2] = 1.f;
  if (g.w > 1e-15f) {
    const float im = 1.f / g.w;
    v[0] = g.x * im + dt * M.g[0];
    v[1] = g.y * im + dt * M.g[1];
    v[2] = g.z * im + dt * M.g[2];
  }
  if (M.damp_scale < 1.f) {
    const float k = 1.f - M.damp_scale;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = v[d] - k * v[d];
      fac[d] *= M.damp_scale;
    }
  }
  for (int k = 0; k < T.n; ++k) collide(T.c[k], ix, iy, iz, M.n, time, dt, M.dx, v, fac);
}

// ---------------------------------------------------------------------------
// Block-owned scatter.
//
// Particles are binned every substep by the cell their 3x3x3 node stencil
// starts at, in blocks of 2x2x2 cells (block-major, cell-minor order).  One
// workgroup owns one block at a time: it scatters the stencils of the block's
// particles into a 4x4x4-node LDS tile (the block's stencil nodes), then
// writes the whole tile to its slot with plain stores.  The grid pass
// gathers, for each node, the <= 8 tiles that cover it.  No global atomics
// are issued.  Every in-grid node of a particle's stencil lies in its bin
// block's tile: the bin's cell is the stencil base clamped to the grid, so a
// base inside the grid puts the whole stencil in the tile, a base below 0
// leaves only nodes 0 and 1 in the grid (tile nodes 0..3 of block 0), and a
// base at n or above leaves none.  Stencil nodes outside the tile are thus
// outside the grid and are dropped.
//
// The waves of a workgroup take the block's particles in chunks of <= 64
// that never straddle a cell, so all lanes of a chunk share one stencil
// base.  Their per-node contributions are summed across the wave by a
// transposed butterfly (v_permlane32_swap, v_permlane16_swap, then DPP
// row/quad exchanges), leaving one LDS atomic per (node, component) per
// chunk.  A particle whose base differs from its bin's (one outside the
// grid, in a first substep) adds its own contributions lane by lane.
// ---------------------------------------------------------------------------
constexpr int kBlkCells = 2;                                    // cells per block edge
constexpr int kTileMargin = 0;                                  // binning is exact: no drift margin
constexpr int kTileN = kBlkCells + 2 + 2 * kTileMargin;         // 4 tile nodes per edge
constexpr int kTileNodes = kTileN * kTileN * kTileN;            // 64
constexpr int kBlkWG = 256;       // default workgroup size of the block-owned scatter kernels (tunable: 128, 256, 512)
constexpr int kBlkGridCap = 1024; // workgroups of a block-owned scatter launch when the occupancy query fails
constexpr int kHeavyChunks = 8;   // default: blocks with more chunks than this are scattered first
constexpr int kDynamicRounds = 4; // default: more blocks per scatter workgroup than this are handed out dynamically
constexpr int kWave = 64;
constexpr int kAdjPayload = 4;    // floats per particle of the G2P-adjoint scatter payload (dL/dv', pad)
constexpr int kSchedCap = 1024;   // substeps per schedule upload (longer calls run in chunks)
constexpr size_t kGraphCache = 8; // instantiated graphs kept per context (least recently used evicted)

// Particle binning of the current segment: perm[k] = k-th particle in
// block order, offs[b * 9 + c] = first position of cell c of block b and
// offs[b * 9 + 8] its end (entries of empty blocks are not written).  The
// blocks' ranges are in block order (empm_sort_by_cell), or in the order
// their workgroups allocated them (the per-substep binning).
struct Binning {
  const int* perm;
  const int* offs;
  // the non-empty blocks: act[0] heavy ones (more than the tuning's heavy_chunks chunks)
  // at act[2 ..], act[1] light ones at act[2 + nblk - 1 ...] downwards; the
  // scatter takes the heavy ones first
  const int* act;
  int nblk;
  const int* occ;      // per block: holds particles
  const int* near;     // per block: some non-empty block's tile covers a node this block owns
  const int* nlist;    // ids of the near blocks, in no particular order
  const int* nnear;    // their count (device scalar)
  int dyn_rounds;      // more active blocks per scatter workgroup than this: handed out by a work counter
};

// ---- cross-lane exchanges (wave64) ----------------------------------------
constexpr int kDppRowRor8 = 0x128;        // lane i <- lane (i + 8) mod 16 within each row of 16
constexpr int kDppRowHalfMirror = 0x141;  // lane i <- lane 7 - i within each group of 8
constexpr int kDppQuadXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppQuadXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]

template <int CTRL>
EMPM_DEV float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// One transposed step over pairs (v[i], v[i + H]), i < H: afterwards v[i] of
// the lanes in one half of each exchange pair holds the pair sum of the old
// v[i], and of the other half the pair sum of the old v[i + H].
template <int H>
EMPM_DEV void tstep_lanes32(float* v) {  // halves: lanes 0-31 keep v[i], 32-63 v[i + H]
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]), __float_as_uint(v[i + H]), false, false);
    v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}
template <int H>
EMPM_DEV void tstep_lanes16(float* v) {  // rows of 16: even rows keep v[i], odd rows v[i + H]
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[i]), __float_as_uint(v[i + H]), false, false);
    v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}
template <int H, int CTRL, int HIBIT>
EMPM_DEV void tstep_dpp(float* v, int lane) {  // lanes with HIBIT set keep v[i + H]
  const bool hi = (lane & HIBIT) != 0;
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const float send = hi ? v[i] : v[i + H];
    const float keep = hi ? v[i + H] : v[i];
    v[i] = keep + dpp_f<CTRL>(send);
  }
}

// Sum over the 64 lanes of NV (32 or 16) values per lane; on return lane L
// holds the wave total of value L >> 1 (NV = 32) or L >> 2 (NV = 16).
template <int NV>
EMPM_DEV float wave_transpose_sum(float* v, int lane) {
  static_assert(NV == 32 || NV == 16, "NV");
  tstep_lanes32<NV / 2>(v);
  tstep_lanes16<NV / 4>(v);
  tstep_dpp<NV / 8, kDppRowRor8, 8>(v, lane);
  tstep_dpp<NV / 16, kDppRowHalfMirror, 4>(v, lane);
  if constexpr (NV == 32) {
    tstep_dpp<1, kDppQuadXor2, 2>(v, lane);
  } else {
    v[0] += dpp_f<kDppQuadXor2>(v[0]);
  }
  return v[0] + dpp_f<kDppQuadXor1>(v[0]);
}

// What one particle adds to the nodes of its stencil: for node offset
// d = (a, b, c) - f, out[k] = w (A_k + B_k . d) + S_k . gw, with w the
// B-spline weight and gw its gradient without the inv_dx factor (folded into
// S by the producer); with NC = 4, out[3] = w * m.
struct ScatterPayload {
  float A[3], B[9], S[9], m;
};

template <int NC>
EMPM_DEV void node_contrib(const ScatterPayload& P, const Stencil& s, int a, int b, int c, float wa, float wb, float wc,
                           float dwa, float dwb, float dwc, float* out) {
  const float d0 = (float)a - s.f[0], d1 = (float)b - s.f[1], d2 = (float)c - s.f[2];
  const float wab = wa * wb;
  const float wt = wab * wc;
  const float g0 = dwa * wb * wc, g1 = wa * dwb * wc, g2 = wab * dwc;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    out[k] = wt * (P.A[k] + P.B[3 * k] * d0 + P.B[3 * k + 1] * d1 + P.B[3 * k + 2] * d2) + P.S[3 * k] * g0 +
             P.S[3 * k + 1] * g1 + P.S[3 * k + 2] * g2;
  if (NC == 4) out[3] = wt * P.m;
}

// The same contribution factored by stencil column: the terms shared by the
// three nodes (a, b, 0..2) are formed once, and each node then costs three
// operations per component: out[k] = wc (X_k + Y_k d2) + dwc Z_k,
// out[3] = wc Wm.
struct ColumnTerms {
  float X[3], Y[3], Z[3], Wm;
};

template <int NC>
EMPM_DEV ColumnTerms column_terms(const ScatterPayload& P, const Stencil& s, int a, int b, float wa, float wb,
                                  float dwa, float dwb) {
  const float d0 = (float)a - s.f[0], d1 = (float)b - s.f[1];
  const float wab = wa * wb, ga = dwa * wb, gb = wa * dwb;
  ColumnTerms t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t.X[k] = wab * (P.A[k] + P.B[3 * k] * d0 + P.B[3 * k + 1] * d1) + P.S[3 * k] * ga + P.S[3 * k + 1] * gb;
    t.Y[k] = wab * P.B[3 * k + 2];
    t.Z[k] = wab * P.S[3 * k + 2];
  }
  t.Wm = NC == 4 ? wab * P.m : 0.f;
  return t;
}

template <int NC>
EMPM_DEV void column_node(const ColumnTerms& t, float d2, float wc, float dwc, float* out) {
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] = wc * (t.X[k] + t.Y[k] * d2) + dwc * t.Z[k];
  if (NC == 4) out[3] = wc * t.Wm;
}

// Wave-level scatter of the lanes flagged `in_group` (all with stencil base =
// the chunk's cell base) into the LDS tile at tile-local node offset t0.
// Nodes are taken NPR per round (NPR * NC <= 32 values); every lane of the
// wave must call this.
template <int NC>
EMPM_DEV void group_scatter(float (*tile)[kTileNodes], const int* t0, bool in_group, const ScatterPayload& P,
                            const Stencil& s, int lane) {
  constexpr int NPR = NC == 4 ? 8 : 10;
  float w0[3], w1[3], w2[3], g0[3], g1[3], g2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    w0[i] = in_group ? s.w[0][i] : 0.f;
    w1[i] = s.w[1][i];
    w2[i] = s.w[2][i];
    g0[i] = in_group ? s.dw[0][i] : 0.f;
    g1[i] = s.dw[1][i];
    g2[i] = s.dw[2][i];
  }
#pragma unroll
  for (int n0 = 0; n0 < 27; n0 += NPR) {
    constexpr int kNV = 32;
    const int nn = (27 - n0) < NPR ? (27 - n0) : NPR;
    float v[kNV];
#pragma unroll
    for (int q = 0; q < kNV; ++q) v[q] = 0.f;
#pragma unroll
    for (int j = 0; j < NPR; ++j) {
      const int node = n0 + j;
      if (node < 27) {
        const int a = node / 9, b = (node / 3) % 3, c = node % 3;
        float o[4];
        const ColumnTerms t = column_terms<NC>(P, s, a, b, w0[a], w1[b], g0[a], g1[b]);
        column_node<NC>(t, (float)c - s.f[2], w2[c], g2[c], o);
#pragma unroll
        for (int k = 0; k < NC; ++k) v[j * NC + k] = o[k];
      }
    }
    float tot;
    int idx;
    if (NC == 4 && nn * NC <= 16) {
      tot = wave_transpose_sum<16>(v, lane);
      idx = (lane & 3) == 0 ? (lane >> 2) : 99;
    } else {
      tot = wave_transpose_sum<32>(v, lane);
      idx = (lane & 1) == 0 ? (lane >> 1) : 99;
    }
    if (idx < nn * NC) {
      const int node = n0 + idx / NC, comp = idx % NC;
      const int a = node / 9, b = (node / 3) % 3, c = node % 3;
      atomicAdd(&tile[comp][((t0[0] + a) * kTileN + (t0[1] + b)) * kTileN + (t0[2] + c)], tot);
    }
  }
}

// One lane's own 27-node scatter into the tile (o: the tile's origin), for a
// particle whose stencil does not start at its bin's cell: nodes outside the
// tile are outside the grid and are dropped.
template <int NC>
EMPM_DEV void lane_scatter(float (*tile)[kTileNodes], const int* o, const ScatterPayload& P, const Stencil& s) {
#pragma unroll 1
  for (int j = 0; j < 27; ++j) {
    const int a = j / 9, b = (j / 3) % 3, c = j % 3;
    const int tx = s.base[0] + a - o[0], ty = s.base[1] + b - o[1], tz = s.base[2] + c - o[2];
    if ((unsigned)tx >= (unsigned)kTileN || (unsigned)ty >= (unsigned)kTileN || (unsigned)tz >= (unsigned)kTileN)
      continue;
    float out[4];
    node_contrib<NC>(P, s, a, b, c, sw(s, 0, a), sw(s, 1, b), sw(s, 2, c), sdw(s, 0, a), sdw(s, 1, b), sdw(s, 2, c),
                     out);
    const int t = (tx * kTileN + ty) * kTileN + tz;
#pragma unroll
    for (int k = 0; k < NC; ++k) atomicAdd(&tile[k][t], out[k]);
  }
}

// Block id -> tile origin (node coordinates of tile node 0)
EMPM_DEV void tile_origin(int blk, int nb, int* o) {
  o[2] = (blk % nb) * kBlkCells - kTileMargin;
  o[1] = ((blk / nb) % nb) * kBlkCells - kTileMargin;
  o[0] = (blk / (nb * nb)) * kBlkCells - kTileMargin;
}

// The block-owned scatter driver.  load(p, xp, P) fills a particle's position
// and payload.  Each active block is taken by one workgroup, which draws its
// next block from the launch's work counter (zero at launch); its waves take
// the block's cell chunks round robin.  Launched with as many workgroups as
// fit on the device at once, so no workgroup waits for another to finish
// before it starts.
// DETW > 0 (deterministic mode, DETW = the workgroup's waves): each wave sums
// into a tile of its own, the lanes of particles scattered lane by lane take
// turns in lane order, and the wave tiles are added in wave order, so every
// float sum has an order fixed by the chunks (the binning orders a cell's
// particles by index in this mode) instead of by the LDS atomics' arrival.
template <int NC, int DETW, typename Load>
EMPM_DEV void block_scatter(const Batch& Bt, const Binning& B, float4* __restrict__ tiles, int* __restrict__ work,
                            Load load) {
  constexpr int NT = DETW > 0 ? DETW : 1;
  __shared__ float tile[NT][NC][kTileNodes];
  __shared__ int next[2][2];
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave), nwaves = blockDim.x / kWave;
  const int last = B.nblk - 1;
  // List entry w goes to workgroup w (read ahead of the counts, as if
  // heavy).  With up to B.dyn_rounds entries per workgroup the rest are
  // taken in strides of gridDim.x, the next id read at the top of each
  // round.  With more, workgroups take entries gridDim.x + (work counter)
  // one at a time: thread 0 requests the next entry and resolves its block
  // id while the current block is scattered (a returning device atomic
  // takes microseconds) and hands both over through alternating LDS slots.
  int ai = blockIdx.x;
  const int first = B.act[2 + (ai < last ? ai : last)];
  const int nh = B.act[0], nact = nh + B.act[1];
  const bool dynamic = nact > B.dyn_rounds * (int)gridDim.x;
  const auto entry = [&](int k) { return k < nh ? B.act[2 + k] : B.act[2 + last - (k - nh)]; };
  int blk = ai < nh ? first : (ai < nact ? entry(ai) : 0);
  for (int it = 0; ai < nact; ++it) {
    int fetched = 0, next_static = 0;
    if (dynamic) {
      if (threadIdx.x == 0) fetched = (int)gridDim.x + atomicAdd(work, 1);
    } else {
      next_static = ai + (int)gridDim.x < nact ? entry(ai + (int)gridDim.x) : 0;
    }
    blk = __builtin_amdgcn_readfirstlane(blk);
    int ci = 0;
    for (int hi = Bt.ncand - 1; ci < hi;) {
      const int mid = (ci + hi + 1) >> 1;
      if (Bt.cand[mid].b0 <= blk) ci = mid;
      else hi = mid - 1;
    }
    const DevModel& M = Bt.cand[__builtin_amdgcn_readfirstlane(ci)];
    const int* __restrict__ sel = M.sel;
    const float inv_dx = M.inv_dx;
    int o[3];
    tile_origin(blk - M.b0, M.nb, o);
    for (int i = threadIdx.x; i < NT * NC * kTileNodes; i += blockDim.x) (&tile[0][0][0])[i] = 0.f;
    __syncthreads();
    float(*wt)[kTileNodes] = tile[DETW > 0 ? wave : 0];
    // the block's cell ranges and the first chunk of each cell (kc[8] =
    // chunks in all); wave-uniform, so the chunk -> cell walk below is a
    // chain of scalar selects
    int cs[9], kc[9];
#pragma unroll
    for (int c = 0; c <= 8; ++c) cs[c] = B.offs[blk * 9 + c];
    kc[0] = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) kc[c + 1] = kc[c] + (cs[c + 1] - cs[c] + kWave - 1) / kWave;
    const int nch = kc[8];
    for (int ch = wave; ch < nch; ch += nwaves) {
      // chunk ch -> cell c (the last cell whose first chunk is <= ch), its
      // range [cb, ce) and the chunk's first particle
      int c = 0, cb = cs[0], ce = cs[1], c0 = 0;
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const bool past = ch >= kc[k];
        c = past ? k : c;
        cb = past ? cs[k] : cb;
        ce = past ? cs[k + 1] : ce;
        c0 = past ? kc[k] : c0;
      }
      const int i = cb + (ch - c0) * kWave + lane;
      const int p = i < ce ? B.perm[i] : -1;
      const bool active = p >= 0 && !(sel && sel[p] != 0);
      // the cell's stencil base and its tile-local offset
      const int t0[3] = {kTileMargin + ((c >> 2) & 1), kTileMargin + ((c >> 1) & 1), kTileMargin + (c & 1)};
      float xp[3] = {(o[0] + t0[0] + 1.f) * (1.f / inv_dx), (o[1] + t0[1] + 1.f) * (1.f / inv_dx),
                     (o[2] + t0[2] + 1.f) * (1.f / inv_dx)};
      ScatterPayload P = {};
      if (active) load(p, xp, P, M);
      Stencil s;
      make_stencil(xp, inv_dx, s);
      const bool home = active && s.base[0] == o[0] + t0[0] && s.base[1] == o[1] + t0[1] && s.base[2] == o[2] + t0[2];
      if (__ballot(home)) group_scatter<NC>(wt, t0, home, P, s, lane);
      if constexpr (DETW > 0) {
        for (unsigned long long m = __ballot(active && !home); m; m &= m - 1)
          if (lane == __builtin_ctzll(m)) lane_scatter<NC>(wt, o, P, s);
      } else if (active && !home) {
        lane_scatter<NC>(wt, o, P, s);
      }
    }
    int fetched_blk = 0;
    if (dynamic && threadIdx.x == 0 && fetched < nact) fetched_blk = entry(fetched);
    __syncthreads();
    float4* dst = tiles + (size_t)blk * kTileNodes;
    for (int k = threadIdx.x; k < kTileNodes; k += blockDim.x) {
      float4 t = make_float4(tile[0][0][k], tile[0][1][k], tile[0][2][k], NC == 4 ? tile[0][NC - 1][k] : 0.f);
#pragma unroll
      for (int w = 1; w < NT; ++w) {
        t.x += tile[w][0][k];
        t.y += tile[w][1][k];
        t.z += tile[w][2][k];
        if (NC == 4) t.w += tile[w][NC - 1][k];
      }
      dst[k] = t;
    }
    if (dynamic && threadIdx.x == 0) {
      next[it & 1][0] = fetched;
      next[it & 1][1] = fetched_blk;
    }
    __syncthreads();
    if (dynamic) {
      ai = next[it & 1][0];
      blk = next[it & 1][1];
    } else {
      ai += (int)gridDim.x;
      blk = next_static;
    }
  }
}

// Sum of the tiles at node (ix, iy, iz), a node of a near block (the node
// passes visit no other).  The node lies in the tiles of blocks
// b = ((i - 2) >> 1) + {0, 1} per axis (a tile spans nodes 2b .. 2b + 3);
// all 8 occupancy flags and tile entries are loaded before any is used, so
// the gather costs one memory latency.
EMPM_DEV float4 gather_node(const Binning& B, const DevModel& M, const float4* __restrict__ tiles, int ix, int iy,
                            int iz) {
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  const int nb = M.nb;
  const int l0 = (ix - 2) >> 1, l1 = (iy - 2) >> 1, l2 = (iz - 2) >> 1;
  bool ok[8];
  int blk[8];
  size_t k[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int b0 = l0 + (j >> 2), b1 = l1 + ((j >> 1) & 1), b2 = l2 + (j & 1);
    ok[j] = (unsigned)b0 < (unsigned)nb && (unsigned)b1 < (unsigned)nb && (unsigned)b2 < (unsigned)nb;
    blk[j] = ok[j] ? M.b0 + (b0 * nb + b1) * nb + b2 : 0;
    const int lx = ix - 2 * b0, ly = iy - 2 * b1, lz = iz - 2 * b2;
    k[j] = ok[j] ? (size_t)blk[j] * kTileNodes + (lx * kTileN + ly) * kTileN + lz : 0;
  }
  int occ[8];
  float4 t[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    occ[j] = B.occ[blk[j]];
    t[j] = tiles[k[j]];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const bool use = ok[j] && occ[j];
    acc.x += use ? t[j].x : 0.f;
    acc.y += use ? t[j].y : 0.f;
    acc.z += use ? t[j].z : 0.f;
    acc.w += use ? t[j].w : 0.f;
  }
  return acc;
}

// Bin of a particle of simulation M at position xp: (M.b0 + block) * 8 +
// cell, the cell being the one its stencil starts at (clamped to the grid).
EMPM_DEV int bin_key(const float* xp, const DevModel& M) {
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int v = (int)(xp[d] * M.inv_dx - 0.5f);
    c[d] = v < 0 ? 0 : (v >= M.n ? M.n - 1 : v);
  }
  const int nb = M.nb;
  const int blk = M.b0 + ((c[0] >> 1) * nb + (c[1] >> 1)) * nb + (c[2] >> 1);
  return blk * 8 + (((c[0] & 1) << 2) | ((c[1] & 1) << 1) | (c[2] & 1));
}

// Counting step of the binning for the calling lanes: count[key] grows by
// one per active lane and key_rank[p] = (key, the lane's rank in its bin).
// Lanes are grouped by key first (ballots only), then every group leader
// issues its atomic in the same instruction: one atomic round trip per wave.
EMPM_DEV void count_bin(bool active, int key, int p, int* __restrict__ count, int2* __restrict__ key_rank) {
  const int lane = __lane_id();
  unsigned long long pending = __ballot(active);
  int my_leader = lane, my_rank = 0, group_size = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lkey = __shfl(key, leader);
    const bool in = ((pending >> lane) & 1ull) && key == lkey;
    const unsigned long long grp = __ballot(in);
    pending &= ~grp;
    if (in) {
      my_leader = leader;
      my_rank = __popcll(grp & ((1ull << lane) - 1ull));
    }
    if (lane == leader) group_size = __popcll(grp);
  }
  int base = 0;
  if (group_size > 0) base = atomicAdd(&count[key], group_size);
  base = __shfl(base, my_leader);
  if (active) key_rank[p] = make_int2(key, base + my_rank);
}

// ---------------------------------------------------------------------------
// Forward kernels
// ---------------------------------------------------------------------------
// compute_stress_from_F_trial (mpm_utils.py:603-665) for one particle: the
// return map writes Fe (and, for material 5 / hardening, mu, lam, yield
// stress in place); S = -dt V tau inv_dx is the stress part of the P2G
// payload.
// Whether a simulation's return map rewrites its yield stress, mu and lam
// (material 5's softening, hardening == 1 of the von Mises maps).
template <int MAT>
EMPM_DEV bool mutates_params(const DevModel& M) {
  return MAT == 5 || ((MAT == 1 || MAT == 7) && M.mp.hardening == 1.f);
}

// A material-state record (empm_particles.mstate): yield stress, mu and lam,
// NA floats each (NA = the arena's particle count).
template <int MAT>
EMPM_DEV void particle_stress(const DevModel& M, float dt, int p, const float* Ftr, float* Fe_out,
                              float* __restrict__ S_out, int allow_write, int NA, const float* mst_in,
                              float* mst_rec) {
  // material 5 and hardening rewrite the model arrays in place (forward only);
  // their values at this substep's start come from mst_in when given (a
  // replay's first substep) and are recorded into mst_rec when given
  const bool mut = mutates_params<MAT>(M);
  const bool write_params = allow_write && mut;
  float mu, lam, ys;
  if (mut && mst_in) {
    ys = mst_in[p];
    mu = mst_in[(size_t)NA + p];
    lam = mst_in[2 * (size_t)NA + p];
  } else {
    mu = M.mu[p];
    lam = M.lam[p];
    ys = M.ys[p];
  }
  if (mut && mst_rec) {
    mst_rec[p] = ys;
    mst_rec[(size_t)NA + p] = mu;
    mst_rec[2 * (size_t)NA + p] = lam;
  }
  MatParams matp = M.mp;
  matp.material = MAT;  // compile-time material: one code path, constitutive state stays in registers
  ConstitState st;
  float tau[9];
  constitutive_fwd(st, Ftr, mu, lam, ys, matp, dt, tau);
  if (write_params) {
    M.mu[p] = mu;
    M.lam[p] = lam;
    M.ys[p] = ys;
  }
  store9(Fe_out, p, st.Fe);
  const float k = -dt * M.vol[p] * M.inv_dx;
  float S[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = k * tau[i];
  store9(S_out, p, S);
}

// The stress stage of the first substep of a forward call (later substeps
// get it from the previous substep's G2P).
template <int MAT>
__global__ __launch_bounds__(kBlock) void k_stress(Batch Bt, float dt,
                                                   const float* __restrict__ F, float* __restrict__ Fe_out,
                                                   float* __restrict__ S_out, int write_params,
                                                   const float* __restrict__ mst_in, float* __restrict__ mst_rec) {
  per_particle(Bt, MAT, [&](const DevModel& M, const int p) {
  if (M.sel && M.sel[p] != 0) return;
  float Ftr[9];
  load9(F, p, Ftr);
  particle_stress<MAT>(M, dt, p, Ftr, Fe_out, S_out, write_params, Bt.N, mst_in, mst_rec);
  });
}

// p2g_apic_with_stress (mpm_utils.py:358-423): APIC momentum, mass and the
// stress impulse into the block tiles.
template <int WG, bool DET = false>
__global__ __launch_bounds__(WG) void k_p2g(Batch Bt, const float* __restrict__ x, const float* __restrict__ v,
                                                const float* __restrict__ C, const float* __restrict__ S, Binning B,
                                                float4* __restrict__ tiles, int* __restrict__ work) {
  block_scatter<4, DET ? WG / kWave : 0>(Bt, B, tiles, work, [&](int p, float* xp, ScatterPayload& P, const DevModel& M) {
    const float dx = M.dx;
#pragma unroll
    for (int d = 0; d < 3; ++d) xp[d] = x[3 * p + d];
    const float mp = M.mass[p];
    float Cp[9], Ct[9];
    load9(C, p, Cp);
    rpic_blend(Cp, M.rpic, Ct);
#pragma unroll
    for (int d = 0; d < 3; ++d) P.A[d] = mp * v[3 * p + d];
#pragma unroll
    for (int k = 0; k < 9; ++k) P.B[k] = mp * dx * Ct[k];
    load9(S, p, P.S);
    P.m = mp;
  });
}

// set_F_C_p2g (mpm_utils.py:843-876) of restart_and_compute_F_C: the mass
// and the mass-weighted displacement target - x into the block tiles (no
// APIC or stress term).
template <int WG, bool DET = false>
__global__ __launch_bounds__(WG) void k_p2g_disp(Batch Bt, const float* __restrict__ x,
                                                 const float* __restrict__ target, Binning B,
                                                 float4* __restrict__ tiles, int* __restrict__ work) {
  block_scatter<4, DET ? WG / kWave : 0>(Bt, B, tiles, work, [&](int p, float* xp, ScatterPayload& P, const DevModel& M) {
    const float mp = M.mass[p];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xp[d] = x[3 * p + d];
      P.A[d] = mp * (target[3 * p + d] - xp[d]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) P.B[k] = P.S[k] = 0.f;
    P.m = mp;
  });
}

// set_F_C_g2p (mpm_utils.py:879-919): F_trial = I + sum_i v_i (x) grad w_i
// over the normalised displacement grid (its C is not stored, as the
// reference; cov is not updated, update_cov_with_F being off).  Particles
// with selection != 0 keep theirs.
__global__ __launch_bounds__(kBlock) void k_restart_F(Batch Bt, const float* __restrict__ x,
                                                      const float4* __restrict__ v_out, float* __restrict__ F) {
  per_particle(Bt, -1, [&](const DevModel& M, const int p) {
    if (M.sel && M.sel[p] != 0) return;
    const float xp[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
    Stencil s;
    make_stencil(xp, M.inv_dx, s);
    float nG[9];
    m_zero(nG);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int ix = s.base[0] + a, iy = s.base[1] + b, iz = s.base[2] + c;
          const bool in = node_in(ix, iy, iz, M.n);
          const float4 g = v_out[M.g0 + (in ? node_id(ix, iy, iz, M.n) : 0)];
          const float gv[3] = {in ? g.x : 0.f, in ? g.y : 0.f, in ? g.z : 0.f};
          const float gw[3] = {s.dw[0][a] * s.w[1][b] * s.w[2][c], s.w[0][a] * s.dw[1][b] * s.w[2][c],
                               s.w[0][a] * s.w[1][b] * s.dw[2][c]};
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) nG[i * 3 + j] += gv[i] * gw[j];
        }
    float Fn[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fn[k] = ((k % 4 == 0) ? 1.f : 0.f) + nG[k] * M.inv_dx;
    store9(F, p, Fn);
  });
}

// Per node: gather (mv, m) from the tiles, then
// grid_normalization_and_gravity, damping and the colliders; (mv, m) is kept
// when `keep` is set (the adjoint needs it).
// The node passes visit only the nodes a scatter can have touched: the 2x2x2
// nodes owned by each block (nodes 2o .. 2o + 1 per axis) whose `near` flag
// is set.  Other nodes keep stale values, and no later kernel reads them:
// G2P and the P2G adjoint read only the stencil nodes of particles, which
// lie in their block's tile.  A capped grid strides over (near block, owned
// node) pairs of the binning's near list, so the cost follows the occupied
// part of the grid, not its size.  The default cap is 4096 one-wave
// workgroups (a 200k-particle 64^3 rollout's near nodes in one pass; a
// 32-simulation batch's ~640k in a few).
constexpr int kNodePassGridCap = 4096;
// ONE: a batch of one simulation, whose table entry is then read with
// scalar loads at once instead of per lane after the list entry
template <bool ONE, typename F>
EMPM_DEV void for_near_nodes_t(const Batch& Bt, const Binning& B, const ColTable* __restrict__ T, ColTable& Ts,
                               F f) {
  const size_t stride = (size_t)gridDim.x * blockDim.x, last = (size_t)B.nblk - 1;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  // list entries are read ahead of the count test (the list has nblk slots),
  // together with the collider table
  const TableCopy tc = table_load(T);
  int blk_next = B.nlist[(q >> 3) < last ? (q >> 3) : last];
  const size_t npairs = (size_t)(*B.nnear) * 8;
  table_store(tc, Ts);
  for (; q < npairs; q += stride) {
    const int blk = blk_next, r = (int)(q & 7);
    const size_t qn = (q + stride) >> 3;
    blk_next = B.nlist[qn < last ? qn : last];
    const DevModel& M = ONE ? Bt.cand[0] : cand_of_block(Bt, blk);
    const int nb = M.nb, lb = blk - M.b0;
    const int ix = 2 * (lb / (nb * nb)) + (r >> 2), iy = 2 * ((lb / nb) % nb) + ((r >> 1) & 1),
              iz = 2 * (lb % nb) + (r & 1);
    if (ix >= M.n || iy >= M.n || iz >= M.n) continue;
    f(M, ix, iy, iz, (size_t)M.g0 + node_id(ix, iy, iz, M.n));
  }
}
template <typename F>
EMPM_DEV void for_near_nodes(const Batch& Bt, const Binning& B, const ColTable* __restrict__ T, ColTable& Ts, F f) {
  if (Bt.ncand == 1) for_near_nodes_t<true>(Bt, B, T, Ts, f);
  else for_near_nodes_t<false>(Bt, B, T, Ts, f);
}

__global__ __launch_bounds__(kBlock) void k_grid(Batch Bt, const ColTable* __restrict__ T, float dt, Binning B,
                                                 const float4* __restrict__ tiles,
                                                 float4* __restrict__ keep, float4* __restrict__ v_out) {
  __shared__ ColTable Ts;
  for_near_nodes(Bt, B, T, Ts, [&](const DevModel& M, int ix, int iy, int iz, size_t id) {
  const float4 g = gather_node(B, M, tiles, ix, iy, iz);
  if (keep) keep[id] = g;
  float v[3], fac[3];
  grid_node(M, Ts, dt, ix, iy, iz, g, v, fac);
  v_out[id] = make_float4(v[0], v[1], v[2], g.w);
  });
}

// g2p_differentiable (mpm_utils.py:502-579).  With `next_stress` set the
// stress stage of the next substep runs on the new F_trial here (writing
// Fe_next and S), so the scatter kernel needs no SVD.
// Fe and Fe_next may be the same buffer (each particle reads its Fe before
// writing Fe_next).
template <int MAT>
__global__ __launch_bounds__(kBlock) void k_g2p(Batch Bt, float dt, const float* __restrict__ x,
                                                const float* Fe, const float4* __restrict__ v_out,
                                                float* __restrict__ x1, float* __restrict__ v1,
                                                float* __restrict__ C1, float* __restrict__ F1, int next_stress,
                                                float* Fe_next, float* __restrict__ S_next, int write_params,
                                                int write_F, int* __restrict__ count, int2* __restrict__ key_rank,
                                                float* __restrict__ mst_rec_next, int* __restrict__ status) {
  // with count set, the counting step of the next substep's binning is done
  // here on the new positions (k_bin_count's work), last: counting in the
  // middle keeps the gather's state live across the stress stage
  per_particle(Bt, MAT, [&](const DevModel& M, const int p) {
  if (M.sel && M.sel[p] != 0) {  // partial_clone defaults (mpm_data_structure.py:356-387)
    float I[9], Z[9];
    m_eye(I);
    m_zero(Z);
    x1[3 * p] = x1[3 * p + 1] = x1[3 * p + 2] = 0.f;
    v1[3 * p] = v1[3 * p + 1] = v1[3 * p + 2] = 0.f;
    store9(C1, p, Z);
    if (write_F) store9(F1, p, I);
    if (count) {
      const float z[3] = {0.f, 0.f, 0.f};
      count_bin(true, bin_key(z, M), p, count, key_rank);
    }
    return;
  }
  const float xp[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
  Stencil s;
  make_stencil(xp, M.inv_dx, s);
  // v' = sum w v_i, C' = 4 inv_dx sum w v_i (x) d_i, grad v = sum v_i (x) grad w_i,
  // factored by stencil column (a, b): per column the sums over c of v Z,
  // v Z d_z and v dZ, then the column's x and y weights once
  float nv[3] = {0.f, 0.f, 0.f}, nC[9], nG[9];
  m_zero(nC);
  m_zero(nG);
  const float c4 = M.inv_dx * 4.f;
  float zd[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) zd[c] = s.w[2][c] * ((float)c - s.f[2]);
  // (the unmasked node loads of the adjoint kernels, for waves whose
  // stencils are all inside the grid, let the compiler hoist all 27 loads
  // here: 84 -> 129 VGPRs, 5 -> 3 waves, slower)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      float P[3] = {0.f, 0.f, 0.f}, Q[3] = {0.f, 0.f, 0.f}, R[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int ix = s.base[0] + a, iy = s.base[1] + b, iz = s.base[2] + c;
        const bool in = node_in(ix, iy, iz, M.n);
        const float4 gv = v_out[M.g0 + (in ? node_id(ix, iy, iz, M.n) : 0)];
        const float gvv[3] = {in ? gv.x : 0.f, in ? gv.y : 0.f, in ? gv.z : 0.f};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          P[i] += gvv[i] * s.w[2][c];
          Q[i] += gvv[i] * zd[c];
          R[i] += gvv[i] * s.dw[2][c];
        }
      }
      const float X = s.w[0][a], Y = s.w[1][b];
      const float XY = X * Y, XYd0 = XY * ((float)a - s.f[0]), XYd1 = XY * ((float)b - s.f[1]);
      const float dXY = s.dw[0][a] * Y, XdY = X * s.dw[1][b];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        nv[i] += XY * P[i];
        nC[i * 3 + 0] += XYd0 * P[i];
        nC[i * 3 + 1] += XYd1 * P[i];
        nC[i * 3 + 2] += XY * Q[i];
        nG[i * 3 + 0] += dXY * P[i];
        nG[i * 3 + 1] += XdY * P[i];
        nG[i * 3 + 2] += XY * R[i];
      }
    }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    nC[k] *= c4;
    nG[k] *= M.inv_dx;
  }
  const float amin = M.dx * 2.f;
  const float amax = M.grid_lim - M.dx * 2.f;
  float xo[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v1[3 * p + i] = nv[i];
    const float xn = xp[i] + dt * nv[i];
    xo[i] = fminf(fmaxf(xn, amin), amax);
    x1[3 * p + i] = xo[i];
  }
  store9(C1, p, nC);
  float IF[9], Fp[9], Fn[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) IF[i] = ((i % 4 == 0) ? 1.f : 0.f) + nG[i] * dt;
  load9(Fe, p, Fp);
  m_mul(IF, Fp, Fn);
  if (write_F) store9(F1, p, Fn);
  // failure flags of the simulation (read once per rollout by the host):
  // a non-finite new state, or a particle faster than a cell per substep
  {
    float acc = xo[0] + xo[1] + xo[2] + nv[0] + nv[1] + nv[2];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc += Fn[i];
    const bool nonfinite = !isfinite(acc);
    const float cfl = dt * fmaxf(fmaxf(fabsf(nv[0]), fabsf(nv[1])), fabsf(nv[2]));
    const bool fast = cfl > M.dx;
    if (nonfinite || fast)
      atomicOr(&status[&M - Bt.cand], (nonfinite ? EMPM_STATUS_NONFINITE : 0) | (fast ? EMPM_STATUS_CFL : 0));
  }
  if (next_stress) particle_stress<MAT>(M, dt, p, Fn, Fe_next, S_next, write_params, Bt.N, nullptr, mst_rec_next);
  if (count) count_bin(true, bin_key(xo, M), p, count, key_rank);
  });
}

// ---------------------------------------------------------------------------
// Adjoint kernels
// ---------------------------------------------------------------------------
// Adjoint of g2p_differentiable, per particle: dL/dx (partial), dL/dFe, and
// the payload of the dL/dv_out scatter (A = dL/dv', B = 4 inv_dx dL/dC',
// S = dt inv_dx dL/dF' Fe^T): k_g2p_adj stores A, the only part that needs
// the gather (the clamp mask), and the scatter forms B and S from the
// state's cotangents itself.
// Particle p's part: the cotangents of its next state (dx1, dv1, dC1, dF1)
// come in registers, so the P2G adjoint of the later substep can run this
// for the same particle right after its own work (k_p2g_adj<.., FUSE>).
EMPM_DEV void g2p_adj_particle(const DevModel& M, float dt, int p, const float* __restrict__ x,
                               const float* __restrict__ Fe, const float4* __restrict__ v_out, const float* dx1,
                               const float* dv1, const float* dC, const float* dF, float* __restrict__ xbar,
                               float* __restrict__ Febar, float* __restrict__ payload) {
  if (M.sel && M.sel[p] != 0) {
    xbar[3 * p] = xbar[3 * p + 1] = xbar[3 * p + 2] = 0.f;
    float Z[9];
    m_zero(Z);
    store9(Febar, p, Z);
    return;
  }
  const float inv_dx = M.inv_dx;
  const float c4 = inv_dx * 4.f;
  const float xp[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
  float Fp[9], Gb[9];
  load9(Fe, p, Fp);
  // F1 = (I + dt G) Fe:  Gbar = dt dF Fe^T ; Febar = (I + dt G)^T dF
  m_mul_bt(dF, Fp, Gb);
#pragma unroll
  for (int k = 0; k < 9; ++k) Gb[k] *= dt;
  Stencil s;
  make_stencil(xp, inv_dx, s);
  // One pass over the stencil: v' and grad v (for the clamp mask and the F
  // adjoint) and every d(w, grad w, d)/dx term.  The term of dL/dv' through
  // the weights, sum_j (dL/dv' . v_j) dw_j/df, is (dL/dv')^T grad v / inv_dx,
  // so it is added after the pass, once dL/dv' (which needs v') is known.
  // Per node j with v_j, w_j = X Y Z and d_j = node - f the pass needs
  //   dL/dw_j  = 4 inv_dx v_j . (dC d_j)            (through C')
  //   dL/dgw_j = Gb^T v_j                           (through grad v)
  // times the first and second derivatives of w_j, and 4 inv_dx w_j dC^T v_j
  // (through d_j).  All are linear in v_j with products of per-axis weights,
  // so each stencil column (nodes (a, b, 0..2)) is summed over c first (P, Q,
  // R, R2, A5 below) and folded in with its x and y weights; the d_j term is
  // 4 inv_dx dC^T v' after the pass.
  float nv[3] = {0.f, 0.f, 0.f}, nG[9], fb[3] = {0.f, 0.f, 0.f};
  m_zero(nG);
  float zd[3], dzd[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    zd[c] = s.w[2][c] * ((float)c - s.f[2]);
    dzd[c] = s.dw[2][c] * ((float)c - s.f[2]);
  }
  const float q[3] = {dC[2], dC[5], dC[8]};  // dC e_z
  // node loads straight from the stencil's first node when every stencil
  // of the wave is inside the grid, else bounds-tested (out-of-grid nodes
  // count as zero)
  const bool inside = wave_stencils_inside(s, M.n);
  const size_t nn = (size_t)M.n * M.n;
  const float4* __restrict__ vb = v_out + M.g0 + (inside ? node_id(s.base[0], s.base[1], s.base[2], M.n) : 0);
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    const float X = sw(s, 0, a), dX = sdw(s, 0, a), ddX = a == 1 ? -2.f : 1.f;
    const float da = (float)a - s.f[0];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      // column sums over c: Z v, Z d_z v, dZ v, dZ d_z v, ddZ v
      float P[3] = {0.f, 0.f, 0.f}, Q[3] = {0.f, 0.f, 0.f}, R[3] = {0.f, 0.f, 0.f}, R2[3] = {0.f, 0.f, 0.f},
            A5[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float vi[3];
        if (inside) {
          const float4 g4 = vb[a * nn + b * M.n + c];
          vi[0] = g4.x, vi[1] = g4.y, vi[2] = g4.z;
        } else {
          const int ix = s.base[0] + a, iy = s.base[1] + b, iz = s.base[2] + c;
          const bool in = node_in(ix, iy, iz, M.n);
          const float4 g4 = v_out[M.g0 + (in ? node_id(ix, iy, iz, M.n) : 0)];
          vi[0] = in ? g4.x : 0.f, vi[1] = in ? g4.y : 0.f, vi[2] = in ? g4.z : 0.f;
        }
        const float Z = s.w[2][c], dZ = s.dw[2][c];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          P[m] += Z * vi[m];
          Q[m] += zd[c] * vi[m];
          R[m] += dZ * vi[m];
          R2[m] += dzd[c] * vi[m];
          A5[m] += c == 1 ? -2.f * vi[m] : vi[m];
        }
      }
      const float Y = s.w[1][b], dY = s.dw[1][b], ddY = b == 1 ? -2.f : 1.f;
      const float db = (float)b - s.f[1];
      const float XY = X * Y, dXY = dX * Y, XdY = X * dY;
      float WZ = 0.f, WdZ = 0.f;
      float gP0 = 0.f, gP1 = 0.f, gR0 = 0.f, gR1 = 0.f, gR2 = 0.f, gA2 = 0.f;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        nv[m] += XY * P[m];
        nG[m * 3 + 0] += dXY * P[m];
        nG[m * 3 + 1] += XdY * P[m];
        nG[m * 3 + 2] += XY * R[m];
        const float pab = dC[m * 3 + 0] * da + dC[m * 3 + 1] * db;
        WZ += pab * P[m] + q[m] * Q[m];
        WdZ += pab * R[m] + q[m] * R2[m];
        gP0 += Gb[m * 3 + 0] * P[m];
        gP1 += Gb[m * 3 + 1] * P[m];
        gR0 += Gb[m * 3 + 0] * R[m];
        gR1 += Gb[m * 3 + 1] * R[m];
        gR2 += Gb[m * 3 + 2] * R[m];
        gA2 += Gb[m * 3 + 2] * A5[m];
      }
      fb[0] += c4 * dXY * WZ + inv_dx * (gP0 * ddX * Y + gP1 * dX * dY + gR2 * dXY);
      fb[1] += c4 * XdY * WZ + inv_dx * (gP0 * dX * dY + gP1 * X * ddY + gR2 * XdY);
      fb[2] += c4 * XY * WdZ + inv_dx * (gR0 * dXY + gR1 * XdY + gA2 * XY);
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) nG[k] *= inv_dx;
#pragma unroll
  for (int k = 0; k < 3; ++k) fb[k] -= c4 * (dC[0 * 3 + k] * nv[0] + dC[1 * 3 + k] * nv[1] + dC[2 * 3 + k] * nv[2]);
  const float amin = M.dx * 2.f, amax = M.grid_lim - M.dx * 2.f;
  float xb[3], nvb[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float xn = xp[k] + dt * nv[k];
    // wp.clamp passes the adjoint to x unless x < lo or x > hi
    const float msk = (xn < amin || xn > amax) ? 0.f : 1.f;
    const float g = dx1[k] * msk;
    xb[k] = g;
    nvb[k] = dv1[k] + dt * g;
  }
  float IF[9], Feb[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) IF[k] = ((k % 4 == 0) ? 1.f : 0.f) + nG[k] * dt;
  m_mul_at(IF, dF, Feb);
  store9(Febar, p, Feb);
#pragma unroll
  for (int l = 0; l < 3; ++l)
    xbar[3 * p + l] = xb[l] + inv_dx * fb[l] + (nvb[0] * nG[l] + nvb[1] * nG[3 + l] + nvb[2] * nG[6 + l]);
  reinterpret_cast<float4*>(payload)[p] = make_float4(nvb[0], nvb[1], nvb[2], 0.f);
}

__global__ __launch_bounds__(kBlock) void k_g2p_adj(Batch Bt, float dt, const float* __restrict__ x,
                                                    const float* __restrict__ Fe, const float4* __restrict__ v_out,
                                                    const float* __restrict__ dx1, const float* __restrict__ dv1,
                                                    const float* __restrict__ dC1, const float* __restrict__ dF1,
                                                    float* __restrict__ xbar, float* __restrict__ Febar,
                                                    float* __restrict__ payload, int* __restrict__ scatter_work) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *scatter_work = 0;  // for the k_g2p_adj_scatter launch that follows
  // the cotangents are read where the body uses them (the position and
  // velocity ones only after the node loop)
  per_particle(Bt, -1, [&](const DevModel& M, const int p) {
    g2p_adj_particle(M, dt, p, x, Fe, v_out, dx1 + 3 * (size_t)p, dv1 + 3 * (size_t)p, dC1 + 9 * (size_t)p,
                     dF1 + 9 * (size_t)p, xbar, Febar, payload);
  });
}

// Scatter of dL/dv_out (the transpose of the G2P gather) into the block tiles.
// Per particle: A = dL/dv' (clamp mask applied, from k_g2p_adj), and from
// the state's cotangents B = 4 inv_dx dL/dC' and S = dt inv_dx dL/dF' Fe^T
// (F' = (I + dt grad v) Fe).
template <int WG, bool DET = false>
__global__ __launch_bounds__(WG) void k_g2p_adj_scatter(Batch Bt, float dt, const float* __restrict__ x,
                                                            const float* __restrict__ payload,
                                                            const float* __restrict__ dC1,
                                                            const float* __restrict__ dF1,
                                                            const float* __restrict__ Fe, Binning B,
                                                            float4* __restrict__ tiles,
                                                            int* __restrict__ work) {
  block_scatter<3, DET ? WG / kWave : 0>(Bt, B, tiles, work, [&](int p, float* xp, ScatterPayload& P, const DevModel& M) {
#pragma unroll
    for (int d = 0; d < 3; ++d) xp[d] = x[3 * p + d];
    const float4 a = reinterpret_cast<const float4*>(payload)[p];
    P.A[0] = a.x, P.A[1] = a.y, P.A[2] = a.z;
    float dC[9], dF[9], Fp[9], Gb[9];
    load9(dC1, p, dC);
    load9(dF1, p, dF);
    load9(Fe, p, Fp);
    m_mul_bt(dF, Fp, Gb);
    const float c4 = M.inv_dx * 4.f, sg = dt * M.inv_dx;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      P.B[k] = c4 * dC[k];
      P.S[k] = sg * Gb[k];
    }
    P.m = 0.f;
  });
}

// Adjoint of the grid stage: gather dL/dv_out from the tiles, replay the
// node's forward (normalisation, damping, colliders) to get dv_out/dv per
// component, and write dL/d(mv, m).  The grid-mass term is dropped when
// `drop_mass` is set (reference Tape behaviour).
__global__ __launch_bounds__(kBlock) void k_grid_adj(Batch Bt, const ColTable* __restrict__ T, float dt, Binning B,
                                                     const float4* __restrict__ tiles,
                                                     const float4* __restrict__ keep, float4* __restrict__ gadj_in,
                                                     int drop_mass) {
  __shared__ ColTable Ts;
  for_near_nodes(Bt, B, T, Ts, [&](const DevModel& M, int ix, int iy, int iz, size_t id) {
  const float4 b = gather_node(B, M, tiles, ix, iy, iz);
  const float4 g = keep[id];
  float v[3], fac[3];
  grid_node(M, Ts, dt, ix, iy, iz, g, v, fac);
  float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
  if (g.w > 1e-15f) {
    const float im = 1.f / g.w;
    const float b0 = b.x * fac[0], b1 = b.y * fac[1], b2 = b.z * fac[2];
    o.x = b0 * im;
    o.y = b1 * im;
    o.z = b2 * im;
    o.w = drop_mass ? 0.f : -(b0 * g.x + b1 * g.y + b2 * g.z) * im * im;
  }
  gadj_in[id] = o;
  });
}

// MUT: the simulations of this material may rewrite (yield stress, mu, lam)
// in their return map (material 5, hardening): those are then state, read
// from the substep's record `mst` (or the model arrays when it is null), and
// d_mu, d_lam, d_ys carry their cotangents backwards through the update.
// The G2P adjoint of the substep before, for the same particle, when the
// reverse sweep fuses the two (k_p2g_adj<.., FUSE = true>): that substep's
// recorded input positions, Fe and grid velocity, and where its partial
// adjoints go (the xbar / Febar this launch reads are overwritten per
// particle after they are read).
struct PrevG2PAdj {
  const float* x;
  const float* Fe;
  const float4* v_out;
  float* xbar;
  float* Febar;
  float* payload;
  int* scatter_work;
};

// With FUSE, the cotangents of the position and velocity of this substep's
// input state go straight into the previous substep's G2P adjoint in
// registers and are not stored (the sweep reads them nowhere else); those of
// C and F_trial are stored for that substep's scatter.
template <int MAT, bool MUT, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_p2g_adj(Batch Bt, float dt, const float* __restrict__ x,
                                                    const float* __restrict__ v, const float* __restrict__ C,
                                                    const float* __restrict__ F, const float4* __restrict__ gadj_in,
                                                    const float* xbar, const float* Febar_in,
                                                    float* __restrict__ dx, float* __restrict__ dv,
                                                    float* __restrict__ dCo, float* __restrict__ dFo,
                                                    float* __restrict__ dmu, float* __restrict__ dlam,
                                                    float* __restrict__ dys, float* __restrict__ dmass,
                                                    float* __restrict__ dvol, const float* __restrict__ mst,
                                                    PrevG2PAdj prev) {
  if (FUSE && blockIdx.x == 0 && threadIdx.x == 0) *prev.scatter_work = 0;  // for the scatter that follows
  per_particle(Bt, MAT, [&](const DevModel& M, const int p) {
  if (M.sel && M.sel[p] != 0) {
    float Z[9];
    m_zero(Z);
    if (!FUSE) {
      dx[3 * p] = dx[3 * p + 1] = dx[3 * p + 2] = 0.f;
      dv[3 * p] = dv[3 * p + 1] = dv[3 * p + 2] = 0.f;
    }
    store9(dCo, p, Z);
    store9(dFo, p, Z);
    if (FUSE) g2p_adj_particle(M, dt, p, prev.x, prev.Fe, prev.v_out, Z, Z, Z, Z, prev.xbar, prev.Febar, prev.payload);
    return;
  }
  const float xp[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
  const float vp[3] = {v[3 * p], v[3 * p + 1], v[3 * p + 2]};
  float Cp[9], Ct[9];
  load9(C, p, Cp);
  rpic_blend(Cp, M.rpic, Ct);
  Stencil s;
  make_stencil(xp, M.inv_dx, s);
  const float mp = M.mass[p], Vp = M.vol[p];
  // The node loop accumulates, in grid units (d = node - f, gf = dw/df):
  //   S0 = sum w a,  S1 = sum w a (x) d,  T = sum a (x) gf,  W = sum w b,
  //   fA = sum (a.u + b) gf  (u = v + h Ct d),  Hs_jk = sum H_jk a
  // with (a, b) = dL/d(mv, m) of the node and H its weight Hessian; every
  // per-particle factor (m, V, dt, h = dx, tau) is applied once afterwards.
  // All are linear in (a, b) with products of per-axis weights, so each
  // stencil column (nodes (a, b, 0..2)) is first summed over c with the z
  // weights, then folded in with the column's x and y weights.
  float S0[3] = {0.f, 0.f, 0.f}, S1[9], T[9], fA[3] = {0.f, 0.f, 0.f}, Hs[6][3];
  m_zero(S1);
  m_zero(T);
#pragma unroll
  for (int j = 0; j < 6; ++j) Hs[j][0] = Hs[j][1] = Hs[j][2] = 0.f;
  float W = 0.f;
  const float h = M.dx;
  float zd[3], dzd[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    zd[c] = s.w[2][c] * ((float)c - s.f[2]);
    dzd[c] = s.dw[2][c] * ((float)c - s.f[2]);
  }
  const float Ct2[3] = {Ct[2], Ct[5], Ct[8]};
  // the nodes of plane a (9 of them, b-major) are loaded together; when a
  // stencil of the wave leaves the grid, its out-of-grid nodes read
  // node 0 of the grid and count as zero
  const bool inside = wave_stencils_inside(s, M.n);
  const size_t nn = (size_t)M.n * M.n;
  const float4* __restrict__ gb =
      gadj_in + M.g0 + (inside ? node_id(s.base[0], s.base[1], s.base[2], M.n) : 0);
  const auto load_plane = [&](int a, float4* pl) {
    if (inside) {
      const float4* __restrict__ pa = gb + a * nn;
#pragma unroll
      for (int j = 0; j < 9; ++j) pl[j] = pa[(j / 3) * M.n + j % 3];
      return;
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const int ix = s.base[0] + a, iy = s.base[1] + j / 3, iz = s.base[2] + j % 3;
      const bool in = node_in(ix, iy, iz, M.n);
      const float4 ga = gadj_in[M.g0 + (in ? node_id(ix, iy, iz, M.n) : 0)];
      pl[j] = in ? ga : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  // one plane in flight at a time (prefetching the next plane costs 42
  // VGPRs and measured slower)
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    float4 cur[9];
    load_plane(a, cur);
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      // column sums over c: Z a, Z d_z a, dZ a, dZ d_z a, ddZ a, Z b, dZ b
      float A1[3] = {0.f, 0.f, 0.f}, A2[3] = {0.f, 0.f, 0.f}, A3[3] = {0.f, 0.f, 0.f}, A4[3] = {0.f, 0.f, 0.f},
            A5[3] = {0.f, 0.f, 0.f};
      float B1 = 0.f, B3 = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float4 ga = cur[b * 3 + c];
        const float av[3] = {ga.x, ga.y, ga.z};
        const float bm = ga.w;
        const float Z = s.w[2][c], dZ = s.dw[2][c];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          A1[i] += Z * av[i];
          A2[i] += zd[c] * av[i];
          A3[i] += dZ * av[i];
          A4[i] += dzd[c] * av[i];
          A5[i] += c == 1 ? -2.f * av[i] : av[i];
        }
        B1 += Z * bm;
        B3 += dZ * bm;
      }
      const float X = sw(s, 0, a), Y = s.w[1][b], dX = sdw(s, 0, a), dY = s.dw[1][b];
      const float ddX = a == 1 ? -2.f : 1.f, ddY = b == 1 ? -2.f : 1.f;
      const float da = (float)a - s.f[0], db = (float)b - s.f[1];
      const float XY = X * Y, XYd0 = XY * da, XYd1 = XY * db, dXY = dX * Y, XdY = X * dY;
      const float ddXY = ddX * Y, XddY = X * ddY, dXdY = dX * dY;
      float e1 = B1, e3 = B3;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float uab = vp[i] + h * (Ct[i * 3 + 0] * da + Ct[i * 3 + 1] * db);
        e1 += uab * A1[i] + h * Ct2[i] * A2[i];
        e3 += uab * A3[i] + h * Ct2[i] * A4[i];
        S0[i] += XY * A1[i];
        S1[i * 3 + 0] += XYd0 * A1[i];
        S1[i * 3 + 1] += XYd1 * A1[i];
        S1[i * 3 + 2] += XY * A2[i];
        T[i * 3 + 0] += dXY * A1[i];
        T[i * 3 + 1] += XdY * A1[i];
        T[i * 3 + 2] += XY * A3[i];
        Hs[0][i] += ddXY * A1[i];  // xx
        Hs[1][i] += dXdY * A1[i];  // xy
        Hs[2][i] += dXY * A3[i];   // xz
        Hs[3][i] += XddY * A1[i];  // yy
        Hs[4][i] += XdY * A3[i];   // yz
        Hs[5][i] += XY * A5[i];    // zz
      }
      W += XY * B1;
      fA[0] += dXY * e1;
      fA[1] += XdY * e1;
      fA[2] += XY * e3;
    }
  }
  // the substep's return map and stress, recomputed (after the node loop,
  // which does not need them, so that their state is not live during it)
  float Ftr[9];
  load9(F, p, Ftr);
  float mu, lam, ys;
  if (MUT && mst) {
    ys = mst[p];
    mu = mst[(size_t)Bt.N + p];
    lam = mst[2 * (size_t)Bt.N + p];
  } else {
    mu = M.mu[p];
    lam = M.lam[p];
    ys = M.ys[p];
  }
  const float mu0 = mu, lam0 = lam, ys0 = ys;
  MatParams matp = M.mp;
  matp.material = MAT;
  ConstitState st;
  float tau[9];
  constitutive_fwd(st, Ftr, mu, lam, ys, matp, dt, tau);
  // fB_j = sum_k H_jk (tau^T a)_k
  const int hm[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  float fB[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i) acc += tau[i * 3 + k] * Hs[hm[j][k]][i];
    fB[j] = acc;
  }
  // dL/d(v, C~, tau, m, V) of the particle and dL/df (f = x inv_dx - base)
  const float sV = -dt * Vp * M.inv_dx;  // d(stress impulse)/d(tau . grad w) in grid units
  float vb[3], Ctb[9], taub[9], fb[3];
  float mb = W, Vb = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    vb[i] = mp * S0[i];
    mb += S0[i] * vp[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Ctb[i * 3 + j] = mp * h * S1[i * 3 + j];
      taub[i * 3 + j] = sV * T[i * 3 + j];
      mb += h * Ct[i * 3 + j] * S1[i * 3 + j];
      Vb += -dt * M.inv_dx * tau[i * 3 + j] * T[i * 3 + j];
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
    fb[j] = mp * fA[j] - mp * h * (Ct[0 * 3 + j] * S0[0] + Ct[1 * 3 + j] * S0[1] + Ct[2 * 3 + j] * S0[2]) +
            sV * fB[j];
  float xo[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) xo[i] = xbar[3 * p + i] + M.inv_dx * fb[i];
  float Cb[9];
  const float r = M.rpic;
  if (r < -0.001f) {
    m_zero(Cb);
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Cb[i * 3 + j] = (1.f - r) * Ctb[i * 3 + j] + r / 2.f * (Ctb[i * 3 + j] - Ctb[j * 3 + i]);
  }
  float taur[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) taur[i * 3 + j] = 0.5f * (taub[i * 3 + j] + taub[j * 3 + i]);
  float Feb[9], Ftrb[9];
  load9(Febar_in, p, Feb);
  m_zero(Ftrb);
  float mub = 0.f, lamb = 0.f, ysb = 0.f;
  if constexpr (MUT) {
    if (mutates_params<MAT>(M)) {
      // the stress reads the updated (mu, lam); the cotangents of the updated
      // (yield stress, mu, lam) are d_ys, d_mu, d_lam on entry, and on return
      // those of the values this substep started from
      float mub1 = dmu ? dmu[p] : 0.f, lamb1 = dlam ? dlam[p] : 0.f;
      const float ysb1 = dys ? dys[p] : 0.f;
      stress_adj(st, MAT, mu, lam, taur, Feb, mub1, lamb1);
      return_map_adj(st, Ftr, mu0, lam0, ys0, matp, dt, Feb, Ftrb, mub, lamb, ysb);
      vm_state_adj(st, mu0, ys0, mu, matp, MAT == 5, ysb1, mub1, lamb1, Ftrb, ysb, mub, lamb);
      if (!FUSE) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          dx[3 * p + i] = xo[i];
          dv[3 * p + i] = vb[i];
        }
      }
      store9(dCo, p, Cb);
      store9(dFo, p, Ftrb);
      if (dmu) dmu[p] = mub;
      if (dlam) dlam[p] = lamb;
      if (dys) dys[p] = ysb;
      if (dmass) dmass[p] += mb;
      if (dvol) dvol[p] += Vb;
      if (FUSE)
        g2p_adj_particle(M, dt, p, prev.x, prev.Fe, prev.v_out, xo, vb, Cb, Ftrb, prev.xbar, prev.Febar,
                         prev.payload);
      return;
    }
  }
  stress_adj(st, MAT, mu0, lam0, taur, Feb, mub, lamb);
  return_map_adj(st, Ftr, mu0, lam0, ys0, matp, dt, Feb, Ftrb, mub, lamb, ysb);
  if (!FUSE) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      dx[3 * p + i] = xo[i];
      dv[3 * p + i] = vb[i];
    }
  }
  store9(dCo, p, Cb);
  store9(dFo, p, Ftrb);
  if (dmu) dmu[p] += mub;
  if (dlam) dlam[p] += lamb;
  if (dys) dys[p] += ysb;
  if (dmass) dmass[p] += mb;
  if (dvol) dvol[p] += Vb;
  if (FUSE)
    g2p_adj_particle(M, dt, p, prev.x, prev.Fe, prev.v_out, xo, vb, Cb, Ftrb, prev.xbar, prev.Febar, prev.payload);
  });
}

// ---------------------------------------------------------------------------
// Binning, every substep: counting sort of the particles by the cell their
// stencil starts at.  Bins are the 8 cells of a 2x2x2-cell block; a block's
// particles are contiguous, cell-minor.
//   k_bin_count   per particle: bin key and rank within the bin (same-bin
//                 lanes of a wave share one atomic); counts are zero on entry
//                 (within a rollout the previous substep's G2P does this)
//   k_bin_alloc   one thread per block (four for a large batch), kBinScanWG
//                 threads per workgroup:
//                 block totals, their workgroup-local scan and one atomic
//                 per workgroup for its particle range; per-bin offsets,
//                 occupancy, active list; re-zeroes the counts
//   k_bin_place   per particle: perm[offs[key] + rank] = p; per block: the
//                 near flag (a gather of the 8 occupancy flags whose tiles
//                 reach the block) and the near list
// empm_sort_by_cell needs the blocks' ranges in block order and runs
// k_bin_scan + k_bin_finish (a grid-wide prefix over the workgroups)
// instead of k_bin_alloc.
// ---------------------------------------------------------------------------
constexpr int kBinScanWG = 1024;

// What the scatter and grid kernels read of one binning (device arrays).
struct BinBuf {
  int* perm;  // N
  int* offs;  // nb^3 * 9
  int* act;   // 2 + nb^3: heavy and light counts, then the non-empty block ids
  int* occ;   // nb^3
  int* near;  // nb^3
  int* nlist; // 1 + nb^3: count, then the ids of the near blocks
};

__global__ void k_fill_i32(int* __restrict__ a, size_t n, int v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_bin_count(Batch Bt, const float* __restrict__ x, int* __restrict__ count,
                                                      int2* __restrict__ key_rank) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < Bt.N;
  float xp[3] = {0.f, 0.f, 0.f};
  if (active) {
#pragma unroll
    for (int d = 0; d < 3; ++d) xp[d] = x[3 * p + d];
  }
  const int key = active ? bin_key(xp, cand_of_particle(Bt, p)) : -1;
  count_bin(active, key, p, count, key_rank);
}

// Exclusive scan of one value per thread over the workgroup (kBinScanWG
// threads); returns the exclusive prefix and sets `total`.
EMPM_DEV int wg_exclusive_scan(int v, int* smem, int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int kWaves = kBinScanWG / 64;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) smem[wid] = incl;
  __syncthreads();
  if (threadIdx.x < 64) {
    int w = threadIdx.x < kWaves ? smem[threadIdx.x] : 0;
    int wi = w;
#pragma unroll
    for (int o = 1; o < kWaves; o <<= 1) {
      const int t = __shfl_up(wi, o);
      if (lane >= o) wi += t;
    }
    if (threadIdx.x < kWaves) smem[threadIdx.x] = wi - w;
    if (threadIdx.x == kWaves - 1) smem[kWaves] = wi;
  }
  __syncthreads();
  total = smem[kWaves];
  const int r = smem[wid] + incl - v;
  __syncthreads();
  return r;
}

EMPM_DEV int block_count(const int* __restrict__ count, int b) {
  const int4* c4 = reinterpret_cast<const int4*>(count + (size_t)b * 8);
  const int4 a = c4[0], c = c4[1];
  return a.x + a.y + a.z + a.w + c.x + c.y + c.z + c.w;
}

__global__ __launch_bounds__(kBinScanWG) void k_bin_scan(int nblk, const int* __restrict__ count,
                                                         int* __restrict__ blkoff, int* __restrict__ wsum,
                                                         int* __restrict__ act, int* __restrict__ nlist,
                                                         int* __restrict__ p2g_work) {
  __shared__ int smem[kBinScanWG / 64 + 1];
  const int b = blockIdx.x * kBinScanWG + threadIdx.x;
  const int tot = b < nblk ? block_count(count, b) : 0;
  int wtot;
  const int ex = wg_exclusive_scan(tot, smem, wtot);
  if (b < nblk) blkoff[b] = tot > 0 ? ex : -1 - ex;  // empty blocks flagged by the sign
  if (threadIdx.x == 0) wsum[blockIdx.x] = wtot;
  if (b == 0) {
    act[0] = act[1] = 0;
    nlist[0] = 0;
    *p2g_work = 0;  // for the k_p2g launch that follows
  }
}

// Block b (non-empty, with per-cell counts c) joins the active list: heavy
// blocks (more than heavy_chunks chunks) from act[2] upwards, light ones
// from act[2 + nblk - 1] downwards; the lists' lengths are counted in
// *nheavy and *nlight, one atomic per wave and list.
EMPM_DEV void append_active(int b, int nblk, const int* c, int heavy_chunks, int* __restrict__ act, int* nheavy,
                            int* nlight) {
  const unsigned long long lt = (1ull << __lane_id()) - 1ull;
  const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
  int nch = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) nch += (c[k] + kWave - 1) / kWave;
  const bool heavy = nch > heavy_chunks;
  const unsigned long long hmask = __ballot(heavy), lmask = __ballot(!heavy);
  int hbase = 0, lbase = 0;
  if (__lane_id() == leader) {
    if (hmask) hbase = atomicAdd(nheavy, __popcll(hmask));
    if (lmask) lbase = atomicAdd(nlight, __popcll(lmask));
  }
  hbase = __shfl(hbase, leader);
  lbase = __shfl(lbase, leader);
  if (heavy)
    act[2 + hbase + __popcll(hmask & lt)] = b;
  else
    act[2 + nblk - 1 - (lbase + __popcll(lmask & lt))] = b;
}

// Per-bin offsets of block b (non-empty, counts c, first position o), its
// end, and its counts re-zeroed for the next binning.
EMPM_DEV void write_block_offsets(int b, int o, const int* c, int* __restrict__ offs, int* __restrict__ count) {
  int* ob = offs + (size_t)b * 9;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    ob[k] = o;
    o += c[k];
  }
  ob[8] = o;
  int4* c4 = reinterpret_cast<int4*>(count + (size_t)b * 8);
  c4[0] = make_int4(0, 0, 0, 0);
  c4[1] = make_int4(0, 0, 0, 0);
}

__global__ __launch_bounds__(kBinScanWG) void k_bin_finish(Batch Bt, int* __restrict__ count,
                                                           const int* __restrict__ blkoff,
                                                           const int* __restrict__ wsum, int* __restrict__ offs,
                                                           int* __restrict__ act, int* __restrict__ occ,
                                                           int heavy_chunks) {
  __shared__ int red[kBinScanWG / 64];
  const int nblk = Bt.nblk;
  const int b = blockIdx.x * kBinScanWG + threadIdx.x;
  // this workgroup's offset: the totals of the workgroups before it
  int part = 0;
  for (int w = threadIdx.x; w < (int)blockIdx.x; w += kBinScanWG) part += wsum[w];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  int prefix = 0;
#pragma unroll
  for (int k = 0; k < kBinScanWG / 64; ++k) prefix += red[k];
  if (b >= nblk) return;
  // an empty block's counts are already zero and its offsets are never read
  // (a block's end is the next block's first offset, which every non-empty
  // block writes itself): only its occupancy flag is written
  const int e = blkoff[b];
  const bool full = e >= 0;
  occ[b] = full ? 1 : 0;
  if (!full) return;
  const int4* c4 = reinterpret_cast<const int4*>(count + (size_t)b * 8);
  const int4 lo = c4[0], hi = c4[1];
  const int c[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  write_block_offsets(b, prefix + e, c, offs, count);
  append_active(b, nblk, c, heavy_chunks, act, &act[0], &act[1]);
}

// Offsets of the per-substep binning in one launch (k_bin_scan +
// k_bin_finish without the grid-wide prefix): each workgroup scans its
// blocks' particle totals and heavy and light block counts, and takes its
// particle range with one atomic on ctr[0] and its places in both halves of
// the active list with one 64-bit atomic on ctr[2..3] (heavy count in the
// low word, light in the high), so the blocks' ranges are in the order the
// workgroups got there and no atomic is issued per wave.  k_bin_place
// copies the list's counts into act[0], act[1] and re-zeroes ctr[0..3].
// Also clears the near-list count and the k_p2g work counter.  Each thread
// owns BPT blocks, kBinScanWG apart (four for a large batch's block arena,
// whose workgroups then fit the device at once).
template <int BPT>
__global__ __launch_bounds__(kBinScanWG) void k_bin_alloc(Batch Bt, int* __restrict__ count, int* __restrict__ offs,
                                                          int* __restrict__ act, int* __restrict__ occ,
                                                          int* __restrict__ nlist, int* __restrict__ p2g_work,
                                                          int* __restrict__ ctr,
                                                          int heavy_chunks) {
  __shared__ int smem[kBinScanWG / 64 + 1];
  __shared__ int base[3];  // the workgroup's particle range and list places (heavy, light)
  const int nblk = Bt.nblk;
  const int b0 = blockIdx.x * (kBinScanWG * BPT) + threadIdx.x;
  int c[BPT][8], tot[BPT];
  bool heavy[BPT];
  int sum = 0, nh = 0, nl = 0;
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    const int b = b0 + j * kBinScanWG;
#pragma unroll
    for (int k = 0; k < 8; ++k) c[j][k] = 0;
    if (b < nblk) {
      const int4* c4 = reinterpret_cast<const int4*>(count + (size_t)b * 8);
      const int4 lo = c4[0], hi = c4[1];
      c[j][0] = lo.x, c[j][1] = lo.y, c[j][2] = lo.z, c[j][3] = lo.w;
      c[j][4] = hi.x, c[j][5] = hi.y, c[j][6] = hi.z, c[j][7] = hi.w;
    }
    int t = 0, nch = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      t += c[j][k];
      nch += (c[j][k] + kWave - 1) / kWave;
    }
    tot[j] = t;
    heavy[j] = nch > heavy_chunks;
    sum += t;
    nh += (t > 0 && heavy[j]) ? 1 : 0;
    nl += (t > 0 && !heavy[j]) ? 1 : 0;
  }
  // the heavy and light counts share one scan as 16-bit fields (a
  // workgroup has at most kBinScanWG * BPT <= 4096 blocks of either kind)
  static_assert(kBinScanWG * BPT < 65536, "BPT");
  int wtot, hltot;
  int ex = wg_exclusive_scan(sum, smem, wtot);
  const int hlx = wg_exclusive_scan(nh | (nl << 16), smem, hltot);
  const int hx = hlx & 0xffff, lx = hlx >> 16, htot = hltot & 0xffff, ltot = hltot >> 16;
  if (threadIdx.x == 0) {
    base[0] = wtot > 0 ? atomicAdd(&ctr[0], wtot) : 0;
    unsigned long long hl = 0;
    if (htot + ltot > 0)
      hl = atomicAdd(reinterpret_cast<unsigned long long*>(ctr + 2),
                     ((unsigned long long)(unsigned)ltot << 32) | (unsigned)htot);
    base[1] = (int)(unsigned)(hl & 0xffffffffull);
    base[2] = (int)(unsigned)(hl >> 32);
  }
  if (b0 == 0) {
    nlist[0] = 0;
    *p2g_work = 0;  // for the k_p2g launch that follows
  }
  __syncthreads();
  const int wb = base[0];
  int hk = base[1] + hx, lk = base[2] + lx;
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    const int b = b0 + j * kBinScanWG;
    if (b >= nblk) break;
    occ[b] = tot[j] > 0 ? 1 : 0;
    if (tot[j] > 0) {
      write_block_offsets(b, wb + ex, c[j], offs, count);
      // heavy blocks from act[2] upwards, light ones from act[2 + nblk - 1] downwards
      if (heavy[j]) act[2 + hk++] = b;
      else act[2 + nblk - 1 - lk++] = b;
    }
    ex += tot[j];
  }
}

// WG = 64: one-wave workgroups, one near-list atomic per wave holding near
// blocks; WG = kBinScanWG (a large batch's block arena, where those
// atomics would queue on one address): one per workgroup, placed by a scan.
template <int WG>
__global__ __launch_bounds__(WG) void k_bin_place(Batch Bt, const int2* __restrict__ key_rank,
                                                  const int* __restrict__ offs, int* __restrict__ perm,
                                                  const int* __restrict__ occ, int* __restrict__ near,
                                                  int* __restrict__ nlist, int* __restrict__ act,
                                                  int* __restrict__ ctr) {
  static_assert(WG == kWave || WG == kBinScanWG, "WG");
  __shared__ int smem[kBinScanWG / 64 + 1];
  __shared__ int nbase;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (ctr && i == 0) {  // after k_bin_alloc: the active-list counts into the record, counters re-zeroed
    act[0] = ctr[2];
    act[1] = ctr[3];
    ctr[0] = ctr[1] = ctr[2] = ctr[3] = 0;
  }
  if (i < Bt.N) {
    const int2 kr = key_rank[i];
    perm[offs[(kr.x >> 3) * 9 + (kr.x & 7)] + kr.y] = i;
  }
  // block i is near when a non-empty block of its simulation has it in its
  // tile (blocks i - {0,1}^3 per axis); near blocks are appended to the
  // near list
  bool is_near = false;
  if (i < Bt.nblk) {
    const DevModel& M = cand_of_block(Bt, i);
    const int nb = M.nb, lb = i - M.b0;
    const int b0 = lb / (nb * nb), b1 = (lb / nb) % nb, b2 = lb % nb;
    int f[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const int o0 = b0 - (d >> 2), o1 = b1 - ((d >> 1) & 1), o2 = b2 - (d & 1);
      f[d] = (o0 >= 0 && o1 >= 0 && o2 >= 0) ? occ[M.b0 + (o0 * nb + o1) * nb + o2] : 0;
    }
#pragma unroll
    for (int d = 0; d < 8; ++d) is_near = is_near || f[d] != 0;
    near[i] = is_near ? 1 : 0;
  }
  if constexpr (WG == kWave) {
    const unsigned long long m = __ballot(is_near);
    if (m == 0) return;
    const int leader = __ffsll(m) - 1;
    int base = 0;
    if (__lane_id() == leader) base = atomicAdd(&nlist[0], __popcll(m));
    base = __shfl(base, leader);
    if (is_near) nlist[1 + base + __popcll(m & ((1ull << __lane_id()) - 1ull))] = i;
  } else {
    int ntot;
    const int nx = wg_exclusive_scan(is_near ? 1 : 0, smem, ntot);
    if (threadIdx.x == 0) nbase = ntot > 0 ? atomicAdd(&nlist[0], ntot) : 0;
    __syncthreads();
    if (is_near) nlist[1 + nbase + nx] = i;
  }
}

// Deterministic mode, after k_bin_place: each cell's particles in perm in
// ascending index order (k_bin_place leaves them in the order of the counting
// atomics).  One workgroup per active block at a time: the block's range is
// copied (to LDS, or for a block of more than kBinDetLds particles to the
// rank fields of key_rank, free once k_bin_place has read them and rewritten
// by the next counting step), then each particle goes to its cell's start +
// the number of smaller indices in the cell.
constexpr int kBinDetWG = 256;
constexpr int kBinDetLds = 4096;
__global__ __launch_bounds__(kBinDetWG) void k_bin_det(const int* __restrict__ act, int nblk,
                                                       const int* __restrict__ offs, int* __restrict__ perm,
                                                       int2* __restrict__ key_rank) {
  __shared__ int lcopy[kBinDetLds];
  int* __restrict__ gcopy = reinterpret_cast<int*>(key_rank);  // gcopy[2 i + 1]: perm[i]
  const int nh = act[0], nact = nh + act[1], last = nblk - 1;
  for (int e = blockIdx.x; e < nact; e += gridDim.x) {
    const int blk = e < nh ? act[2 + e] : act[2 + last - (e - nh)];
    int cs[9];
#pragma unroll
    for (int c = 0; c <= 8; ++c) cs[c] = offs[blk * 9 + c];
    const int b0 = cs[0];
    if (cs[8] - b0 <= kBinDetLds) {
      for (int i = b0 + (int)threadIdx.x; i < cs[8]; i += blockDim.x) lcopy[i - b0] = perm[i];
      __syncthreads();
      for (int i = b0 + (int)threadIdx.x; i < cs[8]; i += blockDim.x) {
        int c = 0;
#pragma unroll
        for (int k = 1; k < 8; ++k) c = i >= cs[k] ? k : c;
        const int p = lcopy[i - b0];
        int r = 0;
#pragma unroll 8
        for (int j = cs[c] - b0; j < cs[c + 1] - b0; ++j) r += lcopy[j] < p ? 1 : 0;
        perm[cs[c] + r] = p;
      }
    } else {
      for (int i = b0 + (int)threadIdx.x; i < cs[8]; i += blockDim.x) gcopy[2 * (size_t)i + 1] = perm[i];
      __syncthreads();
      for (int i = b0 + (int)threadIdx.x; i < cs[8]; i += blockDim.x) {
        int c = 0;
#pragma unroll
        for (int k = 1; k < 8; ++k) c = i >= cs[k] ? k : c;
        const int p = gcopy[2 * (size_t)i + 1];
        int r = 0;
#pragma unroll 8
        for (int j = cs[c]; j < cs[c + 1]; ++j) r += gcopy[2 * (size_t)j + 1] < p ? 1 : 0;
        perm[cs[c] + r] = p;
      }
    }
    __syncthreads();
  }
}

__global__ void k_mu_lam(int n,const float* __restrict__ E, const float* __restrict__ nu, float* __restrict__ mu,
                         float* __restrict__ lam) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float e = E[p], v = nu[p];
  mu[p] = e / (2.f * (1.f + v));
  lam[p] = e * v / ((1.f + v) * (1.f - 2.f * v));
}

__global__ void k_mu_lam_adj(int n, const float* __restrict__ E, const float* __restrict__ nu,
                             const float* __restrict__ dmu, const float* __restrict__ dlam, float* __restrict__ dE,
                             float* __restrict__ dnu) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float e = E[p], v = nu[p];
  const float q = (1.f + v) * (1.f - 2.f * v);
  const float gm = dmu ? dmu[p] : 0.f, gl = dlam ? dlam[p] : 0.f;
  if (dE) dE[p] += gm / (2.f * (1.f + v)) + gl * v / q;
  if (dnu) dnu[p] += gm * (-e / (2.f * (1.f + v) * (1.f + v))) + gl * e * (1.f + 2.f * v * v) / (q * q);
}

__global__ void k_mass(int n, const float* __restrict__ rho, const float* __restrict__ vol, float* __restrict__ mass) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  mass[p] = rho[p] * vol[p];
}

// Nearest neighbour in B of every point of A: min_j |a_i - b_j|^2 and the
// smallest j attaining it (the knn_points(K = 1) step of the Chamfer loss).
// Each thread owns one point of A; B passes through LDS in tiles that every
// lane reads at the same address (a broadcast).  A few thousand points of A
// are too few threads for 256 CUs, so B is split over blockIdx.y in chunks
// of `chunk` points: split s writes its best (distance, index) to
// dist/idx + s * n and k_nn_reduce takes the minimum over the splits, in
// split order with a strict comparison, which keeps the smallest index.
// Workgroups of four waves share each 32 KB tile: one-wave workgroups would
// hold a tile each and fit only five waves on a CU.
constexpr int kNnTile = 2048;
constexpr int kNnBlock = 256;
unsigned nn_blocks(size_t n) { return (unsigned)((n + kNnBlock - 1) / kNnBlock); }
__global__ __launch_bounds__(kNnBlock) void k_nn_sqdist(int n, const float* __restrict__ a, int m,
                                                      const float* __restrict__ b, int chunk,
                                                      float* __restrict__ dist, int* __restrict__ idx) {
  __shared__ float4 tile[kNnTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * chunk;
  const int j1 = min(m, j0 + chunk);
  float ax = 0.f, ay = 0.f, az = 0.f;
  if (i < n) {
    ax = a[3 * (size_t)i];
    ay = a[3 * (size_t)i + 1];
    az = a[3 * (size_t)i + 2];
  }
  float best = __builtin_huge_valf();
  int bj = j0;
  for (int t0 = j0; t0 < j1; t0 += kNnTile) {
    const int cnt = j1 - t0 < kNnTile ? j1 - t0 : kNnTile;
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
      const size_t q = 3 * (size_t)(t0 + k);
      tile[k] = make_float4(b[q], b[q + 1], b[q + 2], 0.f);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < cnt; ++k) {
      const float4 q = tile[k];
      const float dx = ax - q.x, dy = ay - q.y, dz = az - q.z;
      const float d = dx * dx + dy * dy + dz * dz;
      if (d < best) {  // strict: the first (smallest) index of a tie is kept
        best = d;
        bj = t0 + k;
      }
    }
  }
  if (i < n) {
    const size_t o = (size_t)blockIdx.y * n + i;
    dist[o] = best;
    idx[o] = bj;
  }
}

__global__ __launch_bounds__(kBlock) void k_nn_reduce(int n, int splits, const float* __restrict__ pd,
                                                      const int* __restrict__ pi, float* __restrict__ dist,
                                                      int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float best = pd[i];
  int bj = pi[i];
  for (int s = 1; s < splits; ++s) {
    const float d = pd[(size_t)s * n + i];
    if (d < best) {
      best = d;
      bj = pi[(size_t)s * n + i];
    }
  }
  dist[i] = best;
  idx[i] = bj;
}

// The same search over segments: segment g pairs A[a_off[g], a_off[g+1])
// with B[b_off[g], b_off[g+1]) (blockIdx.z = g; the frames of the training
// loss in one launch).  Indices are global into B; a query whose B segment
// is empty gets +inf and -1.  Blocks past their segment's end exit together.
__global__ __launch_bounds__(kNnBlock) void k_nn_sqdist_seg(const int* __restrict__ a_off,
                                                          const int* __restrict__ b_off,
                                                          const float* __restrict__ a,
                                                          const float* __restrict__ b, int splits, int n_a,
                                                          float* __restrict__ dist, int* __restrict__ idx) {
  __shared__ float4 tile[kNnTile];
  const int g = blockIdx.z;
  const int a0 = a_off[g], a1 = a_off[g + 1];
  if (a0 + (int)(blockIdx.x * blockDim.x) >= a1) return;  // uniform over the block
  const int b0 = b_off[g], b1 = b_off[g + 1];
  const int chunk = (b1 - b0 + splits - 1) / splits;
  const int j0 = b0 + (int)blockIdx.y * chunk;
  const int j1 = min(b1, j0 + chunk);
  const int i = a0 + blockIdx.x * blockDim.x + threadIdx.x;
  float ax = 0.f, ay = 0.f, az = 0.f;
  if (i < a1) {
    ax = a[3 * (size_t)i];
    ay = a[3 * (size_t)i + 1];
    az = a[3 * (size_t)i + 2];
  }
  float best = __builtin_huge_valf();
  int bj = -1;
  for (int t0 = j0; t0 < j1; t0 += kNnTile) {
    const int cnt = j1 - t0 < kNnTile ? j1 - t0 : kNnTile;
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
      const size_t q = 3 * (size_t)(t0 + k);
      tile[k] = make_float4(b[q], b[q + 1], b[q + 2], 0.f);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < cnt; ++k) {
      const float4 q = tile[k];
      const float dx = ax - q.x, dy = ay - q.y, dz = az - q.z;
      const float d = dx * dx + dy * dy + dz * dz;
      if (d < best) {
        best = d;
        bj = t0 + k;
      }
    }
  }
  if (i < a1) {
    const size_t o = (size_t)blockIdx.y * n_a + i;
    dist[o] = best;
    idx[o] = bj;
  }
}

// Scratch of empm_nn_sqdist_noalloc: the per-split partial results (a float
// distance and an int index per (split, query) pair), in a device array that
// the library owns for the life of the process, so that the call allocates
// nothing.  nn_splits makes s = ceil(2048 / ceil(n / 256)) splits at most, so
// s * n < 2048 * 256 + n <= 2^20 pairs whenever s > 1 (s > 1 needs
// n <= 2047 * 256): 8 MiB holds every split plan.
constexpr size_t kNnScratchPairs = (size_t)1 << 20;
static __device__ __align__(256) float scratch[2 * kNnScratchPairs];

// Splits of B for n points of A: enough blocks for 8 per CU, at least 512
// points of B per split.
int nn_splits(int n, int m, int n_seg = 1) {
  const int qb = n_seg * ((n + kNnBlock - 1) / kNnBlock);
  int s = (2048 + qb - 1) / qb;
  s = std::min(s, std::max(1, m / 512));
  if (s < 1) s = 1;
  const int chunk = (m + s - 1) / s;
  return chunk > 0 ? (m + chunk - 1) / chunk : 1;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Per-kernel timing with HIP events recorded on the launch stream around the
// kernels selected by a mask (instrumentation for the benchmark; off unless
// empm_profile_enable is called).
struct KernelProfile {
  unsigned mask = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[EMPM_NUM_KERNELS];
  std::vector<hipEvent_t> pool;
  long launches[EMPM_NUM_KERNELS] = {};
  double ms[EMPM_NUM_KERNELS] = {};
};

struct empm_ctx {
  int device = 0;
  int n_grid = 0;            // grid of a single-simulation context (0: batch context)
  int max_cand = 0;
  int max_particles = 0;
  long long max_nodes = 0;   // node arena capacity
  long long max_blocks = 0;  // block arena capacity
  float4* tiles = nullptr;   // max_blocks tiles of kTileNodes float4 (block-owned scatter)
  float4* keep = nullptr;    // node arena: (mv, m) of the substep being differentiated
  float4* v_out = nullptr;   // node arena
  float4* gadj_in = nullptr; // node arena: dL/d(mv, m)
  float* Fe = nullptr;       // N*9 scratch
  float* S = nullptr;        // N*9: stress part of the P2G payload (written by k_stress / k_g2p)
  float* payload = nullptr;  // N*kAdjPayload: G2P-adjoint scatter payload
  float* xbar = nullptr;     // N*3 scratch
  float* Febar = nullptr;    // N*9 scratch
  int* bins = nullptr;       // the context's own binning record
  int* count = nullptr;      // max_blocks * 8: per-bin counts, zero between binnings
  bool counts_pending = false;  // a G2P counted the next substep's bins, not yet binned
  int2* key_rank = nullptr;  // N: bin key and rank of each particle
  int* blkoff = nullptr;     // max_blocks: workgroup-local exclusive block offsets
  int* wsum = nullptr;       // block-scan workgroup totals
  int* bin_ctr = nullptr;    // 4: k_bin_alloc's particle-range counter and (at 2, 64-bit) active-list counts, zero between binnings
  int* work = nullptr;       // 2: work counters of the k_p2g and k_g2p_adj_scatter launches
  empm_tuning tune = {kBlkWG, 0, kHeavyChunks, kDynamicRounds, kNodePassGridCap};
  bool deterministic = false;  // empm_set_deterministic: fixed float summation orders
  int cus = 0;               // compute units of the device
  int scatter_res[3][2] = {};  // scatter workgroups resident per CU: [128, 256, 512 threads][k_p2g, k_g2p_adj_scatter]
  DevModel* cand_dev = nullptr;     // device table of the simulations of the current call
  DevModel* cand_pinned = nullptr;  // its pinned staging copy
  std::vector<DevModel> cand_last;  // what cand_dev holds
  hipEvent_t cand_upload = nullptr; // completion of the last table upload
  int* status = nullptr;            // max_cand: per-simulation failure bits (EMPM_STATUS_*), set by k_g2p
  int* status_host = nullptr;       // its pinned read-back copy
  // per-substep schedule (time, collider table) of the current call, read by
  // the node passes; uploaded from a two-slot pinned ring
  ColTable* sched = nullptr;        // kSchedCap entries
  ColTable* sched_host[2] = {nullptr, nullptr};
  hipEvent_t sched_ev[2] = {nullptr, nullptr};
  int sched_slot = 0;
  hipStream_t last_stream = nullptr;  // stream of the previous call (calls on another stream are ordered after it)
  bool have_last_stream = false;
  hipEvent_t order_ev = nullptr;
  // HIP graph replay of whole calls (empm_set_graphs)
  bool graphs = false;
  hipStream_t gstream = nullptr;    // capture / replay stream (default priority)
  hipEvent_t gev_in = nullptr, gev_out = nullptr;
  struct GraphEntry {
    std::vector<unsigned long long> key;
    hipGraphExec_t exec = nullptr;
    bool pending_after = false;
    unsigned long long used = 0;
  };
  std::vector<GraphEntry> gcache;
  unsigned long long gclock = 0;
  long long graph_captures = 0, graph_replays = 0;
  KernelProfile* prof = nullptr;
};

namespace {

const char* const kKernelNames[EMPM_NUM_KERNELS] = {"k_p2g",     "k_grid",    "k_g2p",    "k_g2p_adj",
                                                    "k_grid_adj", "k_p2g_adj", "k_stress", "k_g2p_adj_scatter",
                                                    "k_bin"};

hipEvent_t pool_get(KernelProfile* p) {
  if (!p->pool.empty()) {
    hipEvent_t e = p->pool.back();
    p->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// RAII guard: records a start event before and a stop event after the launch
// it brackets, when the kernel is selected.
struct ProfScope {
  KernelProfile* p;
  int id;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(empm_ctx* ctx, int id_, hipStream_t st_) : p(ctx->prof), id(id_), st(st_) {
    if (p && (p->mask & (1u << id))) {
      a = pool_get(p);
      b = pool_get(p);
      if (a && b) (void)hipEventRecord(a, st);
    }
  }
  ~ProfScope() {
    if (a && b) {
      (void)hipEventRecord(b, st);
      p->pending[id].push_back({a, b});
    }
  }
};

int blocks_per_axis(int n) { return (n + kBlkCells - 1) / kBlkCells; }

// The simulations of one call on the host: the device-side table entries,
// arena totals and, per material, the contiguous particle ranges whose
// simulations use it (the constitutive kernels are launched per material).
struct HostBatch {
  std::vector<DevModel> cand;
  int N = 0;
  int nblk = 0;
  long long G = 0;
  int nwaves = 0;
  bool has_material[8] = {};
  bool mutates[8] = {};     // a simulation of this material rewrites (yield stress, mu, lam)
  std::vector<int2> mlist;  // Batch::mlist, material-major
  int moff[9] = {};
  int mwaves[8] = {};
  Batch dev{};
};

unsigned waves_grid(int nwaves) {
  const size_t t = (size_t)nwaves * kWaveSize;
  return (unsigned)((t + kBlock - 1) / kBlock > 0 ? (t + kBlock - 1) / kBlock : 1);
}
unsigned particle_grid(const HostBatch& hb) { return waves_grid(hb.nwaves); }
unsigned material_grid(const HostBatch& hb, int mat) { return waves_grid(hb.mwaves[mat]); }

int check_particles(const empm_particles* s, int N, const char* what) {
  if (!s) return set_err(EMPM_E_INVALID_ARG, "null %s", what);
  if (N > 0 && (!s->x || !s->v || !s->C || !s->F)) return set_err(EMPM_E_INVALID_ARG, "null array in %s", what);
  return EMPM_OK;
}

int check_scalars(const empm_model* m, int c) {
  if (m->material < 0 || m->material > 7)
    return set_err(EMPM_E_UNSUPPORTED_MATERIAL, "Undefined material type %d (simulation %d)", m->material, c);
  if (m->n_grid < 4 || m->n_grid > 1024)
    return set_err(EMPM_E_INVALID_ARG, "simulation %d: n_grid=%d outside [4, 1024]", c, m->n_grid);
  if (m->n_particles < 0) return set_err(EMPM_E_INVALID_ARG, "simulation %d: n_particles < 0", c);
  if (!(m->inv_dx > 0.f) || !(m->dx > 0.f))
    return set_err(EMPM_E_INVALID_ARG, "simulation %d: dx/inv_dx must be positive", c);
  return EMPM_OK;
}

// Validate a batch against the context and build its host description.
void index_waves(HostBatch& hb);

int build_batch(const empm_ctx* ctx, const empm_batch* b, HostBatch& hb) {
  if (!ctx || !b) return set_err(EMPM_E_INVALID_ARG, "null context or batch");
  if (b->n_cand < 1 || b->n_cand > ctx->max_cand)
    return set_err(EMPM_E_INVALID_ARG, "n_cand=%d outside [1, %d]", b->n_cand, ctx->max_cand);
  if (!b->models) return set_err(EMPM_E_INVALID_ARG, "null simulation table");
  hb.cand.assign((size_t)b->n_cand, DevModel{});
  long long N = 0, nblk = 0, G = 0;
  for (int c = 0; c < b->n_cand; ++c) {
    const empm_model* m = &b->models[c];
    int rc = check_scalars(m, c);
    if (rc) return rc;
    DevModel& d = hb.cand[(size_t)c];
    d.N = m->n_particles;
    d.n = m->n_grid;
    d.material = m->material;
    d.nb = blocks_per_axis(m->n_grid);
    d.p0 = (int)N;
    d.b0 = (int)nblk;
    d.g0 = G;
    d.grid_lim = m->grid_lim;
    d.dx = m->dx;
    d.inv_dx = m->inv_dx;
    d.rpic = m->rpic_damping;
    d.damp_scale = m->grid_v_damping_scale;
    d.g[0] = m->gravity[0];
    d.g[1] = m->gravity[1];
    d.g[2] = m->gravity[2];
    d.mp.material = m->material;
    d.mp.alpha = m->alpha;
    d.mp.hardening = m->hardening;
    d.mp.xi = m->xi;
    d.mp.plastic_viscosity = m->plastic_viscosity;
    d.mp.softening = m->softening;
    d.mu = b->mu;
    d.lam = b->lam;
    d.ys = b->yield_stress;
    d.mass = b->mass;
    d.vol = b->vol;
    d.sel = b->selection;
    N += m->n_particles;
    nblk += (long long)d.nb * d.nb * d.nb;
    G += (long long)m->n_grid * m->n_grid * m->n_grid;
  }
  if (N != b->n_particles)
    return set_err(EMPM_E_INVALID_ARG, "n_particles=%d but the simulations hold %lld", b->n_particles, N);
  if (N > ctx->max_particles)
    return set_err(EMPM_E_INVALID_ARG, "%lld particles exceed the context capacity %d", N, ctx->max_particles);
  if (G > ctx->max_nodes || nblk > ctx->max_blocks)
    return set_err(EMPM_E_INVALID_ARG, "grids of %lld nodes exceed the context capacity %lld", G, ctx->max_nodes);
  if (N > 0 && (!b->mu || !b->lam || !b->yield_stress || !b->mass || !b->vol))
    return set_err(EMPM_E_INVALID_ARG, "null per-particle model array");
  hb.N = (int)N;
  hb.nblk = (int)nblk;
  hb.G = G;
  index_waves(hb);
  return EMPM_OK;
}

// Wave ranges of per-particle launches: each simulation's first wave (w0),
// the total, and the per-material lists.
void index_waves(HostBatch& hb) {
  int w = 0;
  for (int m = 0; m < 8; ++m) hb.has_material[m] = hb.mutates[m] = false;
  for (DevModel& d : hb.cand) {
    d.w0 = w;
    w += (d.N + kWaveSize - 1) / kWaveSize;
    if (d.N > 0) {
      hb.has_material[d.material] = true;
      if (d.material == 5 || ((d.material == 1 || d.material == 7) && d.mp.hardening == 1.f))
        hb.mutates[d.material] = true;
    }
  }
  hb.nwaves = w;
  hb.mlist.clear();
  for (int m = 0; m < 8; ++m) {
    hb.moff[m] = (int)hb.mlist.size();
    int mw = 0;
    for (size_t c = 0; c < hb.cand.size(); ++c) {
      const DevModel& d = hb.cand[c];
      if (d.material != m || d.N == 0) continue;
      hb.mlist.push_back(make_int2((int)c, mw));
      mw += (d.N + kWaveSize - 1) / kWaveSize;
    }
    hb.mwaves[m] = mw;
  }
  hb.moff[8] = (int)hb.mlist.size();
}

// Make the device table hold hb's simulations (stream-ordered; uploads only
// when the table changed since the last call).
int upload_batch(empm_ctx* ctx, HostBatch& hb, hipStream_t st) {
  const size_t nc = hb.cand.size();
  const bool same = ctx->cand_last.size() == nc && memcmp(ctx->cand_last.data(), hb.cand.data(),
                                                             nc * sizeof(DevModel)) == 0;
  // the material lists follow the table in the same buffers (they are a
  // function of it, so an unchanged table means unchanged lists)
  int2* mlist_dev = reinterpret_cast<int2*>(ctx->cand_dev + ctx->max_cand);
  if (!same) {
    HIP_TRY(hipEventSynchronize(ctx->cand_upload));  // the staging copy is free again
    memcpy(ctx->cand_pinned, hb.cand.data(), nc * sizeof(DevModel));
    int2* mlist_pinned = reinterpret_cast<int2*>(ctx->cand_pinned + ctx->max_cand);
    memcpy(mlist_pinned, hb.mlist.data(), hb.mlist.size() * sizeof(int2));
    HIP_TRY(hipMemcpyAsync(ctx->cand_dev, ctx->cand_pinned, nc * sizeof(DevModel), hipMemcpyHostToDevice, st));
    if (!hb.mlist.empty())
      HIP_TRY(hipMemcpyAsync(mlist_dev, mlist_pinned, hb.mlist.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(ctx->cand_upload, st));
    ctx->cand_last = hb.cand;
  }
  hb.dev = Batch{ctx->cand_dev, (int)nc, hb.N, hb.nblk, hb.nwaves, mlist_dev, {}, {}};
  for (int m = 0; m <= 8; ++m) hb.dev.moff[m] = hb.moff[m];
  for (int m = 0; m < 8; ++m) hb.dev.mwaves[m] = hb.mwaves[m];
  return EMPM_OK;
}

// A single simulation as a batch of one.
empm_batch batch_of(const empm_model* m) {
  empm_batch b;
  memset(&b, 0, sizeof(b));
  b.n_cand = 1;
  b.n_particles = m ? m->n_particles : 0;
  b.models = m;
  if (m) {
    b.mu = m->mu;
    b.lam = m->lam;
    b.yield_stress = m->yield_stress;
    b.mass = m->mass;
    b.vol = m->vol;
    b.selection = m->selection;
  }
  return b;
}

int prepare(empm_ctx* ctx, const empm_batch* b, HostBatch& hb, hipStream_t st) {
  int rc = build_batch(ctx, b, hb);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  return upload_batch(ctx, hb, st);
}

int make_table(const empm_collider* cols, int ncol, ColTable& T) {
  if (ncol < 0 || ncol > EMPM_MAX_COLLIDERS)
    return set_err(EMPM_E_INVALID_ARG, "ncol=%d outside [0, %d]", ncol, EMPM_MAX_COLLIDERS);
  if (ncol > 0 && !cols) return set_err(EMPM_E_INVALID_ARG, "null collider table");
  memset(&T, 0, sizeof(T));
  for (int k = 0; k < ncol; ++k) {
    if (cols[k].kind < 0 || cols[k].kind > 5)
      return set_err(EMPM_E_INVALID_ARG, "collider %d: unknown kind %d", k, cols[k].kind);
    if (cols[k].kind == 5 && !cols[k].mask) return set_err(EMPM_E_INVALID_ARG, "collider %d: null mask", k);
    T.c[k] = cols[k];
  }
  T.n = ncol;
  return EMPM_OK;
}

// Launch KERNEL<material>(args...): the constitutive kernels are compiled
// once per material so each runs a single branch-free code path.
#define EMPM_LAUNCH_MAT(material, KERNEL, grid, block, stream, ...)                              \
  switch (material) {                                                                           \
    case 0: hipLaunchKernelGGL(KERNEL<0>, grid, block, 0, stream, __VA_ARGS__); break;         \
    case 1: hipLaunchKernelGGL(KERNEL<1>, grid, block, 0, stream, __VA_ARGS__); break;         \
    case 2: hipLaunchKernelGGL(KERNEL<2>, grid, block, 0, stream, __VA_ARGS__); break;         \
    case 3: hipLaunchKernelGGL(KERNEL<3>, grid, block, 0, stream, __VA_ARGS__); break;         \
    case 4: hipLaunchKernelGGL(KERNEL<4>, grid, block, 0, stream, __VA_ARGS__); break;         \
    case 5: hipLaunchKernelGGL(KERNEL<5>, grid, block, 0, stream, __VA_ARGS__); break;         \
    case 6: hipLaunchKernelGGL(KERNEL<6>, grid, block, 0, stream, __VA_ARGS__); break;         \
    default: hipLaunchKernelGGL(KERNEL<7>, grid, block, 0, stream, __VA_ARGS__); break;        \
  }

// KERNEL<mat>(Bt, args...) over the particle waves of material mat, once
// per material present
#define EMPM_LAUNCH_RANGES(hb, KERNEL, stream, ...)                                              \
  for (int mat_ = 0; mat_ < 8; ++mat_)                                                          \
    if ((hb).has_material[mat_]) {                                                              \
      EMPM_LAUNCH_MAT(mat_, KERNEL, dim3(material_grid(hb, mat_)), dim3(kBlock), stream,        \
                      (hb).dev, __VA_ARGS__);                                                   \
    }

// k_p2g_adj<mat, mutating, FUSE> over the particle waves of each material
// present; the mutating variant only for materials 1, 5 and 7 when one of
// their simulations rewrites its parameters
template <bool FUSE, typename... A>
void launch_p2g_adj(const HostBatch& hb, hipStream_t stream, A... args) {
  for (int mat = 0; mat < 8; ++mat) {
    if (!hb.has_material[mat]) continue;
    const dim3 g(material_grid(hb, mat)), b(kBlock);
    const bool mut = hb.mutates[mat];
    switch (mat) {
      case 0: hipLaunchKernelGGL((k_p2g_adj<0, false, FUSE>), g, b, 0, stream, hb.dev, args...); break;
      case 1:
        if (mut) hipLaunchKernelGGL((k_p2g_adj<1, true, FUSE>), g, b, 0, stream, hb.dev, args...);
        else hipLaunchKernelGGL((k_p2g_adj<1, false, FUSE>), g, b, 0, stream, hb.dev, args...);
        break;
      case 2: hipLaunchKernelGGL((k_p2g_adj<2, false, FUSE>), g, b, 0, stream, hb.dev, args...); break;
      case 3: hipLaunchKernelGGL((k_p2g_adj<3, false, FUSE>), g, b, 0, stream, hb.dev, args...); break;
      case 4: hipLaunchKernelGGL((k_p2g_adj<4, false, FUSE>), g, b, 0, stream, hb.dev, args...); break;
      case 5: hipLaunchKernelGGL((k_p2g_adj<5, true, FUSE>), g, b, 0, stream, hb.dev, args...); break;
      case 6: hipLaunchKernelGGL((k_p2g_adj<6, false, FUSE>), g, b, 0, stream, hb.dev, args...); break;
      default:
        if (mut) hipLaunchKernelGGL((k_p2g_adj<7, true, FUSE>), g, b, 0, stream, hb.dev, args...);
        else hipLaunchKernelGGL((k_p2g_adj<7, false, FUSE>), g, b, 0, stream, hb.dev, args...);
        break;
    }
  }
}

size_t pad4(size_t k) { return (k + 3) & ~(size_t)3; }

// Layout of one binning record: [perm N | offs 9 nblk | act 2 + nblk |
// occ nblk | near nblk | nlist 1 + nblk], each section padded to a multiple
// of 4 ints so that the int4 accesses of the binning kernels stay 16-byte
// aligned.
size_t bin_record_ints(long long nblk, long long n_particles) {
  return pad4((size_t)n_particles) + pad4((size_t)nblk * 9) + pad4(2 + (size_t)nblk) +
         pad4(1 + (size_t)nblk) + 2 * pad4((size_t)nblk);
}

BinBuf bin_buf(int* rec, int nblk, int n_particles) {
  BinBuf b;
  b.perm = rec;
  b.offs = b.perm + pad4((size_t)n_particles);
  b.act = b.offs + pad4((size_t)nblk * 9);
  b.occ = b.act + pad4(2 + (size_t)nblk);
  b.near = b.occ + pad4((size_t)nblk);
  b.nlist = b.near + pad4((size_t)nblk);
  return b;
}

Binning binning(const empm_ctx* ctx, const BinBuf& b, int nblk) {
  return Binning{b.perm, b.offs, b.act, nblk, b.occ, b.near, b.nlist + 1, b.nlist, ctx->tune.dynamic_rounds};
}

int wg_variant(int wg) { return wg == 128 ? 0 : (wg == 512 ? 2 : 1); }

// Launch width of a block-owned scatter (idx 0: k_p2g, 1: k_g2p_adj_scatter)
// with the tuned workgroup size: the workgroups resident at once, or the
// tuning's count per CU when that is lower.
unsigned scatter_grid(const empm_ctx* ctx, int idx, int wg) {
  int per_cu = ctx->scatter_res[wg_variant(wg)][idx];
  if (ctx->tune.scatter_per_cu > 0 && (per_cu <= 0 || ctx->tune.scatter_per_cu < per_cu))
    per_cu = ctx->tune.scatter_per_cu;
  const long long w = (long long)per_cu * ctx->cus;
  return w > 0 ? (unsigned)w : (unsigned)kBlkGridCap;
}

// KERNEL<tuned workgroup size> of the block-owned scatters; in deterministic
// mode KERNEL<256, true> (per-wave tiles; the workgroup size is fixed so the
// summation order does not follow the tuning)
#define EMPM_LAUNCH_SCATTER(ctx, idx, KERNEL, stream, ...)                                                    \
  if ((ctx)->deterministic) {                                                                                \
    hipLaunchKernelGGL((KERNEL<256, true>), dim3(scatter_grid(ctx, idx, 256)), dim3(256), 0, stream,        \
                       __VA_ARGS__);                                                                         \
  } else {                                                                                                   \
    switch ((ctx)->tune.scatter_wg) {                                                                        \
      case 128:                                                                                              \
        hipLaunchKernelGGL(KERNEL<128>, dim3(scatter_grid(ctx, idx, 128)), dim3(128), 0, stream, __VA_ARGS__); \
        break;                                                                                               \
      case 512:                                                                                              \
        hipLaunchKernelGGL(KERNEL<512>, dim3(scatter_grid(ctx, idx, 512)), dim3(512), 0, stream, __VA_ARGS__); \
        break;                                                                                               \
      default:                                                                                               \
        hipLaunchKernelGGL(KERNEL<256>, dim3(scatter_grid(ctx, idx, 256)), dim3(256), 0, stream, __VA_ARGS__); \
        break;                                                                                               \
    }                                                                                                        \
  }


unsigned nblocks(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }


dim3 node_grid(const empm_ctx* ctx, int nblk) {
  const size_t wgs = ((size_t)nblk * 8 + kBlock - 1) / kBlock, cap = (size_t)ctx->tune.node_grid_cap;
  return dim3((unsigned)(wgs < cap ? (wgs > 0 ? wgs : 1) : cap));
}

// Bin the particles of a state (positions x) into `out`.
// With `counted`, the counting step was done by the previous substep's G2P
// on these positions (ctx->counts_pending); else it runs here, after
// clearing the counts of a G2P count whose binning never ran (a call that
// stopped between the two).  `ordered` lays the blocks' particle ranges out
// in block order (k_bin_scan + k_bin_finish: the simulations of a batch
// keep their arena ranges); else they are allocated in one launch
// (k_bin_alloc), in no particular block order.
int do_bin(empm_ctx* ctx, const HostBatch& hb, const float* x, const BinBuf& out, hipStream_t st,
           bool counted = false, bool ordered = false) {
  ProfScope ps(ctx, EMPM_KERNEL_BIN, st);
  const int nblk = hb.nblk;
  const unsigned wgs = (unsigned)((nblk + kBinScanWG - 1) / kBinScanWG);
  // a large batch's block arena: more than two k_bin_alloc workgroups per CU
  // at one block per thread
  const bool large = wgs > 2u * (unsigned)(ctx->cus > 0 ? ctx->cus : 256);
  if (!(counted && ctx->counts_pending)) {
    if (ctx->counts_pending) {
      const size_t nc = (size_t)ctx->max_blocks * 8;
      hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nc + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, ctx->count,
                         nc, 0);
    }
    if (hb.N > 0)
      hipLaunchKernelGGL(k_bin_count, dim3(nblocks(hb.N)), dim3(kBlock), 0, st, hb.dev, x, ctx->count,
                         ctx->key_rank);
  }
  ctx->counts_pending = false;
  if (ordered) {
    hipLaunchKernelGGL(k_bin_scan, dim3(wgs), dim3(kBinScanWG), 0, st, nblk, ctx->count, ctx->blkoff, ctx->wsum,
                       out.act, out.nlist, ctx->work);
    hipLaunchKernelGGL(k_bin_finish, dim3(wgs), dim3(kBinScanWG), 0, st, hb.dev, ctx->count, ctx->blkoff,
                       ctx->wsum, out.offs, out.act, out.occ, ctx->tune.heavy_chunks);
  } else {
    // four blocks per thread for a large batch's block arena
    if (large) {
      const unsigned wgs4 = (unsigned)((nblk + 4 * kBinScanWG - 1) / (4 * kBinScanWG));
      hipLaunchKernelGGL(k_bin_alloc<4>, dim3(wgs4), dim3(kBinScanWG), 0, st, hb.dev, ctx->count, out.offs,
                         out.act, out.occ, out.nlist, ctx->work, ctx->bin_ctr, ctx->tune.heavy_chunks);
    } else {
      hipLaunchKernelGGL(k_bin_alloc<1>, dim3(wgs), dim3(kBinScanWG), 0, st, hb.dev, ctx->count, out.offs,
                         out.act, out.occ, out.nlist, ctx->work, ctx->bin_ctr, ctx->tune.heavy_chunks);
    }
  }
  const size_t nplace = (size_t)(hb.N > nblk ? hb.N : nblk);
  if (large)
    hipLaunchKernelGGL(k_bin_place<kBinScanWG>, dim3((unsigned)((nplace + kBinScanWG - 1) / kBinScanWG)),
                       dim3(kBinScanWG), 0, st, hb.dev, ctx->key_rank, out.offs, out.perm, out.occ, out.near,
                       out.nlist, out.act, ordered ? nullptr : ctx->bin_ctr);
  else
    hipLaunchKernelGGL(k_bin_place<kWave>, dim3(nblocks(nplace)), dim3(kWave), 0, st, hb.dev, ctx->key_rank,
                       out.offs, out.perm, out.occ, out.near, out.nlist, out.act, ordered ? nullptr : ctx->bin_ctr);
  if (ctx->deterministic && hb.N > 0) {
    const unsigned cap = 8u * (unsigned)(ctx->cus > 0 ? ctx->cus : 256);
    hipLaunchKernelGGL(k_bin_det, dim3((unsigned)nblk < cap ? (unsigned)(nblk > 0 ? nblk : 1) : cap),
                       dim3(kBinDetWG), 0, st, out.act, nblk, out.offs, out.perm, ctx->key_rank);
  }
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

// One forward substep in -> out, binning `in` first (into the record's bins
// when given).  `first`: run the stress stage here (else the previous
// substep's G2P did); `last`: no stress stage for a next substep.
int launch_fwd(empm_ctx* ctx, const HostBatch& hb, const ColTable* T, float dt, const empm_particles* in,
               const empm_particles* out, const empm_grid_rec* rec, bool first, bool last, hipStream_t st,
               bool replay = false) {
  float* Fe = in->Fe ? in->Fe : ctx->Fe;
  float* Fe_next = out->Fe ? out->Fe : ctx->Fe;
  float4* keep = rec ? reinterpret_cast<float4*>(rec->mv_m) : nullptr;
  float4* v_out = rec && rec->v_out ? reinterpret_cast<float4*>(rec->v_out) : ctx->v_out;
  const BinBuf bb = bin_buf(rec && rec->bins ? rec->bins : ctx->bins, hb.nblk, hb.N);
  int rc = do_bin(ctx, hb, in->x, bb, st, !first);
  if (rc) return rc;
  const Binning B = binning(ctx, bb, hb.nblk);
  if (hb.N > 0 && first) {
    ProfScope ps(ctx, EMPM_KERNEL_STRESS, st);
    // a replay starts from the material state recorded with its first state
    EMPM_LAUNCH_RANGES(hb, k_stress, st, dt, in->F, Fe, ctx->S, 1, replay ? in->mstate : nullptr, in->mstate);
  }
  if (hb.N > 0) {
    ProfScope ps(ctx, EMPM_KERNEL_P2G, st);
    EMPM_LAUNCH_SCATTER(ctx, 0, k_p2g, st, hb.dev, in->x, in->v, in->C, ctx->S, B, ctx->tiles, ctx->work);
  }
  {
    ProfScope ps(ctx, EMPM_KERNEL_GRID, st);
    hipLaunchKernelGGL(k_grid, node_grid(ctx, hb.nblk), dim3(kBlock), 0, st, hb.dev, T, dt, B, ctx->tiles, keep,
                       v_out);
  }
  if (hb.N > 0) {
    ProfScope ps(ctx, EMPM_KERNEL_G2P, st);
    // F_trial of an intermediate state is read only through a recorded tape
    // (the next substep starts from Fe), so it is written for the call's
    // last substep and when recording
    // a next substep also gets its binning's counting step from here
    EMPM_LAUNCH_RANGES(hb, k_g2p, st, dt, in->x, Fe, v_out, out->x, out->v, out->C, out->F, last ? 0 : 1, Fe_next,
                       ctx->S, 1, (last || rec) ? 1 : 0, last ? nullptr : ctx->count,
                       last ? nullptr : ctx->key_rank, last ? nullptr : out->mstate, ctx->status);
    if (!last) ctx->counts_pending = true;
  }
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

// Adjoint of one substep.  With a recorded grid and Fe (rec->mv_m, v_out and
// in->Fe) the forward stress, P2G and boundary pass are not recomputed; with
// recorded bins (rec->bins) the particles are not re-binned.
// Substep k of a reverse sweep may be fused with substep k - 1's G2P adjoint
// (`prev` non-null: that substep's recorded state and grid), and its own G2P
// adjoint may already have been done that way by substep k + 1
// (`g2p_adj_done`).
bool fully_recorded(const empm_particles* in, const empm_grid_rec* rec) {
  return rec && rec->mv_m && rec->v_out && in->Fe;
}

int launch_adj(empm_ctx* ctx, const HostBatch& hb, const ColTable* T, float dt, const empm_particles* in,
               const empm_grid_rec* rec, const empm_particles* d_out, const empm_particles* d_in, float* d_mu,
               float* d_lam, float* d_ys, float* d_mass, float* d_vol, int flags, hipStream_t st,
               const empm_particles* prev = nullptr, const empm_grid_rec* prev_rec = nullptr,
               bool g2p_adj_done = false) {
  const bool recorded = fully_recorded(in, rec);
  const bool have_bins = recorded && rec->bins;
  float* Fe = recorded ? in->Fe : ctx->Fe;
  float4* keep = recorded ? reinterpret_cast<float4*>(rec->mv_m) : ctx->keep;
  float4* v_out = recorded ? reinterpret_cast<float4*>(rec->v_out) : ctx->v_out;
  const BinBuf bb = bin_buf(have_bins ? rec->bins : ctx->bins, hb.nblk, hb.N);
  if (!have_bins) {
    int rc = do_bin(ctx, hb, in->x, bb, st);
    if (rc) return rc;
  }
  const Binning B = binning(ctx, bb, hb.nblk);
  if (!recorded) {
    if (hb.N > 0) {
      ProfScope ps(ctx, EMPM_KERNEL_STRESS, st);
      EMPM_LAUNCH_RANGES(hb, k_stress, st, dt, in->F, Fe, ctx->S, 0, in->mstate, nullptr);
    }
    if (hb.N > 0) {
      ProfScope ps(ctx, EMPM_KERNEL_P2G, st);
      EMPM_LAUNCH_SCATTER(ctx, 0, k_p2g, st, hb.dev, in->x, in->v, in->C, ctx->S, B, ctx->tiles, ctx->work);
    }
    {
      ProfScope ps(ctx, EMPM_KERNEL_GRID, st);
      hipLaunchKernelGGL(k_grid, node_grid(ctx, hb.nblk), dim3(kBlock), 0, st, hb.dev, T, dt, B, ctx->tiles, keep,
                         v_out);
    }
  }
  if (hb.N > 0) {
    if (!(g2p_adj_done && recorded)) {
      ProfScope ps(ctx, EMPM_KERNEL_G2P_ADJ, st);
      hipLaunchKernelGGL(k_g2p_adj, dim3(particle_grid(hb)), dim3(kBlock), 0, st, hb.dev, dt, in->x, Fe, v_out,
                         d_out->x, d_out->v, d_out->C, d_out->F, ctx->xbar, ctx->Febar, ctx->payload, ctx->work + 1);
    }
    {
      ProfScope ps(ctx, EMPM_KERNEL_G2P_ADJ_SCATTER, st);
      EMPM_LAUNCH_SCATTER(ctx, 1, k_g2p_adj_scatter, st, hb.dev, dt, in->x, ctx->payload, d_out->C, d_out->F, Fe, B,
                          ctx->tiles, ctx->work + 1);
    }
  }
  {
    ProfScope ps(ctx, EMPM_KERNEL_GRID_ADJ, st);
    hipLaunchKernelGGL(k_grid_adj, node_grid(ctx, hb.nblk), dim3(kBlock), 0, st, hb.dev, T, dt, B, ctx->tiles, keep,
                       ctx->gadj_in, (flags & EMPM_ADJ_DROP_GRID_MASS) ? 1 : 0);
  }
  if (hb.N > 0) {
    ProfScope ps(ctx, EMPM_KERNEL_P2G_ADJ, st);
    if (prev && fully_recorded(prev, prev_rec)) {
      const PrevG2PAdj pg{prev->x, prev->Fe, reinterpret_cast<const float4*>(prev_rec->v_out), ctx->xbar,
                          ctx->Febar, ctx->payload, ctx->work + 1};
      launch_p2g_adj<true>(hb, st, dt, (const float*)in->x, (const float*)in->v, (const float*)in->C,
                           (const float*)in->F, (const float4*)ctx->gadj_in, (const float*)ctx->xbar,
                           (const float*)ctx->Febar, d_in->x, d_in->v, d_in->C, d_in->F, d_mu, d_lam, d_ys, d_mass,
                           d_vol, (const float*)in->mstate, pg);
    } else {
      launch_p2g_adj<false>(hb, st, dt, (const float*)in->x, (const float*)in->v, (const float*)in->C,
                            (const float*)in->F, (const float4*)ctx->gadj_in, (const float*)ctx->xbar,
                            (const float*)ctx->Febar, d_in->x, d_in->v, d_in->C, d_in->F, d_mu, d_lam, d_ys, d_mass,
                            d_vol, (const float*)in->mstate, PrevG2PAdj{});
    }
  }
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

bool any_mutates(const HostBatch& hb) {
  for (int m = 0; m < 8; ++m)
    if (hb.mutates[m]) return true;
  return false;
}

// Materials that rewrite (yield stress, mu, lam) need each substep's starting
// values (states[k].mstate): the model arrays only hold the latest ones.
int check_mstate(const HostBatch& hb, const empm_particles* states, int n, const char* what) {
  if (!any_mutates(hb) || hb.N == 0) return EMPM_OK;
  for (int k = 0; k < n; ++k)
    if (!states[k].mstate)
      return set_err(EMPM_E_INVALID_ARG,
                     "%s: states[%d].mstate is null, but material 5 / hardening == 1 rewrite yield stress, mu and "
                     "lam inside the substep: each substep's starting values are needed",
                     what, k);
  return EMPM_OK;
}

void advance_colliders(empm_collider* cols, int ncol, double t, double dt) {
  for (int k = 0; k < ncol; ++k) {
    empm_collider& c = cols[k];
    if (c.kind != 1 && c.kind != 4) continue;
    if (t >= (double)c.start_time && t < (double)c.end_time)
      for (int d = 0; d < 3; ++d) c.point[d] = (float)((double)c.point[d] + dt * (double)c.velocity[d]);
  }
}

// Allocate every buffer of a context for the given capacities.
int alloc_ctx(empm_ctx* c) {
  const size_t np = c->max_particles > 0 ? (size_t)c->max_particles : 1;
  const size_t nblk = (size_t)c->max_blocks, G = (size_t)c->max_nodes;
  const size_t nwsum = (nblk + kBinScanWG - 1) / kBinScanWG;
  hipError_t e = hipSuccess;
  e = e == hipSuccess ? hipMalloc((void**)&c->tiles, nblk * kTileNodes * sizeof(float4)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->keep, G * sizeof(float4)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->v_out, G * sizeof(float4)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->gadj_in, G * sizeof(float4)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->Fe, np * 9 * sizeof(float)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->S, np * 9 * sizeof(float)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->payload, np * kAdjPayload * sizeof(float)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->xbar, np * 3 * sizeof(float)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->Febar, np * 9 * sizeof(float)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->bins, bin_record_ints((long long)nblk, (long long)np) * sizeof(int)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->count, nblk * 8 * sizeof(int)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->key_rank, np * sizeof(int2)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->blkoff, nblk * sizeof(int)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->wsum, nwsum * sizeof(int)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->bin_ctr, 4 * sizeof(int)) : e;
  e = e == hipSuccess ? hipMemset(c->bin_ctr, 0, 4 * sizeof(int)) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->work, 2 * sizeof(int)) : e;
  const size_t cand_bytes = (size_t)c->max_cand * (sizeof(DevModel) + sizeof(int2));  // table + material lists
  e = e == hipSuccess ? hipMalloc((void**)&c->cand_dev, cand_bytes) : e;
  e = e == hipSuccess ? hipHostMalloc((void**)&c->cand_pinned, cand_bytes) : e;
  e = e == hipSuccess ? hipEventCreateWithFlags(&c->cand_upload, hipEventDisableTiming) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->sched, kSchedCap * sizeof(ColTable)) : e;
  for (int i = 0; i < 2; ++i) {
    e = e == hipSuccess ? hipHostMalloc((void**)&c->sched_host[i], kSchedCap * sizeof(ColTable)) : e;
    e = e == hipSuccess ? hipEventCreateWithFlags(&c->sched_ev[i], hipEventDisableTiming) : e;
  }
  e = e == hipSuccess ? hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming) : e;
  e = e == hipSuccess ? hipMalloc((void**)&c->status, (size_t)c->max_cand * sizeof(int)) : e;
  e = e == hipSuccess ? hipHostMalloc((void**)&c->status_host, (size_t)c->max_cand * sizeof(int)) : e;
  e = e == hipSuccess ? hipMemset(c->status, 0, (size_t)c->max_cand * sizeof(int)) : e;
  e = e == hipSuccess ? hipMemset(c->count, 0, nblk * 8 * sizeof(int)) : e;
  e = e == hipSuccess ? hipDeviceSynchronize() : e;
  // the block-owned scatters are launched as wide as the device holds them
  // (per workgroup size of the tuning)
  e = e == hipSuccess ? hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, c->device) : e;
  const void* scatter_kernels[3][2] = {
      {reinterpret_cast<const void*>(k_p2g<128>), reinterpret_cast<const void*>(k_g2p_adj_scatter<128>)},
      {reinterpret_cast<const void*>(k_p2g<256>), reinterpret_cast<const void*>(k_g2p_adj_scatter<256>)},
      {reinterpret_cast<const void*>(k_p2g<512>), reinterpret_cast<const void*>(k_g2p_adj_scatter<512>)}};
  const int wgs[3] = {128, 256, 512};
  for (int v = 0; v < 3; ++v)
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      int per_cu = 0;
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, scatter_kernels[v][i], wgs[v], 0);
      if (e == hipSuccess) c->scatter_res[v][i] = per_cu;
    }
  if (e != hipSuccess)
    return set_err(e == hipErrorOutOfMemory ? EMPM_E_OOM : EMPM_E_HIP, "empm_create: %s", hipGetErrorString(e));
  return EMPM_OK;
}

// A call on a stream other than the previous call's is ordered after that
// call (the context's scratch and schedule buffers are shared by its calls).
int order_after_last(empm_ctx* ctx, hipStream_t st) {
  if (ctx->have_last_stream && ctx->last_stream != st) {
    HIP_TRY(hipEventRecord(ctx->order_ev, ctx->last_stream));
    HIP_TRY(hipStreamWaitEvent(st, ctx->order_ev, 0));
  }
  ctx->last_stream = st;
  ctx->have_last_stream = true;
  return EMPM_OK;
}

// The pinned staging slot for the next schedule upload (waits until the
// upload that last used it has been copied).
ColTable* sched_stage(empm_ctx* ctx) {
  (void)hipEventSynchronize(ctx->sched_ev[ctx->sched_slot]);
  return ctx->sched_host[ctx->sched_slot];
}

// Copy n staged entries to the device schedule, stream-ordered on st.
int sched_commit(empm_ctx* ctx, int n, hipStream_t st) {
  const int slot = ctx->sched_slot;
  HIP_TRY(hipMemcpyAsync(ctx->sched, ctx->sched_host[slot], (size_t)n * sizeof(ColTable), hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(ctx->sched_ev[slot], st));
  ctx->sched_slot = slot ^ 1;
  return EMPM_OK;
}

// What a captured launch sequence bakes in besides the per-substep
// pointers: the batch's launch shapes (read at run time from the device
// table are the simulations' scalars and arrays, so they are not part of it).
void key_batch(std::vector<unsigned long long>& k, const HostBatch& hb) {
  const Batch& b = hb.dev;
  k.push_back((unsigned long long)b.ncand);
  k.push_back((unsigned long long)b.N);
  k.push_back((unsigned long long)b.nblk);
  k.push_back((unsigned long long)b.nwaves);
  for (int m = 0; m <= 8; ++m) k.push_back((unsigned long long)b.moff[m]);
  unsigned long long bits = 0;
  for (int m = 0; m < 8; ++m) {
    k.push_back((unsigned long long)b.mwaves[m]);
    bits |= (hb.has_material[m] ? 1ull : 0ull) << m;
    bits |= (hb.mutates[m] ? 1ull : 0ull) << (8 + m);
  }
  k.push_back(bits);
}
void key_ptr(std::vector<unsigned long long>& k, const void* p) { k.push_back((unsigned long long)(uintptr_t)p); }
void key_state(std::vector<unsigned long long>& k, const empm_particles& s) {
  key_ptr(k, s.x), key_ptr(k, s.v), key_ptr(k, s.C), key_ptr(k, s.F), key_ptr(k, s.Fe), key_ptr(k, s.mstate);
}
void key_grid(std::vector<unsigned long long>& k, const empm_grid_rec* g) {
  if (!g) {
    k.push_back(0);
    return;
  }
  key_ptr(k, g->mv_m), key_ptr(k, g->v_out), key_ptr(k, g->bins);
}
unsigned long long fbits(double d) {
  unsigned long long u;
  memcpy(&u, &d, sizeof(u));
  return u;
}

// Run one call's launches: eagerly on st, or (graph mode) as an instantiated
// HIP graph cached under `key` and replayed on the context's own stream,
// ordered after st's earlier work and before its later work by events.  The
// graph's kernels read the schedule buffer and the simulation table at run
// time, so replays with new times, collider positions and parameters need no
// re-capture; what the launch sequence bakes in (pointers, counts, flags) is
// in the key.
template <typename Body>
int run_call(empm_ctx* ctx, hipStream_t st, std::vector<unsigned long long>& key, Body&& body) {
  if (!ctx->graphs || (ctx->prof && ctx->prof->mask)) return body(st);
  key.push_back(ctx->counts_pending ? 1 : 0);
  empm_ctx::GraphEntry* e = nullptr;
  for (auto& g : ctx->gcache)
    if (g.key == key) e = &g;
  if (e) {
    ctx->counts_pending = e->pending_after;
    ++ctx->graph_replays;
  } else {
    const bool pending_before = ctx->counts_pending;
    HIP_TRY(hipStreamBeginCapture(ctx->gstream, hipStreamCaptureModeThreadLocal));
    const int rc = body(ctx->gstream);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(ctx->gstream, &graph);
    if (rc || ce != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      ctx->counts_pending = pending_before;
      return rc ? rc : set_err(EMPM_E_HIP, "graph capture: %s", hipGetErrorString(ce));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      ctx->counts_pending = pending_before;
      return set_err(EMPM_E_HIP, "graph instantiate: %s", hipGetErrorString(ie));
    }
    if (ctx->gcache.size() >= kGraphCache) {
      size_t lru = 0;
      for (size_t i = 1; i < ctx->gcache.size(); ++i)
        if (ctx->gcache[i].used < ctx->gcache[lru].used) lru = i;
      HIP_TRY(hipStreamSynchronize(ctx->gstream));
      (void)hipGraphExecDestroy(ctx->gcache[lru].exec);
      ctx->gcache.erase(ctx->gcache.begin() + (long)lru);
    }
    empm_ctx::GraphEntry ne;
    ne.key = key;
    ne.exec = exec;
    ne.pending_after = ctx->counts_pending;
    ctx->gcache.push_back(std::move(ne));
    e = &ctx->gcache.back();
    ++ctx->graph_captures;
  }
  e->used = ++ctx->gclock;
  HIP_TRY(hipEventRecord(ctx->gev_in, st));
  HIP_TRY(hipStreamWaitEvent(ctx->gstream, ctx->gev_in, 0));
  HIP_TRY(hipGraphLaunch(e->exec, ctx->gstream));
  HIP_TRY(hipEventRecord(ctx->gev_out, ctx->gstream));
  HIP_TRY(hipStreamWaitEvent(st, ctx->gev_out, 0));
  return EMPM_OK;
}

// Forward rollout of a batch over n_steps substeps (see empm_rollout_fwd).
// With cols (and time) the colliders are advanced on the host per substep
// and the table / time each substep saw optionally recorded; with
// cols_hist / times_hist given instead (a replay) they are replayed.
int forward_calls(empm_ctx* ctx, const HostBatch& hb, empm_collider* cols, int ncol, double* time,
                  const empm_collider* hist_in, const double* times_in, double dt, int n_steps,
                  const empm_particles* states, const empm_grid_rec* grids, empm_collider* cols_hist,
                  double* times_hist, bool replay, hipStream_t st) {
  for (int k0 = 0; k0 < n_steps; k0 += kSchedCap) {
    const int n = std::min(kSchedCap, n_steps - k0);
    ColTable* stage = sched_stage(ctx);
    for (int j = 0; j < n; ++j) {
      const int k = k0 + j;
      int rc;
      if (replay) {
        if ((rc = make_table(&hist_in[(size_t)k * ncol], ncol, stage[j]))) return rc;
        stage[j].time = (float)times_in[k];
      } else {
        if ((rc = make_table(cols, ncol, stage[j]))) return rc;
        stage[j].time = (float)*time;
        if (cols_hist) memcpy(&cols_hist[(size_t)k * ncol], cols, sizeof(empm_collider) * ncol);
        if (times_hist) times_hist[k] = *time;
        advance_colliders(cols, ncol, *time, dt);
        *time = *time + dt;
      }
    }
    int rc = sched_commit(ctx, n, st);
    if (rc) return rc;
    std::vector<unsigned long long> key = {replay ? 2ull : 1ull, (unsigned long long)k0, (unsigned long long)n,
                                           (unsigned long long)n_steps, fbits(dt)};
    key_batch(key, hb);
    for (int k = k0; k <= k0 + n; ++k) key_state(key, states[k]);
    for (int k = k0; k < k0 + n; ++k) key_grid(key, grids ? &grids[k] : nullptr);
    rc = run_call(ctx, st, key, [&](hipStream_t cs) {
      for (int j = 0; j < n; ++j) {
        const int k = k0 + j;
        const int r = launch_fwd(ctx, hb, ctx->sched + j, (float)dt, &states[k], &states[k + 1],
                                 grids ? &grids[k] : nullptr, k == 0, k == n_steps - 1, cs, replay);
        if (r) return r;
      }
      return (int)EMPM_OK;
    });
    if (rc) return rc;
  }
  return EMPM_OK;
}

int rollout_fwd(empm_ctx* ctx, const empm_batch* batch, empm_collider* cols, int ncol, double* time, double dt,
                int n_steps, const empm_particles* states, const empm_grid_rec* grids, empm_collider* cols_hist,
                double* times_hist, hipStream_t st) {
  HostBatch hb;
  int rc = build_batch(ctx, batch, hb);
  if (rc) return rc;
  if (!time || !states || n_steps < 0) return set_err(EMPM_E_INVALID_ARG, "empm_rollout_fwd: bad arguments");
  for (int k = 0; k <= n_steps; ++k)
    if ((rc = check_particles(&states[k], hb.N, "states[k]"))) return rc;
  for (int k = 0; k < n_steps; ++k)
    if (hb.N > 0 && states[k].x == states[k + 1].x)
      return set_err(EMPM_E_INVALID_ARG, "states[%d] and states[%d] alias", k, k + 1);
  ColTable T;
  if ((rc = make_table(cols, ncol, T))) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = order_after_last(ctx, st))) return rc;
  if ((rc = upload_batch(ctx, hb, st))) return rc;
  return forward_calls(ctx, hb, cols, ncol, time, nullptr, nullptr, dt, n_steps, states, grids, cols_hist, times_hist,
                       false, st);
}

int rollout_fwd_replay(empm_ctx* ctx, const empm_batch* batch, const empm_collider* cols_hist, int ncol,
                       const double* times_hist, double dt, int n_steps, const empm_particles* states,
                       const empm_grid_rec* grids, hipStream_t st) {
  HostBatch hb;
  int rc = build_batch(ctx, batch, hb);
  if (rc) return rc;
  if (!states || !times_hist || n_steps < 0 || (ncol > 0 && !cols_hist))
    return set_err(EMPM_E_INVALID_ARG, "empm_rollout_fwd_replay: bad arguments");
  for (int k = 0; k <= n_steps; ++k)
    if ((rc = check_particles(&states[k], hb.N, "states[k]"))) return rc;
  for (int k = 0; k < n_steps; ++k)
    if (hb.N > 0 && states[k].x == states[k + 1].x)
      return set_err(EMPM_E_INVALID_ARG, "states[%d] and states[%d] alias", k, k + 1);
  if (n_steps > 0 && (rc = check_mstate(hb, states, 1, "empm_rollout_fwd_replay"))) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = order_after_last(ctx, st))) return rc;
  if ((rc = upload_batch(ctx, hb, st))) return rc;
  return forward_calls(ctx, hb, nullptr, ncol, nullptr, cols_hist, times_hist, dt, n_steps, states, grids, nullptr,
                       nullptr, true, st);
}

int rollout_adj(empm_ctx* ctx, const empm_batch* batch, const empm_collider* cols_hist, int ncol,
                const double* times_hist, double dt, int n_steps, const empm_particles* states,
                const empm_grid_rec* grids, const empm_particles* adj, float* d_mu, float* d_lam, float* d_ys,
                float* d_mass, float* d_vol, int flags, hipStream_t st) {
  HostBatch hb;
  int rc = build_batch(ctx, batch, hb);
  if (rc) return rc;
  if (!states || !adj || !times_hist || n_steps < 0 || (ncol > 0 && !cols_hist))
    return set_err(EMPM_E_INVALID_ARG, "empm_rollout_adj: bad arguments");
  for (int k = 0; k < n_steps; ++k)
    if ((rc = check_particles(&states[k], hb.N, "states[k]"))) return rc;
  if ((rc = check_mstate(hb, states, n_steps, "empm_rollout_adj"))) return rc;
  if ((rc = check_particles(&adj[0], hb.N, "adj[0]")) || (rc = check_particles(&adj[1], hb.N, "adj[1]"))) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = order_after_last(ctx, st))) return rc;
  if ((rc = upload_batch(ctx, hb, st))) return rc;
  // reverse order: sweep j = 0.. handles substep k = n_steps - 1 - j
  for (int j0 = 0; j0 < n_steps; j0 += kSchedCap) {
    const int n = std::min(kSchedCap, n_steps - j0);
    ColTable* stage = sched_stage(ctx);
    for (int i = 0; i < n; ++i) {
      const int k = n_steps - 1 - (j0 + i);
      if ((rc = make_table(&cols_hist[(size_t)k * ncol], ncol, stage[i]))) return rc;
      stage[i].time = (float)times_hist[k];
    }
    if ((rc = sched_commit(ctx, n, st))) return rc;
    std::vector<unsigned long long> key = {3ull, (unsigned long long)j0, (unsigned long long)n,
                                           (unsigned long long)n_steps, fbits(dt), (unsigned long long)flags};
    key_batch(key, hb);
    key_state(key, adj[0]);
    key_state(key, adj[1]);
    for (const float* q : {d_mu, d_lam, d_ys, d_mass, d_vol}) key_ptr(key, q);
    for (int i = 0; i < n; ++i) {
      const int k = n_steps - 1 - (j0 + i);
      key_state(key, states[k]);
      key_grid(key, grids ? &grids[k] : nullptr);
    }
    rc = run_call(ctx, st, key, [&](hipStream_t cs) {
      // within a schedule chunk, substep k's P2G adjoint also runs substep
      // k - 1's G2P adjoint when both are fully recorded
      bool done = false;
      for (int i = 0; i < n; ++i) {
        const int j = j0 + i;
        const int k = n_steps - 1 - j;
        const empm_particles* dout = &adj[j % 2];
        const empm_particles* din = &adj[(j + 1) % 2];
        const bool fuse = i + 1 < n && grids && fully_recorded(&states[k], &grids[k]) &&
                          fully_recorded(&states[k - 1], &grids[k - 1]);
        const int r = launch_adj(ctx, hb, ctx->sched + i, (float)dt, &states[k], grids ? &grids[k] : nullptr, dout,
                                 din, d_mu, d_lam, d_ys, d_mass, d_vol, flags, cs, fuse ? &states[k - 1] : nullptr,
                                 fuse ? &grids[k - 1] : nullptr, done);
        if (r) return r;
        done = fuse;
      }
      return (int)EMPM_OK;
    });
    if (rc) return rc;
  }
  return EMPM_OK;
}

}  // namespace

// Error reporting for the library's other translation units: stores msg as
// the thread's last error and returns code.
__attribute__((visibility("hidden"))) int empm_internal_error(int code, const char* msg) {
  return set_err(code, "%s", msg);
}

extern "C" {

int empm_abi_version(void) { return EMPM_ABI_VERSION; }

const char* empm_last_error(void) { return g_err; }

int empm_create(empm_ctx** out, const empm_create_desc* desc) {
  if (!out || !desc) return set_err(EMPM_E_INVALID_ARG, "null argument");
  *out = nullptr;
  if (desc->n_grid < 4 || desc->n_grid > 1024)
    return set_err(EMPM_E_INVALID_ARG, "n_grid=%d outside [4, 1024]", desc->n_grid);
  if (desc->max_particles < 0) return set_err(EMPM_E_INVALID_ARG, "max_particles < 0");
  HIP_TRY(hipSetDevice(desc->device));
  empm_ctx* c = new empm_ctx();
  c->device = desc->device;
  c->n_grid = desc->n_grid;
  c->max_cand = 1;
  c->max_particles = desc->max_particles;
  const long long nb = blocks_per_axis(desc->n_grid);
  c->max_nodes = (long long)desc->n_grid * desc->n_grid * desc->n_grid;
  c->max_blocks = nb * nb * nb;
  int rc = alloc_ctx(c);
  if (rc) {
    empm_destroy(c);
    return rc;
  }
  *out = c;
  return EMPM_OK;
}

int empm_create_batch(empm_ctx** out, const empm_batch_create_desc* desc) {
  if (!out || !desc) return set_err(EMPM_E_INVALID_ARG, "null argument");
  *out = nullptr;
  if (desc->max_grid < 4 || desc->max_grid > 1024)
    return set_err(EMPM_E_INVALID_ARG, "max_grid=%d outside [4, 1024]", desc->max_grid);
  if (desc->max_cand < 1 || desc->max_particles < 0)
    return set_err(EMPM_E_INVALID_ARG, "max_cand must be >= 1 and max_particles >= 0");
  const long long nb = blocks_per_axis(desc->max_grid);
  const long long nodes = (long long)desc->max_cand * desc->max_grid * desc->max_grid * desc->max_grid;
  const long long blocks = (long long)desc->max_cand * nb * nb * nb;
  if (blocks * 8 + 1 > INT32_MAX || nodes > ((long long)1 << 40))
    return set_err(EMPM_E_INVALID_ARG, "batch capacity too large (%lld blocks)", blocks);
  HIP_TRY(hipSetDevice(desc->device));
  empm_ctx* c = new empm_ctx();
  c->device = desc->device;
  c->n_grid = 0;
  c->max_cand = desc->max_cand;
  c->max_particles = desc->max_particles;
  c->max_nodes = nodes;
  c->max_blocks = blocks;
  int rc = alloc_ctx(c);
  if (rc) {
    empm_destroy(c);
    return rc;
  }
  *out = c;
  return EMPM_OK;
}

void empm_destroy(empm_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  if (c->prof) {
    for (int k = 0; k < EMPM_NUM_KERNELS; ++k)
      for (auto& pr : c->prof->pending[k]) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
      }
    for (hipEvent_t e : c->prof->pool) (void)hipEventDestroy(e);
    delete c->prof;
  }
  (void)hipFree(c->tiles);
  (void)hipFree(c->keep);
  (void)hipFree(c->v_out);
  (void)hipFree(c->gadj_in);
  (void)hipFree(c->Fe);
  (void)hipFree(c->S);
  (void)hipFree(c->payload);
  (void)hipFree(c->xbar);
  (void)hipFree(c->Febar);
  (void)hipFree(c->bins);
  (void)hipFree(c->count);
  (void)hipFree(c->key_rank);
  (void)hipFree(c->blkoff);
  (void)hipFree(c->wsum);
  (void)hipFree(c->bin_ctr);
  (void)hipFree(c->work);
  (void)hipFree(c->cand_dev);
  if (c->cand_pinned) (void)hipHostFree(c->cand_pinned);
  if (c->cand_upload) (void)hipEventDestroy(c->cand_upload);
  (void)hipFree(c->status);
  if (c->status_host) (void)hipHostFree(c->status_host);
  (void)hipFree(c->sched);
  for (int i = 0; i < 2; ++i) {
    if (c->sched_host[i]) (void)hipHostFree(c->sched_host[i]);
    if (c->sched_ev[i]) (void)hipEventDestroy(c->sched_ev[i]);
  }
  if (c->order_ev) (void)hipEventDestroy(c->order_ev);
  for (auto& g : c->gcache) (void)hipGraphExecDestroy(g.exec);
  if (c->gstream) (void)hipStreamDestroy(c->gstream);
  if (c->gev_in) (void)hipEventDestroy(c->gev_in);
  if (c->gev_out) (void)hipEventDestroy(c->gev_out);
  delete c;
}

int empm_set_graphs(empm_ctx* ctx, int enable) {
  if (!ctx) return set_err(EMPM_E_INVALID_ARG, "empm_set_graphs: null context");
  HIP_TRY(hipSetDevice(ctx->device));
  if (enable && !ctx->gstream) {
    HIP_TRY(hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&ctx->gev_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ctx->gev_out, hipEventDisableTiming));
  }
  if (!enable && !ctx->gcache.empty()) {
    HIP_TRY(hipStreamSynchronize(ctx->gstream));
    for (auto& g : ctx->gcache) (void)hipGraphExecDestroy(g.exec);
    ctx->gcache.clear();
  }
  ctx->graphs = enable != 0;
  return EMPM_OK;
}

int empm_set_tuning(empm_ctx* ctx, const empm_tuning* t) {
  if (!ctx || !t) return set_err(EMPM_E_INVALID_ARG, "empm_set_tuning: null argument");
  if (t->scatter_wg != 128 && t->scatter_wg != 256 && t->scatter_wg != 512)
    return set_err(EMPM_E_INVALID_ARG, "empm_set_tuning: scatter_wg=%d (128, 256 or 512)", t->scatter_wg);
  if (t->scatter_per_cu < 0 || t->scatter_per_cu > 32)
    return set_err(EMPM_E_INVALID_ARG, "empm_set_tuning: scatter_per_cu=%d outside [0, 32]", t->scatter_per_cu);
  if (t->heavy_chunks < 0) return set_err(EMPM_E_INVALID_ARG, "empm_set_tuning: heavy_chunks=%d < 0", t->heavy_chunks);
  if (t->dynamic_rounds < 1 || t->dynamic_rounds > 1024)
    return set_err(EMPM_E_INVALID_ARG, "empm_set_tuning: dynamic_rounds=%d outside [1, 1024]", t->dynamic_rounds);
  if (t->node_grid_cap < 1 || t->node_grid_cap > 65536)
    return set_err(EMPM_E_INVALID_ARG, "empm_set_tuning: node_grid_cap=%d outside [1, 65536]", t->node_grid_cap);
  if (!ctx->gcache.empty()) {
    // the cached graphs bake the old launch shapes
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->gstream));
    for (auto& g : ctx->gcache) (void)hipGraphExecDestroy(g.exec);
    ctx->gcache.clear();
  }
  ctx->tune = *t;
  return EMPM_OK;
}

int empm_get_tuning(empm_ctx* ctx, empm_tuning* t) {
  if (!ctx || !t) return set_err(EMPM_E_INVALID_ARG, "empm_get_tuning: null argument");
  *t = ctx->tune;
  return EMPM_OK;
}

int empm_set_deterministic(empm_ctx* ctx, int enable) {
  if (!ctx) return set_err(EMPM_E_INVALID_ARG, "empm_set_deterministic: null context");
  if ((enable != 0) != ctx->deterministic && !ctx->gcache.empty()) {
    // the cached graphs bake the other mode's kernels
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->gstream));
    for (auto& g : ctx->gcache) (void)hipGraphExecDestroy(g.exec);
    ctx->gcache.clear();
  }
  ctx->deterministic = enable != 0;
  return EMPM_OK;
}

int empm_graph_stats(empm_ctx* ctx, long long* captures, long long* replays) {
  if (!ctx) return set_err(EMPM_E_INVALID_ARG, "empm_graph_stats: null context");
  if (captures) *captures = ctx->graph_captures;
  if (replays) *replays = ctx->graph_replays;
  return EMPM_OK;
}

int empm_read_status(empm_ctx* ctx, int* status, int n, int clear, void* stream) {
  if (!ctx || n < 0 || (n > 0 && !status))
    return set_err(EMPM_E_INVALID_ARG, "empm_read_status: bad arguments");
  if (n > ctx->max_cand) n = ctx->max_cand;
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpyAsync(ctx->status_host, ctx->status, (size_t)ctx->max_cand * sizeof(int), hipMemcpyDeviceToHost,
                         st));
  if (clear)  // a fill kernel, not a memset node, so that callers may capture this call
    hipLaunchKernelGGL(k_fill_i32, dim3(nblocks((size_t)ctx->max_cand)), dim3(kBlock), 0, st, ctx->status,
                       (size_t)ctx->max_cand, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  int any = 0, first = -1;
  for (int c = 0; c < ctx->max_cand; ++c) {
    if (c < n) status[c] = ctx->status_host[c];
    any |= ctx->status_host[c];
    if (first < 0 && (ctx->status_host[c] & EMPM_STATUS_NONFINITE)) first = c;
  }
  if (any & EMPM_STATUS_NONFINITE)
    return set_err(EMPM_E_NONFINITE,
                   "non-finite particle state (x, v or F_trial) produced by simulation %d since the last status read",
                   first);
  return EMPM_OK;
}

int empm_sort_by_cell(empm_ctx* ctx, int n_particles, const float* x, float inv_dx, int* perm, void* stream) {
  if (!ctx || ctx->n_grid == 0 || n_particles < 0 || n_particles > ctx->max_particles ||
      (n_particles > 0 && (!x || !perm)) || !(inv_dx > 0.f))
    return set_err(EMPM_E_INVALID_ARG, "empm_sort_by_cell: bad arguments");
  if (n_particles == 0) return EMPM_OK;
  empm_model m;
  memset(&m, 0, sizeof(m));
  m.n_particles = n_particles;
  m.n_grid = ctx->n_grid;
  m.inv_dx = inv_dx;
  m.dx = 1.f / inv_dx;
  m.grid_lim = (float)ctx->n_grid / inv_dx;
  empm_batch b = batch_of(&m);
  HostBatch hb;
  hipStream_t st = (hipStream_t)stream;
  // positions only: the per-particle model arrays are not read by the binning
  hb.cand.clear();
  DevModel d{};
  d.N = n_particles;
  d.n = ctx->n_grid;
  d.nb = blocks_per_axis(ctx->n_grid);
  d.inv_dx = inv_dx;
  d.dx = m.dx;
  hb.cand.push_back(d);
  hb.N = n_particles;
  hb.nblk = d.nb * d.nb * d.nb;
  index_waves(hb);
  (void)b;
  HIP_TRY(hipSetDevice(ctx->device));
  int rc = upload_batch(ctx, hb, st);
  if (rc) return rc;
  const BinBuf bb = bin_buf(ctx->bins, hb.nblk, n_particles);
  if ((rc = do_bin(ctx, hb, x, bb, st, false, true))) return rc;
  HIP_TRY(hipMemcpyAsync(perm, bb.perm, (size_t)n_particles * sizeof(int), hipMemcpyDeviceToDevice, st));
  return EMPM_OK;
}

int empm_batch_sort_by_cell(empm_ctx* ctx, const empm_batch* batch, const float* x, int* perm, void* stream) {
  HostBatch hb;
  hipStream_t st = (hipStream_t)stream;
  int rc = prepare(ctx, batch, hb, st);
  if (rc) return rc;
  if (hb.N == 0) return EMPM_OK;
  if (!x || !perm) return set_err(EMPM_E_INVALID_ARG, "empm_batch_sort_by_cell: null array");
  const BinBuf bb = bin_buf(ctx->bins, hb.nblk, hb.N);
  if ((rc = do_bin(ctx, hb, x, bb, st, false, true))) return rc;
  HIP_TRY(hipMemcpyAsync(perm, bb.perm, (size_t)hb.N * sizeof(int), hipMemcpyDeviceToDevice, st));
  return EMPM_OK;
}

long long empm_bin_record_ints(int n_grid, int n_particles) {
  if (n_grid < 4 || n_grid > 1024 || n_particles < 0) return -1;
  const long long nb = blocks_per_axis(n_grid);
  return (long long)bin_record_ints(nb * nb * nb, n_particles);
}

int empm_batch_sizes(const empm_batch* batch, long long* n_nodes, long long* bin_record_ints_out) {
  if (!batch || !batch->models || batch->n_cand < 1) return set_err(EMPM_E_INVALID_ARG, "empm_batch_sizes: bad batch");
  long long G = 0, nblk = 0, N = 0;
  for (int c = 0; c < batch->n_cand; ++c) {
    const empm_model* m = &batch->models[c];
    int rc = check_scalars(m, c);
    if (rc) return rc;
    const long long nb = blocks_per_axis(m->n_grid);
    G += (long long)m->n_grid * m->n_grid * m->n_grid;
    nblk += nb * nb * nb;
    N += m->n_particles;
  }
  if (n_nodes) *n_nodes = G;
  if (bin_record_ints_out) *bin_record_ints_out = (long long)bin_record_ints(nblk, N);
  return EMPM_OK;
}

int empm_profile_enable(empm_ctx* ctx, unsigned kernel_mask) {
  if (!ctx) return set_err(EMPM_E_INVALID_ARG, "null context");
  if (!ctx->prof) ctx->prof = new KernelProfile();
  ctx->prof->mask = kernel_mask;
  return EMPM_OK;
}

int empm_profile_read(empm_ctx* ctx, int kernel_id, long* launches, double* total_ms) {
  if (!ctx || kernel_id < 0 || kernel_id >= EMPM_NUM_KERNELS || !launches || !total_ms)
    return set_err(EMPM_E_INVALID_ARG, "empm_profile_read: bad arguments");
  *launches = 0;
  *total_ms = 0.0;
  KernelProfile* p = ctx->prof;
  if (!p) return EMPM_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  for (auto& pr : p->pending[kernel_id]) {
    HIP_TRY(hipEventSynchronize(pr.second));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second));
    p->ms[kernel_id] += ms;
    p->launches[kernel_id] += 1;
    p->pool.push_back(pr.first);
    p->pool.push_back(pr.second);
  }
  p->pending[kernel_id].clear();
  *launches = p->launches[kernel_id];
  *total_ms = p->ms[kernel_id];
  return EMPM_OK;
}

int empm_profile_reset(empm_ctx* ctx) {
  if (!ctx) return set_err(EMPM_E_INVALID_ARG, "null context");
  if (!ctx->prof) return EMPM_OK;
  for (int k = 0; k < EMPM_NUM_KERNELS; ++k) {
    long l;
    double ms;
    int rc = empm_profile_read(ctx, k, &l, &ms);
    if (rc) return rc;
    ctx->prof->launches[k] = 0;
    ctx->prof->ms[k] = 0.0;
  }
  return EMPM_OK;
}

const char* empm_kernel_name(int kernel_id) {
  return (kernel_id >= 0 && kernel_id < EMPM_NUM_KERNELS) ? kKernelNames[kernel_id] : "";
}

int empm_mu_lam(int n, const float* E, const float* nu, float* mu, float* lam, void* stream) {
  if (n < 0 || (n > 0 && (!E || !nu || !mu || !lam))) return set_err(EMPM_E_INVALID_ARG, "empm_mu_lam: bad arguments");
  if (n == 0) return EMPM_OK;
  hipLaunchKernelGGL(k_mu_lam, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n, E, nu, mu, lam);
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_mu_lam_adj(int n, const float* E, const float* nu, const float* dmu, const float* dlam, float* dE, float* dnu,
                    void* stream) {
  if (n < 0 || (n > 0 && (!E || !nu))) return set_err(EMPM_E_INVALID_ARG, "empm_mu_lam_adj: bad arguments");
  if (n == 0) return EMPM_OK;
  hipLaunchKernelGGL(k_mu_lam_adj, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n, E, nu, dmu, dlam, dE, dnu);
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_mass_from_density(int n, const float* density, const float* vol, float* mass, void* stream) {
  if (n < 0 || (n > 0 && (!density || !vol || !mass)))
    return set_err(EMPM_E_INVALID_ARG, "empm_mass_from_density: bad arguments");
  if (n == 0) return EMPM_OK;
  hipLaunchKernelGGL(k_mass, dim3(nblocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n, density, vol, mass);
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

size_t empm_nn_sqdist_segments_workspace(int n_seg, int n_a, int max_a, int max_b) {
  if (n_seg <= 0 || n_a <= 0 || max_a <= 0 || max_b <= 0) return 0;
  const int s = nn_splits(max_a, max_b, n_seg);
  return s > 1 ? (size_t)s * (size_t)n_a * 8 : 0;
}

int empm_nn_sqdist_segments(int n_seg, const int* a_off, const int* b_off, int n_a, int max_a, int max_b,
                            const float* a, const float* b, float* dist, int* idx, void* workspace,
                            size_t workspace_bytes, void* stream) {
  if (n_seg < 0 || n_a < 0 || max_a < 0 || max_b < 0 || n_seg > 65535 ||
      (n_seg > 0 && n_a > 0 && (!a_off || !b_off || !a || !b || !dist || !idx)))
    return set_err(EMPM_E_INVALID_ARG, "empm_nn_sqdist_segments: bad arguments (n_seg=%d, n_a=%d)", n_seg, n_a);
  if (n_seg == 0 || n_a == 0 || max_a == 0) return EMPM_OK;
  hipStream_t st = (hipStream_t)stream;
  int s = max_b > 0 ? nn_splits(max_a, max_b, n_seg) : 1;
  if (s > 1 && (!workspace || workspace_bytes < (size_t)s * (size_t)n_a * 8 || ((uintptr_t)workspace & 3u))) s = 1;
  const dim3 grid(nn_blocks(max_a), s, n_seg);
  if (s == 1) {
    hipLaunchKernelGGL(k_nn_sqdist_seg, grid, dim3(kNnBlock), 0, st, a_off, b_off, a, b, 1, n_a, dist, idx);
  } else {
    float* pd = (float*)workspace;
    int* pi = (int*)((char*)workspace + (size_t)s * n_a * 4);
    hipLaunchKernelGGL(k_nn_sqdist_seg, grid, dim3(kNnBlock), 0, st, a_off, b_off, a, b, s, n_a, pd, pi);
    hipLaunchKernelGGL(k_nn_reduce, dim3(nblocks(n_a)), dim3(kBlock), 0, st, n_a, s, pd, pi, dist, idx);
  }
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

size_t empm_nn_sqdist_workspace(int n, int m) {
  if (n <= 0 || m <= 0) return 0;
  const int s = nn_splits(n, m);
  return s > 1 ? (size_t)s * (size_t)n * 8 : 0;
}

int empm_nn_sqdist(int n, const float* a, int m, const float* b, float* dist, int* idx, void* workspace,
                   size_t workspace_bytes, void* stream) {
  if (n < 0 || m < 0 || (n > 0 && (m < 1 || !a || !b || !dist || !idx)))
    return set_err(EMPM_E_INVALID_ARG, "empm_nn_sqdist: bad arguments (n=%d, m=%d)", n, m);
  if (n == 0) return EMPM_OK;
  hipStream_t st = (hipStream_t)stream;
  int s = nn_splits(n, m);
  if (s > 1 && (!workspace || workspace_bytes < (size_t)s * (size_t)n * 8 || ((uintptr_t)workspace & 3u))) s = 1;
  const int chunk = (m + s - 1) / s;
  if (s == 1) {
    hipLaunchKernelGGL(k_nn_sqdist, dim3(nn_blocks(n), 1), dim3(kNnBlock), 0, st, n, a, m, b, m, dist, idx);
  } else {
    float* pd = (float*)workspace;
    int* pi = (int*)((char*)workspace + (size_t)s * n * 4);
    hipLaunchKernelGGL(k_nn_sqdist, dim3(nn_blocks(n), s), dim3(kNnBlock), 0, st, n, a, m, b, chunk, pd, pi);
    hipLaunchKernelGGL(k_nn_reduce, dim3(nblocks(n)), dim3(kBlock), 0, st, n, s, pd, pi, dist, idx);
  }
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address; the
// array's name is a host-side stand-in that must not reach a kernel).
static int nn_scratch_ptr(float** out) {
  static std::mutex mu;
  static void* ptrs[256] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_err(EMPM_E_HIP, "empm_nn_sqdist_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= 256) return set_err(EMPM_E_INVALID_ARG, "empm_nn_sqdist_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(EMPM_E_HIP, "empm_nn_sqdist_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                     hipGetErrorString(e));
    ptrs[dev] = p;
  }
  *out = static_cast<float*>(ptrs[dev]);
  return EMPM_OK;
}

int empm_nn_sqdist_noalloc(int n, const float* a, int m, const float* b, float* dist, int* idx, void* stream) {
  if (n < 0 || m < 0 || (n > 0 && (m < 1 || !a || !b || !dist || !idx)))
    return set_err(EMPM_E_INVALID_ARG, "empm_nn_sqdist_noalloc: bad arguments (n=%d, m=%d)", n, m);
  if (n == 0) return EMPM_OK;
  hipStream_t st = (hipStream_t)stream;
  const int s = nn_splits(n, m);
  if (s == 1) {
    hipLaunchKernelGGL(k_nn_sqdist, dim3(nn_blocks(n), 1), dim3(kNnBlock), 0, st, n, a, m, b, m, dist, idx);
    HIP_TRY(hipGetLastError());
    return EMPM_OK;
  }
  if ((size_t)s * (size_t)n > kNnScratchPairs)
    return set_err(EMPM_E_INVALID_ARG,
                   "empm_nn_sqdist_noalloc: %d splits of %d queries need %zu bytes of scratch, more than its %zu "
                   "(use empm_nn_sqdist with a workspace)",
                   s, n, (size_t)s * (size_t)n * 8, kNnScratchPairs * 8);
  float* base = nullptr;
  int rc = nn_scratch_ptr(&base);
  if (rc) return rc;
  // every (split, query) entry is written by k_nn_sqdist before k_nn_reduce
  // reads it, so nothing needs clearing between calls
  float* pd = base;
  int* pi = reinterpret_cast<int*>(base + (size_t)s * n);
  const int chunk = (m + s - 1) / s;
  hipLaunchKernelGGL(k_nn_sqdist, dim3(nn_blocks(n), s), dim3(kNnBlock), 0, st, n, a, m, b, chunk, pd, pi);
  hipLaunchKernelGGL(k_nn_reduce, dim3(nblocks(n)), dim3(kBlock), 0, st, n, s, pd, pi, dist, idx);
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_substep_fwd(empm_ctx* ctx, const empm_model* model, const empm_collider* cols, int ncol, float time, float dt,
                     const empm_particles* in, const empm_particles* out, void* stream) {
  if (!model) return set_err(EMPM_E_INVALID_ARG, "null model");
  const empm_batch b = batch_of(model);
  HostBatch hb;
  hipStream_t st = (hipStream_t)stream;
  int rc = build_batch(ctx, &b, hb);
  if (rc) return rc;
  if ((rc = check_particles(in, hb.N, "in")) || (rc = check_particles(out, hb.N, "out"))) return rc;
  if (hb.N > 0 && (in->x == out->x || in->F == out->F)) return set_err(EMPM_E_INVALID_ARG, "in and out states alias");
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = order_after_last(ctx, st))) return rc;
  if ((rc = upload_batch(ctx, hb, st))) return rc;
  ColTable* T = sched_stage(ctx);
  if ((rc = make_table(cols, ncol, *T))) return rc;
  T->time = time;
  if ((rc = sched_commit(ctx, 1, st))) return rc;
  return launch_fwd(ctx, hb, ctx->sched, dt, in, out, nullptr, true, true, st);
}

int empm_substep_adj(empm_ctx* ctx, const empm_model* model, const empm_collider* cols, int ncol, float time, float dt,
                     const empm_particles* in, const empm_particles* d_out, const empm_particles* d_in, float* d_mu,
                     float* d_lam, float* d_ys, float* d_mass, float* d_vol, int flags, void* stream) {
  if (!model) return set_err(EMPM_E_INVALID_ARG, "null model");
  const empm_batch b = batch_of(model);
  HostBatch hb;
  hipStream_t st = (hipStream_t)stream;
  int rc = build_batch(ctx, &b, hb);
  if (rc) return rc;
  if ((rc = check_particles(in, hb.N, "in")) || (rc = check_particles(d_out, hb.N, "d_out")) ||
      (rc = check_particles(d_in, hb.N, "d_in")))
    return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = order_after_last(ctx, st))) return rc;
  if ((rc = upload_batch(ctx, hb, st))) return rc;
  ColTable* T = sched_stage(ctx);
  if ((rc = make_table(cols, ncol, *T))) return rc;
  T->time = time;
  if ((rc = sched_commit(ctx, 1, st))) return rc;
  return launch_adj(ctx, hb, ctx->sched, dt, in, nullptr, d_out, d_in, d_mu, d_lam, d_ys, d_mass, d_vol, flags, st);
}

int empm_restart_deformation(empm_ctx* ctx, const empm_model* model, const float* x, const float* target, float* F_trial,
                   void* stream) {
  if (!ctx || !model) return set_err(EMPM_E_INVALID_ARG, "empm_restart_deformation: null context or model");
  if (model->n_particles > 0 && (!x || !target || !F_trial))
    return set_err(EMPM_E_INVALID_ARG, "empm_restart_deformation: null x, target or F_trial");
  // the normalisation alone: no damping, no colliders, dt = 0 (no gravity)
  empm_model m = *model;
  m.grid_v_damping_scale = 1.1f;
  const empm_batch b = batch_of(&m);
  hipStream_t st = (hipStream_t)stream;
  HostBatch hb;
  int rc = build_batch(ctx, &b, hb);
  if (rc) return rc;
  if (hb.N == 0) return EMPM_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = order_after_last(ctx, st))) return rc;
  if ((rc = upload_batch(ctx, hb, st))) return rc;
  ColTable* stage = sched_stage(ctx);
  if ((rc = make_table(nullptr, 0, stage[0]))) return rc;
  stage[0].time = 0.f;
  if ((rc = sched_commit(ctx, 1, st))) return rc;
  const BinBuf bb = bin_buf(ctx->bins, hb.nblk, hb.N);
  if ((rc = do_bin(ctx, hb, x, bb, st))) return rc;
  const Binning B = binning(ctx, bb, hb.nblk);
  EMPM_LAUNCH_SCATTER(ctx, 0, k_p2g_disp, st, hb.dev, x, target, B, ctx->tiles, ctx->work);
  hipLaunchKernelGGL(k_grid, node_grid(ctx, hb.nblk), dim3(kBlock), 0, st, hb.dev, ctx->sched, 0.f, B, ctx->tiles,
                     nullptr, ctx->v_out);
  hipLaunchKernelGGL(k_restart_F, dim3(particle_grid(hb)), dim3(kBlock), 0, st, hb.dev, x, ctx->v_out, F_trial);
  HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_rollout_fwd(empm_ctx* ctx, const empm_model* model, empm_collider* cols, int ncol, double* time, double dt,
                     int n_steps, const empm_particles* states, const empm_grid_rec* grids, empm_collider* cols_hist,
                     double* times_hist, void* stream) {
  if (!model) return set_err(EMPM_E_INVALID_ARG, "null model");
  const empm_batch b = batch_of(model);
  return rollout_fwd(ctx, &b, cols, ncol, time, dt, n_steps, states, grids, cols_hist, times_hist,
                     (hipStream_t)stream);
}

int empm_rollout_fwd_replay(empm_ctx* ctx, const empm_model* model, const empm_collider* cols_hist, int ncol,
                            const double* times_hist, double dt, int n_steps, const empm_particles* states,
                            const empm_grid_rec* grids, void* stream) {
  if (!model) return set_err(EMPM_E_INVALID_ARG, "null model");
  const empm_batch b = batch_of(model);
  return rollout_fwd_replay(ctx, &b, cols_hist, ncol, times_hist, dt, n_steps, states, grids, (hipStream_t)stream);
}

int empm_rollout_adj(empm_ctx* ctx, const empm_model* model, const empm_collider* cols_hist, int ncol,
                     const double* times_hist, double dt, int n_steps, const empm_particles* states,
                     const empm_grid_rec* grids, const empm_particles* adj, float* d_mu, float* d_lam, float* d_ys,
                     float* d_mass, float* d_vol, int flags, void* stream) {
  if (!model) return set_err(EMPM_E_INVALID_ARG, "null model");
  const empm_batch b = batch_of(model);
  return rollout_adj(ctx, &b, cols_hist, ncol, times_hist, dt, n_steps, states, grids, adj, d_mu, d_lam, d_ys, d_mass,
                     d_vol, flags, (hipStream_t)stream);
}

int empm_batch_rollout_fwd(empm_ctx* ctx, const empm_batch* batch, empm_collider* cols, int ncol, double* time,
                           double dt, int n_steps, const empm_particles* states, const empm_grid_rec* grids,
                           empm_collider* cols_hist, double* times_hist, void* stream) {
  return rollout_fwd(ctx, batch, cols, ncol, time, dt, n_steps, states, grids, cols_hist, times_hist,
                     (hipStream_t)stream);
}

int empm_batch_rollout_fwd_replay(empm_ctx* ctx, const empm_batch* batch, const empm_collider* cols_hist, int ncol,
                                  const double* times_hist, double dt, int n_steps, const empm_particles* states,
                                  const empm_grid_rec* grids, void* stream) {
  return rollout_fwd_replay(ctx, batch, cols_hist, ncol, times_hist, dt, n_steps, states, grids, (hipStream_t)stream);
}

int empm_batch_rollout_adj(empm_ctx* ctx, const empm_batch* batch, const empm_collider* cols_hist, int ncol,
                           const double* times_hist, double dt, int n_steps, const empm_particles* states,
                           const empm_grid_rec* grids, const empm_particles* adj, float* d_mu, float* d_lam,
                           float* d_ys, float* d_mass, float* d_vol, int flags, void* stream) {
  return rollout_adj(ctx, batch, cols_hist, ncol, times_hist, dt, n_steps, states, grids, adj, d_mu, d_lam, d_ys,
                     d_mass, d_vol, flags, (hipStream_t)stream);
}

}  // extern "C"
