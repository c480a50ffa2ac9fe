// Linear-blend-skinning (LBS) motion interpolation of Gaussians from
// simulated particles, as hand-written HIP kernels for gfx950.  The
// particles are the "bones"; every Gaussian follows its K nearest bones.
//
// Reference: third_party/gaussian_splatting/dynamic_utils.py
//   get_topk_indices :248, knn_weights :268, calc_weights_vals_from_indices
//   :282, knn_weights_sparse :295, interpolate_motions_speedup :305-493, as
//   called once per frame by the interactive loop
//   (scripts/interactive_viser.py:259-280) and by the trainer's rendering
//   loop (phys_sim/engine/trainer_warp_mpm.py:1630-1656).
//
// Kernels:
//   k_knn<KMAX>    brute-force K nearest (K <= 32) with the reference points
//                  staged through LDS; a sorted list of K in registers per
//                  query (setup, run once per scene);
//   k_lbs_weights  inverse-distance weights of the K chosen bones;
//   k_lbs_bones    per bone: F = sum_j (new_j - new_b)(old_j - old_b)^T over
//                  its adjacent bones, the closest rotation R (Kabsch), the
//                  rotation's quaternion, written as a 64-byte record;
//   k_lbs_points   per Gaussian: gather of its K bone records, blended
//                  position and quaternion, product with its own quaternion.
// The per-frame path is k_lbs_bones (n_bones threads) + k_lbs_points (one
// thread per Gaussian, bound by gathering its 16 bone records).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "empm.h"
#include "empm_math.h"

// error reporting of the library (defined next to empm_last_error)
int empm_internal_error(int code, const char* msg);

using namespace empm;

namespace {

constexpr int kLbsBlock = 256;

// native 4-vectors for the non-temporal builtins (HIP's int4 / float4 are structs)
typedef int iv4 __attribute__((ext_vector_type(4)));
typedef float fv4 __attribute__((ext_vector_type(4)));
constexpr int kKnnTile = 2048;
constexpr int kMaxK = 32;

unsigned lbs_blocks(long long n) { return (unsigned)((n + kLbsBlock - 1) / kLbsBlock); }

int lbs_err(int code, const char* fmt, long long a = 0, long long b = 0) {
  char msg[256];
  snprintf(msg, sizeof(msg), fmt, a, b);
  return empm_internal_error(code, msg);
}

#define LBS_HIP_TRY(expr)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return empm_internal_error(EMPM_E_HIP, hipGetErrorString(e_));      \
  } while (0)

// ---------------------------------------------------------------------------
// K nearest neighbours.  One thread per query keeps the kk = k + skip nearest
// seen so far as a list sorted by squared distance in registers (KMAX slots,
// slots >= kk hold -inf so no insertion reaches them).  Candidates arrive in
// increasing index order and the comparisons are strict, so among equal
// distances the smaller index comes first.  The first `skip` entries are
// dropped from the output (get_topk_indices drops the query itself).
// dist[] receives Euclidean distances (the reference ranks torch.norm).
// ---------------------------------------------------------------------------
template <int KMAX>
__global__ __launch_bounds__(kLbsBlock) void k_knn(int n, const float* __restrict__ q, int m,
                                                   const float* __restrict__ r, int k, int skip,
                                                   int* __restrict__ idx, float* __restrict__ dist) {
  __shared__ float4 tile[kKnnTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = k + skip;
  float qx = 0.f, qy = 0.f, qz = 0.f;
  if (i < n) {
    qx = q[3 * (size_t)i];
    qy = q[3 * (size_t)i + 1];
    qz = q[3 * (size_t)i + 2];
  }
  float bd[KMAX];
  int bi[KMAX];
#pragma unroll
  for (int s = 0; s < KMAX; ++s) {
    bd[s] = s < kk ? __builtin_huge_valf() : -__builtin_huge_valf();
    bi[s] = -1;
  }
  float thr = __builtin_huge_valf();  // bd[kk - 1]
  for (int t0 = 0; t0 < m; t0 += kKnnTile) {
    const int cnt = m - t0 < kKnnTile ? m - t0 : kKnnTile;
    __syncthreads();
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
      const size_t o = 3 * (size_t)(t0 + j);
      tile[j] = make_float4(r[o], r[o + 1], r[o + 2], 0.f);
    }
    __syncthreads();
#pragma unroll 4
    for (int j = 0; j < cnt; ++j) {
      const float4 p = tile[j];
      const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
      const float d = dx * dx + dy * dy + dz * dz;
      if (d < thr) {
        float cd = d;
        int ci = t0 + j;
#pragma unroll
        for (int s = 0; s < KMAX; ++s) {
          const bool sw = cd < bd[s];
          const float td = bd[s];
          const int ti = bi[s];
          bd[s] = sw ? cd : td;
          bi[s] = sw ? ci : ti;
          cd = sw ? td : cd;
          ci = sw ? ti : ci;
        }
        float t = bd[0];
#pragma unroll
        for (int s = 1; s < KMAX; ++s) t = s < kk ? bd[s] : t;
        thr = t;
      }
    }
  }
  if (i < n) {
    const size_t o = (size_t)i * k;
#pragma unroll
    for (int s = 0; s < KMAX; ++s) {
      if (s >= skip && s < kk) {
        idx[o + s - skip] = bi[s];
        dist[o + s - skip] = sqrtf(bd[s]);
      }
    }
  }
}

// Index rows are int32 or int64 (torch.topk's dtype), read as given; KC = 16
// with int32 reads the 64-byte row as four 16-byte loads (plain loads: the
// four pieces of a row share a cache line, which a streaming hint would let
// go between them).  An index outside [0, nb) is
// replaced by 0 and reported (the caller then writes NaN for that row, as no
// memory outside the bones may be read).
template <int KC, typename I>
__device__ __forceinline__ bool load_row(const I* __restrict__ idx, int i, int k, int nb, int* out) {
  constexpr int KM = KC ? KC : kMaxK;
  bool ok = true;
  if constexpr (KC == 16 && sizeof(I) == 4) {
    const iv4* row = reinterpret_cast<const iv4*>(idx + (size_t)i * 16);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const iv4 v = row[c];
      out[4 * c] = v.x;
      out[4 * c + 1] = v.y;
      out[4 * c + 2] = v.z;
      out[4 * c + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      const bool in = (unsigned)out[j] < (unsigned)nb;
      ok = ok && in;
      out[j] = in ? out[j] : 0;
    }
  } else {
    const I* row = idx + (size_t)i * k;
#pragma unroll
    for (int j = 0; j < KM; ++j) {
      if (j < k) {
        const long long v = (long long)row[j];
        const bool in = v >= 0 && v < nb;
        ok = ok && in;
        out[j] = in ? (int)v : 0;
      }
    }
  }
  return ok;
}

// calc_weights_vals_from_indices (dynamic_utils.py:282-292):
// w_j = 1 / (|x - b_idx_j| + 1e-6), normalised to sum 1 over the row.  KC = 16
// writes the row as four 16-byte stores.
template <int KC, typename I>
__global__ __launch_bounds__(kLbsBlock) void k_lbs_weights(int n, const float* __restrict__ xyz, int k_rt,
                                                           const I* __restrict__ idx, int nb,
                                                           const float* __restrict__ bones,
                                                           float* __restrict__ w) {
  constexpr int KM = KC ? KC : kMaxK;
  const int k = KC ? KC : k_rt;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
  int nbr[KM];
  const bool ok = load_row<KC, I>(idx, i, k, nb, nbr);
  float v[KM];
  float sum = 0.f;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    if (j < k) {
      const size_t b = 3 * (size_t)nbr[j];
      const float dx = x - bones[b], dy = y - bones[b + 1], dz = z - bones[b + 2];
      v[j] = 1.f / (sqrtf(dx * dx + dy * dy + dz * dz) + 1e-6f);
      sum += v[j];
    }
  }
  const float inv = ok ? 1.f / sum : __builtin_nanf("");
  if constexpr (KC == 16) {
    fv4* row = reinterpret_cast<fv4*>(w + (size_t)i * 16);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      fv4 o;
      o.x = v[4 * c] * inv;
      o.y = v[4 * c + 1] * inv;
      o.z = v[4 * c + 2] * inv;
      o.w = v[4 * c + 3] * inv;
      row[c] = o;
    }
  } else {
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (j < k) w[(size_t)i * k + j] = v[j] * inv;
  }
}

// Quaternion (w, x, y, z) of a rotation matrix by Shepperd's method with the
// branch order and the eps = 1e-8 under the square root of kornia 0.8.1
// rotation_matrix_to_quaternion, which the reference calls
// (dynamic_utils.py:469-475): the branch is chosen by trace > 0, then by the
// largest diagonal entry (ties going to the later axis), and the component of
// that branch is the positive one.
__device__ __forceinline__ float4 rot_to_quat(const float* R) {
  const float m00 = R[0], m01 = R[1], m02 = R[2], m10 = R[3], m11 = R[4], m12 = R[5], m20 = R[6], m21 = R[7],
              m22 = R[8];
  const float tr = m00 + m11 + m22;
  float4 q;
  if (tr > 0.f) {
    const float sq = sqrtf(tr + 1.f + 1e-8f) * 2.f;
    q = make_float4(0.25f * sq, (m21 - m12) / sq, (m02 - m20) / sq, (m10 - m01) / sq);
  } else if (m00 > m11 && m00 > m22) {
    const float sq = sqrtf(1.f + m00 - m11 - m22 + 1e-8f) * 2.f;
    q = make_float4((m21 - m12) / sq, 0.25f * sq, (m01 + m10) / sq, (m02 + m20) / sq);
  } else if (m11 > m22) {
    const float sq = sqrtf(1.f + m11 - m00 - m22 + 1e-8f) * 2.f;
    q = make_float4((m02 - m20) / sq, (m01 + m10) / sq, 0.25f * sq, (m12 + m21) / sq);
  } else {
    const float sq = sqrtf(1.f + m22 - m00 - m11 + 1e-8f) * 2.f;
    q = make_float4((m10 - m01) / sq, (m02 + m20) / sq, (m12 + m21) / sq, 0.25f * sq);
  }
  return q;
}

// torch.nn.functional.normalize: q / max(|q|, 1e-12)
__device__ __forceinline__ float4 normalize4(float4 q) {
  const float nrm = fmaxf(sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w), 1e-12f);
  return make_float4(q.x / nrm, q.y / nrm, q.z / nrm, q.w / nrm);
}

// Per-bone transform (dynamic_utils.py:321-416).  With old offsets
// o_j = b_rel(j) - b and new offsets n_j = (b + m)_rel(j) - (b + m),
// F = sum_j n_j o_j^T and R is the rotation closest to F: U diag(1, 1, d) V^T
// of the SVD of F with d fixing det R = +1, which is what the reference's two
// determinant corrections produce.  svd3 returns U, V in SO(3) with the sign
// of det F on the last singular value, so that rotation is U V^T.
//
// The record of bone b is 64 bytes (4 float4, one 64-byte line segment for
// a gather):
//   [q.w q.x q.y q.z] [b.x b.y b.z 0] [m.x m.y m.z 0] [0 0 0 0]
// q = normalize(quaternion of R); the blend rotates by q, whose matrix is R
// to float32 round-off (the conversion's eps = 1e-8 is below it).
//
// Ranks (torch.linalg.matrix_rank with its default tolerance: singular
// values above 3 * eps_f32 * s_max count; dynamic_utils.py:331-338) decide
// which rotations the reference keeps (:342-405): one per rank-2/3 bone, or,
// when any bone has rank 1, one per rank-1 bone instead; that set is then
// assigned to all bones, which works only when it holds every bone or exactly
// one (a broadcast), else every rotation is the identity.  A rank-1 bone's
// rotation takes e_x to its F's range direction `axis` (sign: the largest-
// magnitude component positive; the reference's is torch.svd's) about
// axis x e_x (:373-397), and is the identity when that is shorter than 1e-6.
// The rank test uses sigma_0 sigma_1 ~ |cof F| and sigma_0^2 ~ |F|^2 in
// double precision.  The bones fill the rank state (8 device ints) that
// k_lbs_points reads:
//   [0] bones of rank < 2, [1] of rank 1, [2] of rank >= 2 (and bones with a
//   relation outside [0, n_bones), whose record is NaN), [3] the largest
//   rank-1 bone index, [4] the largest rank >= 2 bone index, [5] rank-1 bones
//   whose second singular value exceeds 1e-8 (the reference's allclose
//   assert fails then, :375, and every rotation is the identity).
enum { kRankLt2 = 0, kRank1 = 1, kRankFull = 2, kRank1Bone = 3, kRankFullBone = 4, kRank1Loose = 5, kRankState = 8 };

// The rank-1 rotation: x = e_x, perp = axis x x / |axis x x|,
// R = axis x^T + perp perp^T + (axis x perp)(x x perp)^T.
__device__ __forceinline__ void rank1_rotation(const float* F, float* R) {
  double cn[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    cn[c] = (double)F[c] * F[c] + (double)F[3 + c] * F[3 + c] + (double)F[6 + c] * F[6 + c];
  const int c = cn[0] >= cn[1] ? (cn[0] >= cn[2] ? 0 : 2) : (cn[1] >= cn[2] ? 1 : 2);
  const double inv = 1.0 / sqrt(cn[c]);
  double a[3] = {F[c] * inv, F[3 + c] * inv, F[6 + c] * inv};
  const int big = fabs(a[0]) >= fabs(a[1]) ? (fabs(a[0]) >= fabs(a[2]) ? 0 : 2) : (fabs(a[1]) >= fabs(a[2]) ? 1 : 2);
  if (a[big] < 0.0) a[0] = -a[0], a[1] = -a[1], a[2] = -a[2];
  // perp = axis x e_x = (0, a2, -a1)
  const double pl = sqrt(a[2] * a[2] + a[1] * a[1]);
  if (pl < 1e-6) {
    m_eye(R);
    return;
  }
  const double p[3] = {0.0, a[2] / pl, -a[1] / pl};
  const double th[3] = {0.0, -p[2], p[1]};  // e_x x perp
  const double ta[3] = {a[1] * p[2] - a[2] * p[1], a[2] * p[0] - a[0] * p[2], a[0] * p[1] - a[1] * p[0]};  // axis x perp
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) R[r * 3 + q] = (float)(a[r] * (q == 0 ? 1.0 : 0.0) + p[r] * p[q] + ta[r] * th[q]);
}
template <typename I>
__global__ __launch_bounds__(kLbsBlock) void k_lbs_bones(int nb, const float* __restrict__ bones,
                                                         const float* __restrict__ motions, int nadj,
                                                         const I* __restrict__ rel, float4* __restrict__ rec,
                                                         int* __restrict__ rank_state) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const float bx = bones[3 * (size_t)b], by = bones[3 * (size_t)b + 1], bz = bones[3 * (size_t)b + 2];
  const float mx = motions[3 * (size_t)b], my = motions[3 * (size_t)b + 1], mz = motions[3 * (size_t)b + 2];
  const float cx = bx + mx, cy = by + my, cz = bz + mz;
  float F[9];
  m_zero(F);
  const I* row = rel + (size_t)b * nadj;
  bool ok = true;
  for (int j = 0; j < nadj; ++j) {
    const long long v = (long long)row[j];
    if (v < 0 || v >= nb) {
      ok = false;
      continue;
    }
    const size_t a = 3 * (size_t)v;
    const float ax = bones[a], ay = bones[a + 1], az = bones[a + 2];
    const float o[3] = {ax - bx, ay - by, az - bz};
    const float nw[3] = {(ax + motions[a]) - cx, (ay + motions[a + 1]) - cy, (az + motions[a + 2]) - cz};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) F[r * 3 + c] += nw[r] * o[c];
  }
  // rank < 2 test in double
  double f2 = 0.0;
#pragma unroll
  for (int e = 0; e < 9; ++e) f2 += (double)F[e] * (double)F[e];
  double c2 = 0.0;
  {
    const double d0 = F[0], d1 = F[1], d2 = F[2], d3 = F[3], d4 = F[4], d5 = F[5], d6 = F[6], d7 = F[7], d8 = F[8];
    const double cf[9] = {d4 * d8 - d5 * d7, d5 * d6 - d3 * d8, d3 * d7 - d4 * d6, d2 * d7 - d1 * d8, d0 * d8 - d2 * d6,
                          d1 * d6 - d0 * d7, d1 * d5 - d2 * d4, d2 * d3 - d0 * d5, d0 * d4 - d1 * d3};
#pragma unroll
    for (int e = 0; e < 9; ++e) c2 += cf[e] * cf[e];
  }
  const double rtol = 3.0 * 1.1920928955078125e-07;
  const int rank = !ok ? 3 : (!(f2 > 0.0) ? 0 : (c2 > rtol * rtol * f2 * f2 ? 3 : 1));
  float R[9];
  if (rank < 2) atomicAdd(&rank_state[kRankLt2], 1);
  if (!ok) {
    // a relation index outside [0, n_bones): the record is NaN
    atomicAdd(&rank_state[kRankFull], 1);
    atomicMax(&rank_state[kRankFullBone], b);
#pragma unroll
    for (int e = 0; e < 9; ++e) R[e] = __builtin_nanf("");
  } else if (rank == 0) {
    m_eye(R);
  } else if (rank == 1) {
    atomicAdd(&rank_state[kRank1], 1);
    atomicMax(&rank_state[kRank1Bone], b);
    if (c2 > 1e-16 * f2) atomicAdd(&rank_state[kRank1Loose], 1);  // sigma_1 ~ |cof F| / |F| > 1e-8
    rank1_rotation(F, R);
  } else {
    atomicAdd(&rank_state[kRankFull], 1);
    atomicMax(&rank_state[kRankFullBone], b);
    // scale F to unit max entry (R is invariant) so F^T F stays in range
    float mx_abs = 0.f;
#pragma unroll
    for (int e = 0; e < 9; ++e) mx_abs = fmaxf(mx_abs, fabsf(F[e]));
    const float is = 1.f / mx_abs;
#pragma unroll
    for (int e = 0; e < 9; ++e) F[e] *= is;
    float U[9], s[3], V[9];
    svd3(F, U, s, V);
    m_mul_bt(U, V, R);
  }
  const float4 qn = normalize4(rot_to_quat(R));
  float4* o = rec + 4 * (size_t)b;
  o[0] = qn;
  o[1] = make_float4(bx, by, bz, 0.f);
  o[2] = make_float4(mx, my, mz, 0.f);
  o[3] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Hamilton product q1 * q2, (w, x, y, z) (dynamic_utils.py:452-460)
__device__ __forceinline__ float4 quat_mul(float4 a, float4 b) {
  return make_float4(a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w, a.x * b.y + a.y * b.x + a.z * b.w - a.w * b.z,
                     a.x * b.z - a.y * b.w + a.z * b.x + a.w * b.y, a.x * b.w + a.y * b.z - a.z * b.y + a.w * b.x);
}

// v rotated by the unit quaternion q = (w, u): v + 2w (u x v) + 2 u x (u x v)
__device__ __forceinline__ void quat_rotate(float4 q, float vx, float vy, float vz, float& ox, float& oy, float& oz) {
  const float tx = 2.f * (q.z * vz - q.w * vy), ty = 2.f * (q.w * vx - q.y * vz), tz = 2.f * (q.y * vy - q.z * vx);
  ox = vx + q.x * tx + (q.z * tz - q.w * ty);
  oy = vy + q.x * ty + (q.w * tx - q.y * tz);
  oz = vz + q.x * tz + (q.y * ty - q.z * tx);
}

// Per Gaussian (dynamic_utils.py:434-485):
//   x' = sum_j w_j (R_j (x - b_j) + m_j + b_j)
//   q' = normalize(sum_j w_j q_j) * quat   (when quat is given)
// with w either given (n x k) or, when null, the inverse-distance weights of
// calc_weights_vals_from_indices computed from the same bones and x.  The
// rotations follow the rank state of k_lbs_bones: every bone's own, one
// bone's for all (a broadcast), or the identity for all (q_j = (1, 0, 0, 0)).
// One thread per Gaussian; KC = 16 is the reference's K with its index row
// read as four 16-byte loads, KC = 0 any K <= 32; W: weights given, Q:
// quaternions given.  Every branch inside the neighbour loop is resolved at
// compile time (the identity fallback is a select of q = (1, 0, 0, 0), whose
// rotation is exact), so the gathers of all neighbours are in flight
// together.  The quaternion stream and the outputs, touched once per
// frame, are read and written non-temporally so that the bone records stay
// in L2 for the gathers; the 64-byte index and weight rows use plain loads
// (their four 16-byte pieces share a cache line).
template <int KC, typename I, bool W, bool Q>
__global__ __launch_bounds__(kLbsBlock) void k_lbs_points(int n, const float* __restrict__ xyz,
                                                          const float* __restrict__ quat, int k_rt,
                                                          const I* __restrict__ idx,
                                                          const float* __restrict__ weights, int nbones,
                                                          const float4* __restrict__ rec,
                                                          const int* __restrict__ rank_state,
                                                          float* __restrict__ xyz_out,
                                                          float* __restrict__ quat_out) {
  constexpr int KM = KC ? KC : kMaxK;
  const int k = KC ? KC : k_rt;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // which set of rotations the reference keeps and whether it is
  // assignable to all bones (dynamic_utils.py:365-405)
  const int n1 = rank_state[kRank1], nfull = rank_state[kRankFull];
  const int kept = n1 > 0 ? n1 : nfull;
  const bool ident = (n1 > 0 && rank_state[kRank1Loose] > 0) || !(kept == nbones || kept == 1);
  const int bsel = kept == 1 ? rank_state[n1 > 0 ? kRank1Bone : kRankFullBone] : -1;
  const float4 qsel = (!ident && bsel >= 0) ? rec[4 * (size_t)bsel] : make_float4(1.f, 0.f, 0.f, 0.f);
  const float x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
  int nb[KM];
  const bool ok = load_row<KC, I>(idx, i, k, nbones, nb);
  float wv[KM];
  if constexpr (W) {
    if constexpr (KC == 16) {
      const fv4* row = reinterpret_cast<const fv4*>(weights + (size_t)i * 16);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const fv4 v = row[c];
        wv[4 * c] = v.x;
        wv[4 * c + 1] = v.y;
        wv[4 * c + 2] = v.z;
        wv[4 * c + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int j = 0; j < KM; ++j)
        if (j < k) wv[j] = weights[(size_t)i * k + j];
    }
  }
  // One pass over the neighbours.  Without given weights the blend uses the
  // unnormalised v_j = 1 / (|x - b_j| + 1e-6) and divides by their sum at
  // the end (the quaternion is normalised anyway), which is the reference's
  // w_j = v_j / sum_j v_j up to rounding.
  float px = 0.f, py = 0.f, pz = 0.f, vsum = 0.f;
  float4 qa = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    if (j < k) {
      const float4* r = rec + 4 * (size_t)nb[j];
      const float4 b = r[1], m = r[2];
      const float4 rq = r[0];
      const float4 bq = ident ? make_float4(1.f, 0.f, 0.f, 0.f) : (bsel >= 0 ? qsel : rq);
      const float lx = x - b.x, ly = y - b.y, lz = z - b.z;
      float wj;
      if constexpr (W) {
        wj = wv[j];
      } else {
        wj = 1.f / (sqrtf(lx * lx + ly * ly + lz * lz) + 1e-6f);
        vsum += wj;
      }
      float ox, oy, oz;
      quat_rotate(bq, lx, ly, lz, ox, oy, oz);
      px += ((ox + m.x) + b.x) * wj;
      py += ((oy + m.y) + b.y) * wj;
      pz += ((oz + m.z) + b.z) * wj;
      if constexpr (Q) {
        qa.x += bq.x * wj;
        qa.y += bq.y * wj;
        qa.z += bq.z * wj;
        qa.w += bq.w * wj;
      }
    }
  }
  if constexpr (!W) {
    const float inv = 1.f / vsum;
    px *= inv;
    py *= inv;
    pz *= inv;
  }
  if (!ok) px = py = pz = qa.x = __builtin_nanf("");
  __builtin_nontemporal_store(px, xyz_out + 3 * (size_t)i);
  __builtin_nontemporal_store(py, xyz_out + 3 * (size_t)i + 1);
  __builtin_nontemporal_store(pz, xyz_out + 3 * (size_t)i + 2);
  if constexpr (Q) {
    const fv4 gv = __builtin_nontemporal_load(reinterpret_cast<const fv4*>(quat) + i);
    const float4 r = quat_mul(normalize4(qa), make_float4(gv.x, gv.y, gv.z, gv.w));
    fv4 rv;
    rv.x = r.x;
    rv.y = r.y;
    rv.z = r.z;
    rv.w = r.w;
    __builtin_nontemporal_store(rv, reinterpret_cast<fv4*>(quat_out) + i);
  }
}

// the rank state before k_lbs_bones: counts 0, bone indices -1
__global__ void k_init_rank_state(int* p) {
  const int t = threadIdx.x;
  p[t] = (t == kRank1Bone || t == kRankFullBone) ? -1 : 0;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int empm_knn(int n, const float* q, int m, const float* r, int k, int skip_first, int* idx, float* dist,
             void* stream) {
  if (n < 0 || m < 0 || k < 1 || (skip_first != 0 && skip_first != 1) || k + skip_first > kMaxK)
    return lbs_err(EMPM_E_INVALID_ARG, "empm_knn: need n, m >= 0 and 1 <= k + skip_first <= 32 (k=%lld)", k);
  if (n == 0) return EMPM_OK;
  if (m < k + skip_first)
    return lbs_err(EMPM_E_INVALID_ARG, "empm_knn: %lld reference points for %lld neighbours", m, k + skip_first);
  if (!q || !r || !idx || !dist) return lbs_err(EMPM_E_INVALID_ARG, "empm_knn: null array");
  hipStream_t st = (hipStream_t)stream;
  if (k + skip_first <= 16)
    hipLaunchKernelGGL(k_knn<16>, dim3(lbs_blocks(n)), dim3(kLbsBlock), 0, st, n, q, m, r, k, skip_first, idx, dist);
  else
    hipLaunchKernelGGL(k_knn<32>, dim3(lbs_blocks(n)), dim3(kLbsBlock), 0, st, n, q, m, r, k, skip_first, idx, dist);
  LBS_HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_lbs_weights(int n, const float* xyz, int k, const void* idx, int idx_bytes, int n_bones, const float* bones,
                     float* weights, void* stream) {
  if (n < 0 || k < 1 || k > kMaxK || (idx_bytes != 4 && idx_bytes != 8))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_weights: need 1 <= k <= 32 and 4- or 8-byte indices (k=%lld)", k);
  if (n == 0) return EMPM_OK;
  if (n_bones < 1) return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_weights: n_bones=%lld", n_bones);
  if (!xyz || !idx || !bones || !weights) return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_weights: null array");
  hipStream_t st = (hipStream_t)stream;
  const dim3 g(lbs_blocks(n)), blk(kLbsBlock);
  const bool v16 = k == 16 && (idx_bytes == 8 || aligned16(idx)) && aligned16(weights);
  if (idx_bytes == 4 && v16)
    hipLaunchKernelGGL((k_lbs_weights<16, int32_t>), g, blk, 0, st, n, xyz, k, (const int32_t*)idx, n_bones, bones,
                       weights);
  else if (idx_bytes == 4)
    hipLaunchKernelGGL((k_lbs_weights<0, int32_t>), g, blk, 0, st, n, xyz, k, (const int32_t*)idx, n_bones, bones,
                       weights);
  else if (v16)
    hipLaunchKernelGGL((k_lbs_weights<16, int64_t>), g, blk, 0, st, n, xyz, k, (const int64_t*)idx, n_bones, bones,
                       weights);
  else
    hipLaunchKernelGGL((k_lbs_weights<0, int64_t>), g, blk, 0, st, n, xyz, k, (const int64_t*)idx, n_bones, bones,
                       weights);
  LBS_HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_lbs_bones(int n_bones, const float* bones, const float* motions, int n_adj, const void* relations,
                   int idx_bytes, float* bone_rec, int* rank_state, void* stream) {
  if (n_bones < 0 || n_adj < 0 || (idx_bytes != 4 && idx_bytes != 8))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: bad arguments (n_bones=%lld, n_adj=%lld)", n_bones, n_adj);
  if (!rank_state) return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: null rank state");
  if (n_bones > 0 && (!bones || !motions || !bone_rec || (n_adj > 0 && !relations)))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: null array");
  if (n_bones > 0 && ((uintptr_t)bone_rec & 63u))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: bone_rec not 64-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_init_rank_state, dim3(1), dim3(kRankState), 0, st, rank_state);
  if (n_bones > 0) {
    if (idx_bytes == 4)
      hipLaunchKernelGGL(k_lbs_bones<int32_t>, dim3(lbs_blocks(n_bones)), dim3(kLbsBlock), 0, st, n_bones, bones,
                         motions, n_adj, (const int32_t*)relations, (float4*)bone_rec, rank_state);
    else
      hipLaunchKernelGGL(k_lbs_bones<int64_t>, dim3(lbs_blocks(n_bones)), dim3(kLbsBlock), 0, st, n_bones, bones,
                         motions, n_adj, (const int64_t*)relations, (float4*)bone_rec, rank_state);
  }
  LBS_HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_lbs_points(int n, const float* xyz, const float* quat, int k, const void* idx, int idx_bytes,
                    const float* weights, int n_bones, const float* bone_rec, const int* rank_state, float* xyz_out,
                    float* quat_out, void* stream) {
  if (n < 0 || k < 1 || k > kMaxK || (idx_bytes != 4 && idx_bytes != 8))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: need 1 <= k <= 32 and 4- or 8-byte indices (k=%lld)", k);
  if (n == 0) return EMPM_OK;
  if (n_bones < 1) return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: n_bones=%lld", n_bones);
  if (!xyz || !idx || !bone_rec || !rank_state || !xyz_out || (quat && !quat_out))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: null array");
  if (xyz_out == xyz || (quat && quat_out == quat))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: outputs alias inputs");
  if (((uintptr_t)bone_rec & 63u) || (quat && (!aligned16(quat) || !aligned16(quat_out))))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: bone_rec must be 64-byte, quaternions 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  const dim3 g(lbs_blocks(n)), blk(kLbsBlock);
  const float4* r4 = (const float4*)bone_rec;
  const bool v16 = k == 16 && (idx_bytes == 8 || aligned16(idx)) && (!weights || aligned16(weights));
  const int sel = (v16 ? 8 : 0) | (idx_bytes == 8 ? 4 : 0) | (weights ? 2 : 0) | (quat ? 1 : 0);
#define EMPM_LBS_POINTS(KC, I, W, Q)                                                                          \
  hipLaunchKernelGGL((k_lbs_points<KC, I, W, Q>), g, blk, 0, st, n, xyz, quat, k, (const I*)idx, weights, n_bones, \
                     r4, rank_state, xyz_out, quat_out)
  switch (sel) {
    case 0: EMPM_LBS_POINTS(0, int32_t, false, false); break;
    case 1: EMPM_LBS_POINTS(0, int32_t, false, true); break;
    case 2: EMPM_LBS_POINTS(0, int32_t, true, false); break;
    case 3: EMPM_LBS_POINTS(0, int32_t, true, true); break;
    case 4: EMPM_LBS_POINTS(0, int64_t, false, false); break;
    case 5: EMPM_LBS_POINTS(0, int64_t, false, true); break;
    case 6: EMPM_LBS_POINTS(0, int64_t, true, false); break;
    case 7: EMPM_LBS_POINTS(0, int64_t, true, true); break;
    case 8: EMPM_LBS_POINTS(16, int32_t, false, false); break;
    case 9: EMPM_LBS_POINTS(16, int32_t, false, true); break;
    case 10: EMPM_LBS_POINTS(16, int32_t, true, false); break;
    case 11: EMPM_LBS_POINTS(16, int32_t, true, true); break;
    case 12: EMPM_LBS_POINTS(16, int64_t, false, false); break;
    case 13: EMPM_LBS_POINTS(16, int64_t, false, true); break;
    case 14: EMPM_LBS_POINTS(16, int64_t, true, false); break;
    default: EMPM_LBS_POINTS(16, int64_t, true, true); break;
  }
#undef EMPM_LBS_POINTS
  LBS_HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

}  // extern "C"
