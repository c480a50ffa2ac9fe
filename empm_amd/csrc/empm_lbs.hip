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
//                  rotation's quaternion, written as an 80-byte record;
//   k_lbs_points   per Gaussian: gather of its K bone records, blended
//                  position and quaternion, product with its own quaternion.
// The per-frame path is k_lbs_bones (n_bones threads) + k_lbs_points (one
// thread per Gaussian, HBM-bound on its index row and state).
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

// Index rows are int32 or int64 (torch.topk's dtype), read as given.  An
// index outside [0, nb) is replaced by 0 and reported (the caller then
// writes NaN for that row, as no memory outside the bones may be read).
template <typename I>
__device__ __forceinline__ bool load_row(const I* __restrict__ row, int k, int nb, int* out) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) {
    if (j < k) {
      const long long v = (long long)row[j];
      const bool in = v >= 0 && v < nb;
      ok = ok && in;
      out[j] = in ? (int)v : 0;
    }
  }
  return ok;
}

// calc_weights_vals_from_indices (dynamic_utils.py:282-292):
// w_j = 1 / (|x - b_idx_j| + 1e-6), normalised to sum 1 over the row.
template <typename I>
__global__ __launch_bounds__(kLbsBlock) void k_lbs_weights(int n, const float* __restrict__ xyz, int k,
                                                           const I* __restrict__ idx, int nb,
                                                           const float* __restrict__ bones,
                                                           float* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
  int nbr[kMaxK];
  const bool ok = load_row(idx + (size_t)i * k, k, nb, nbr);
  float v[kMaxK];
  float sum = 0.f;
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) {
    if (j < k) {
      const size_t b = 3 * (size_t)nbr[j];
      const float dx = x - bones[b], dy = y - bones[b + 1], dz = z - bones[b + 2];
      v[j] = 1.f / (sqrtf(dx * dx + dy * dy + dz * dz) + 1e-6f);
      sum += v[j];
    }
  }
  const float inv = ok ? 1.f / sum : __builtin_nanf("");
#pragma unroll
  for (int j = 0; j < kMaxK; ++j)
    if (j < k) w[(size_t)i * k + j] = v[j] * inv;
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
// The record of bone b (5 float4):
//   [R00 R01 R02 b.x] [R10 R11 R12 b.y] [R20 R21 R22 b.z] [m.x m.y m.z 0] [q]
// q = normalize(quaternion of R).
//
// A bone whose F has rank < 2 (torch.linalg.matrix_rank with its default
// tolerance: singular values above 3 * eps_f32 * s_max count) sends the
// reference into a fallback that leaves every bone's rotation at the
// identity; such bones are counted in *degenerate and k_lbs_points then uses
// R = I for all bones.  The rank test uses sigma_1 sigma_2 ~ |cof F| and
// sigma_1^2 ~ |F|^2, in double precision.
template <typename I>
__global__ __launch_bounds__(kLbsBlock) void k_lbs_bones(int nb, const float* __restrict__ bones,
                                                         const float* __restrict__ motions, int nadj,
                                                         const I* __restrict__ rel, float4* __restrict__ rec,
                                                         int* __restrict__ degenerate) {
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
  const bool degen = ok && !(c2 > rtol * rtol * f2 * f2);
  float R[9];
  if (!ok) {
    // a relation index outside [0, n_bones): the record is NaN
#pragma unroll
    for (int e = 0; e < 9; ++e) R[e] = __builtin_nanf("");
  } else if (degen) {
    atomicAdd(degenerate, 1);
    m_eye(R);
  } else {
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
  float4* o = rec + 5 * (size_t)b;
  o[0] = make_float4(R[0], R[1], R[2], bx);
  o[1] = make_float4(R[3], R[4], R[5], by);
  o[2] = make_float4(R[6], R[7], R[8], bz);
  o[3] = make_float4(mx, my, mz, 0.f);
  o[4] = qn;
}

// Hamilton product q1 * q2, (w, x, y, z) (dynamic_utils.py:452-460)
__device__ __forceinline__ float4 quat_mul(float4 a, float4 b) {
  return make_float4(a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w, a.x * b.y + a.y * b.x + a.z * b.w - a.w * b.z,
                     a.x * b.z - a.y * b.w + a.z * b.x + a.w * b.y, a.x * b.w + a.y * b.z - a.z * b.y + a.w * b.x);
}

// Per Gaussian (dynamic_utils.py:434-485):
//   x' = sum_j w_j (R_j (x - b_j) + m_j + b_j)
//   q' = normalize(sum_j w_j q_j) * quat   (when quat is given)
// with w either given (n x k) or, when null, the inverse-distance weights of
// calc_weights_vals_from_indices computed from the same bones and x.  With
// *degenerate != 0 every R_j is the identity and every q_j = (1, 0, 0, 0).
template <typename I>
__global__ __launch_bounds__(kLbsBlock) void k_lbs_points(int n, const float* __restrict__ xyz,
                                                          const float* __restrict__ quat, int k,
                                                          const I* __restrict__ idx,
                                                          const float* __restrict__ weights, int nbones,
                                                          const float4* __restrict__ rec,
                                                          const int* __restrict__ degenerate,
                                                          float* __restrict__ xyz_out,
                                                          float* __restrict__ quat_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool ident = *degenerate != 0;
  const float x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
  int nb[kMaxK];
  const bool ok = load_row(idx + (size_t)i * k, k, nbones, nb);
  float w[kMaxK];
  if (weights) {
#pragma unroll
    for (int j = 0; j < kMaxK; ++j)
      if (j < k) w[j] = weights[(size_t)i * k + j];
  } else {
    float sum = 0.f;
#pragma unroll
    for (int j = 0; j < kMaxK; ++j) {
      if (j < k) {
        const float4* r = rec + 5 * (size_t)nb[j];
        const float dx = x - r[0].w, dy = y - r[1].w, dz = z - r[2].w;
        w[j] = 1.f / (sqrtf(dx * dx + dy * dy + dz * dz) + 1e-6f);
        sum += w[j];
      }
    }
    const float inv = 1.f / sum;
#pragma unroll
    for (int j = 0; j < kMaxK; ++j)
      if (j < k) w[j] *= inv;
  }
  float px = 0.f, py = 0.f, pz = 0.f;
  float4 qa = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) {
    if (j < k) {
      const float4* r = rec + 5 * (size_t)nb[j];
      const float4 r0 = r[0], r1 = r[1], r2 = r[2], t = r[3];
      const float lx = x - r0.w, ly = y - r1.w, lz = z - r2.w;
      float ox, oy, oz;
      if (ident) {
        ox = lx;
        oy = ly;
        oz = lz;
      } else {
        ox = r0.x * lx + r0.y * ly + r0.z * lz;
        oy = r1.x * lx + r1.y * ly + r1.z * lz;
        oz = r2.x * lx + r2.y * ly + r2.z * lz;
      }
      px += ((ox + t.x) + r0.w) * w[j];
      py += ((oy + t.y) + r1.w) * w[j];
      pz += ((oz + t.z) + r2.w) * w[j];
      if (quat) {
        const float4 bq = ident ? make_float4(1.f, 0.f, 0.f, 0.f) : r[4];
        qa.x += bq.x * w[j];
        qa.y += bq.y * w[j];
        qa.z += bq.z * w[j];
        qa.w += bq.w * w[j];
      }
    }
  }
  if (!ok) px = py = pz = qa.x = __builtin_nanf("");
  xyz_out[3 * (size_t)i] = px;
  xyz_out[3 * (size_t)i + 1] = py;
  xyz_out[3 * (size_t)i + 2] = pz;
  if (quat) {
    const float4 g = reinterpret_cast<const float4*>(quat)[i];
    reinterpret_cast<float4*>(quat_out)[i] = quat_mul(normalize4(qa), g);
  }
}

__global__ void k_zero_i32(int* p) { *p = 0; }

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
  if (idx_bytes == 4)
    hipLaunchKernelGGL(k_lbs_weights<int32_t>, dim3(lbs_blocks(n)), dim3(kLbsBlock), 0, st, n, xyz, k,
                       (const int32_t*)idx, n_bones, bones, weights);
  else
    hipLaunchKernelGGL(k_lbs_weights<int64_t>, dim3(lbs_blocks(n)), dim3(kLbsBlock), 0, st, n, xyz, k,
                       (const int64_t*)idx, n_bones, bones, weights);
  LBS_HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_lbs_bones(int n_bones, const float* bones, const float* motions, int n_adj, const void* relations,
                   int idx_bytes, float* bone_rec, int* degenerate, void* stream) {
  if (n_bones < 0 || n_adj < 0 || (idx_bytes != 4 && idx_bytes != 8))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: bad arguments (n_bones=%lld, n_adj=%lld)", n_bones, n_adj);
  if (!degenerate) return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: null degenerate counter");
  if (n_bones > 0 && (!bones || !motions || !bone_rec || (n_adj > 0 && !relations)))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: null array");
  if (n_bones > 0 && !aligned16(bone_rec)) return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_bones: bone_rec not 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, degenerate);
  if (n_bones > 0) {
    if (idx_bytes == 4)
      hipLaunchKernelGGL(k_lbs_bones<int32_t>, dim3(lbs_blocks(n_bones)), dim3(kLbsBlock), 0, st, n_bones, bones,
                         motions, n_adj, (const int32_t*)relations, (float4*)bone_rec, degenerate);
    else
      hipLaunchKernelGGL(k_lbs_bones<int64_t>, dim3(lbs_blocks(n_bones)), dim3(kLbsBlock), 0, st, n_bones, bones,
                         motions, n_adj, (const int64_t*)relations, (float4*)bone_rec, degenerate);
  }
  LBS_HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

int empm_lbs_points(int n, const float* xyz, const float* quat, int k, const void* idx, int idx_bytes,
                    const float* weights, int n_bones, const float* bone_rec, const int* degenerate, float* xyz_out,
                    float* quat_out, void* stream) {
  if (n < 0 || k < 1 || k > kMaxK || (idx_bytes != 4 && idx_bytes != 8))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: need 1 <= k <= 32 and 4- or 8-byte indices (k=%lld)", k);
  if (n == 0) return EMPM_OK;
  if (n_bones < 1) return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: n_bones=%lld", n_bones);
  if (!xyz || !idx || !bone_rec || !degenerate || !xyz_out || (quat && !quat_out))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: null array");
  if (xyz_out == xyz || (quat && quat_out == quat))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: outputs alias inputs");
  if (!aligned16(bone_rec) || (quat && (!aligned16(quat) || !aligned16(quat_out))))
    return lbs_err(EMPM_E_INVALID_ARG, "empm_lbs_points: bone_rec and quaternions must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  if (idx_bytes == 4)
    hipLaunchKernelGGL(k_lbs_points<int32_t>, dim3(lbs_blocks(n)), dim3(kLbsBlock), 0, st, n, xyz, quat, k,
                       (const int32_t*)idx, weights, n_bones, (const float4*)bone_rec, degenerate, xyz_out, quat_out);
  else
    hipLaunchKernelGGL(k_lbs_points<int64_t>, dim3(lbs_blocks(n)), dim3(kLbsBlock), 0, st, n, xyz, quat, k,
                       (const int64_t*)idx, weights, n_bones, (const float4*)bone_rec, degenerate, xyz_out, quat_out);
  LBS_HIP_TRY(hipGetLastError());
  return EMPM_OK;
}

}  // extern "C"
