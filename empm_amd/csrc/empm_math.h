// Device math for the MLS-MPM substep on gfx950: 3x3 algebra, a branch-light
// 3x3 SVD with U, V in SO(3), the constitutive models and return maps of the
// reference (phys_sim/diff_simulator/warp_mpm/mpm_utils.py:8-274,603-665)
// and the hand-written reverse-mode of each.
//
// Matrices are float[9], row-major (m[r*3+c]), held in registers: every loop
// below has a compile-time trip count and is fully unrolled.
#pragma once
#include <hip/hip_runtime.h>

#define EMPM_DEV __device__ __forceinline__

namespace empm {

// Hardware transcendentals: v_rcp_f32, v_log_f32 (log2) and v_exp_f32
// (exp2), about 1 ulp on normal inputs (denormals are flushed in this
// library).  The guarded reciprocal `c ? 1.f / x : 0.f` would otherwise
// compile to the correctly rounded division (the divide is hoisted above
// the select without its accuracy tag), and the library's powf, log1pf and
// expm1f expand to 100-250 instructions each.
EMPM_DEV float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
EMPM_DEV float hw_log2(float x) { return __builtin_amdgcn_logf(x); }
EMPM_DEV float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
// x^y for x > 0 (NaN for x < 0, +inf for x = 0 and y < 0, as powf)
EMPM_DEV float hw_pow(float x, float y) { return hw_exp2(y * hw_log2(x)); }

EMPM_DEV void m_zero(float* a) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = 0.f;
}
EMPM_DEV void m_eye(float* a) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = (i % 4 == 0) ? 1.f : 0.f;
}
EMPM_DEV void m_copy(const float* a, float* o) {
#pragma unroll
  for (int i = 0; i < 9; ++i) o[i] = a[i];
}
// o = a * b
EMPM_DEV void m_mul(const float* a, const float* b, float* o) {
  float t[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      t[r * 3 + c] = a[r * 3 + 0] * b[0 * 3 + c] + a[r * 3 + 1] * b[1 * 3 + c] + a[r * 3 + 2] * b[2 * 3 + c];
  m_copy(t, o);
}
// o = a * b^T
EMPM_DEV void m_mul_bt(const float* a, const float* b, float* o) {
  float t[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      t[r * 3 + c] = a[r * 3 + 0] * b[c * 3 + 0] + a[r * 3 + 1] * b[c * 3 + 1] + a[r * 3 + 2] * b[c * 3 + 2];
  m_copy(t, o);
}
// o = a^T * b
EMPM_DEV void m_mul_at(const float* a, const float* b, float* o) {
  float t[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      t[r * 3 + c] = a[0 * 3 + r] * b[0 * 3 + c] + a[1 * 3 + r] * b[1 * 3 + c] + a[2 * 3 + r] * b[2 * 3 + c];
  m_copy(t, o);
}
EMPM_DEV float m_det(const float* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
// cofactor matrix: d det(F) / dF
EMPM_DEV void m_cof(const float* m, float* o) {
  o[0] = m[4] * m[8] - m[5] * m[7];
  o[1] = m[5] * m[6] - m[3] * m[8];
  o[2] = m[3] * m[7] - m[4] * m[6];
  o[3] = m[2] * m[7] - m[1] * m[8];
  o[4] = m[0] * m[8] - m[2] * m[6];
  o[5] = m[1] * m[6] - m[0] * m[7];
  o[6] = m[1] * m[5] - m[2] * m[4];
  o[7] = m[2] * m[3] - m[0] * m[5];
  o[8] = m[0] * m[4] - m[1] * m[3];
}
// o = U diag(d) V^T
EMPM_DEV void m_udvt(const float* U, const float* d, const float* V, float* o) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      o[r * 3 + c] = U[r * 3 + 0] * d[0] * V[c * 3 + 0] + U[r * 3 + 1] * d[1] * V[c * 3 + 1] +
                     U[r * 3 + 2] * d[2] * V[c * 3 + 2];
}

// ---------------------------------------------------------------------------
// SVD F = U diag(s) V^T with U, V in SO(3), |s0| >= |s1| >= |s2|, only s2 may
// be negative.  Cyclic Jacobi on the symmetric F^T F (fixed 4 sweeps: the 3x3
// problem converges to FP32 round-off in 3), eigenvalues sorted with a sign
// flip per column swap to keep det V = +1, then B = F V is QR-factorised by
// Gram-Schmidt with u2 = u0 x u1 so that det U = +1 and s2 = u2 . b2 carries
// the sign of det F.
// ---------------------------------------------------------------------------
// One Jacobi rotation zeroing a_pq of the symmetric A = F^T F, held as its
// diagonal and upper triangle (app, aqq, apq; arp, arq for the third index
// r): with t = tan of the rotation angle, a_pp -= t a_pq, a_qq += t a_pq,
// a_pq = 0 and (a_rp, a_rq) rotate by (c, s); the rotation is accumulated
// into the columns p, q of W.
template <int P, int Q>
EMPM_DEV void jacobi_rot(float& app, float& aqq, float& apq, float& arp, float& arq, float* W) {
  // t = tan of the angle that zeroes apq: the root of t^2 + 2 theta t - 1 = 0
  // of smaller magnitude, theta = (aqq - app) / (2 apq) (scale-free).  A
  // negligible apq (below 1e-30, so its reciprocal stays finite) selects
  // t = 0, the identity, instead of branching around the update.
  const float theta = 0.5f * (aqq - app) * hw_rcp(apq);
  const float tt = copysignf(1.f, theta) * hw_rcp(fabsf(theta) + sqrtf(fmaf(theta, theta, 1.f)));
  const float t = fabsf(apq) < 1e-30f ? 0.f : tt;
  const float c = rsqrtf(fmaf(t, t, 1.f));
  const float s = t * c;
  app -= t * apq;
  aqq += t * apq;
  apq = 0.f;
  const float rp = arp, rq = arq;
  arp = c * rp - s * rq;
  arq = s * rp + c * rq;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float wkp = W[k * 3 + P], wkq = W[k * 3 + Q];
    W[k * 3 + P] = c * wkp - s * wkq;
    W[k * 3 + Q] = s * wkp + c * wkq;
  }
}

template <int J>
EMPM_DEV void swap_cols_neg(float* W, float* ev) {
  if (ev[J] < ev[J + 1]) {
    const float te = ev[J];
    ev[J] = ev[J + 1];
    ev[J + 1] = te;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float tw = W[k * 3 + J];
      W[k * 3 + J] = W[k * 3 + J + 1];
      W[k * 3 + J + 1] = -tw;
    }
  }
}

EMPM_DEV void svd3(const float* F, float* U, float* s, float* V) {
  const auto ata = [&](int r, int c) {
    return F[0 * 3 + r] * F[0 * 3 + c] + F[1 * 3 + r] * F[1 * 3 + c] + F[2 * 3 + r] * F[2 * 3 + c];
  };
  float a00 = ata(0, 0), a11 = ata(1, 1), a22 = ata(2, 2), a01 = ata(0, 1), a02 = ata(0, 2), a12 = ata(1, 2);
  m_eye(V);
#pragma unroll
  for (int sweep = 0; sweep < 4; ++sweep) {
    jacobi_rot<0, 1>(a00, a11, a01, a02, a12, V);
    jacobi_rot<0, 2>(a00, a22, a02, a01, a12, V);
    jacobi_rot<1, 2>(a11, a22, a12, a01, a02, V);
  }
  float ev[3] = {a00, a11, a22};
  swap_cols_neg<0>(V, ev);
  swap_cols_neg<1>(V, ev);
  swap_cols_neg<0>(V, ev);
  float B[9];
  m_mul(F, V, B);
  float b0[3] = {B[0], B[3], B[6]}, b1[3] = {B[1], B[4], B[7]}, b2[3] = {B[2], B[5], B[8]};
  float n0 = sqrtf(b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]);
  float u0[3], u1[3], u2[3];
  if (n0 > 0.f) {
    const float in0 = 1.f / n0;
    u0[0] = b0[0] * in0;
    u0[1] = b0[1] * in0;
    u0[2] = b0[2] * in0;
  } else {
    u0[0] = 1.f;
    u0[1] = 0.f;
    u0[2] = 0.f;
  }
  const float d01 = u0[0] * b1[0] + u0[1] * b1[1] + u0[2] * b1[2];
  float w1[3] = {b1[0] - d01 * u0[0], b1[1] - d01 * u0[1], b1[2] - d01 * u0[2]};
  const float n1 = sqrtf(w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]);
  if (n1 > 1e-30f) {
    const float in1 = 1.f / n1;
    u1[0] = w1[0] * in1;
    u1[1] = w1[1] * in1;
    u1[2] = w1[2] * in1;
  } else {
    // any unit vector orthogonal to u0
    const float ax = fabsf(u0[0]), ay = fabsf(u0[1]), az = fabsf(u0[2]);
    float a[3] = {0.f, 0.f, 0.f};
    if (ax <= ay && ax <= az) a[0] = 1.f;
    else if (ay <= az) a[1] = 1.f;
    else a[2] = 1.f;
    const float d = u0[0] * a[0] + u0[1] * a[1] + u0[2] * a[2];
    a[0] -= d * u0[0];
    a[1] -= d * u0[1];
    a[2] -= d * u0[2];
    const float ia = rsqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    u1[0] = a[0] * ia;
    u1[1] = a[1] * ia;
    u1[2] = a[2] * ia;
  }
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  s[0] = n0;
  s[1] = u1[0] * b1[0] + u1[1] * b1[1] + u1[2] * b1[2];
  s[2] = u2[0] * b2[0] + u2[1] * b2[1] + u2[2] * b2[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    U[k * 3 + 0] = u0[k];
    U[k * 3 + 1] = u1[k];
    U[k * 3 + 2] = u2[k];
  }
}

// ---------------------------------------------------------------------------
// Constitutive parameters that are scalars of MPMModelStruct
// (mpm_data_structure.py:390-422).
// ---------------------------------------------------------------------------
struct MatParams {
  int material;
  float alpha, hardening, xi, plastic_viscosity, softening;
};

// Return-map branch codes
enum : int { RM_IDENTITY = 0, RM_POWER = 1, RM_ROTATION = 2, RM_DAMAGED = 3 };

// The result of the constitutive forward pass for one particle.  Every
// return map of the reference has the form Fe = U diag(g) V^T with
// g_i = A * sbar_i^k (A, k symmetric in the singular values), the identity
// (Fe = F_trial) or the rotation U V^T; the stress is evaluated on the SVD of
// Fe, which is (U, g, V) itself, so one SVD per particle serves both stages.
struct ConstitState {
  float U[9], V[9];
  float s[3];    // singular values of F_trial (signed, s2 may be < 0)
  float se[3];   // singular values of Fe in the (U, V) frame
  float Fe[9];
  float J;       // det(Fe), as wp.determinant at mpm_utils.py:636
  int branch;
  float k;       // exponent of the power map
  float lo;      // clamp threshold of sbar (0.01, or 1e-14 for sand on |s|)
  int use_abs;   // sand uses |s|
  float ys_soft; // von Mises: the yield stress after softening, before hardening (decides mu, lam = 0)
  bool have_svd;
};

EMPM_DEV bool is_plastic(int m) { return m == 1 || m == 2 || m == 3 || m == 5 || m == 7; }

// ------------------------- return maps (forward) ---------------------------
// von_mises_return_mapping (mpm_utils.py:87-130) and the damage variant
// (mpm_utils.py:133-186); material 7 uses the former.
EMPM_DEV void rm_von_mises(ConstitState& st, const float* Ftr, float& mu, float& lam, float& ys, const MatParams& P,
                           bool damage) {
  // Every branch only sets locals; the state and the in/out parameters are
  // written once at the end (stores to different fields on different
  // branches otherwise merge into one store through a selected address,
  // which keeps the fields out of registers).
  float e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) e[i] = logf(fmaxf(st.s[i], 0.01f));
  const float temp = (e[0] + e[1] + e[2]) / 3.f;
  const float tr = e[0] + e[1] + e[2];
  float tau[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) tau[i] = 2.f * mu * e[i] + lam * tr;
  const float st3 = (tau[0] + tau[1] + tau[2]) / 3.f;
  const float c0 = tau[0] - st3, c1 = tau[1] - st3, c2 = tau[2] - st3;
  const float cn = sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
  float g[3] = {st.s[0], st.s[1], st.s[2]};
  float mu_n = mu, lam_n = lam, ys_n = ys, ys_soft = ys, k = 1.f;
  int branch = RM_IDENTITY;
  if (cn > ys) {
    if (damage && ys <= 0.f) {
      branch = RM_DAMAGED;
    } else {
      float eh[3] = {e[0] - temp, e[1] - temp, e[2] - temp};
      const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]) + 1e-6f;
      const float dg = ehn - ys / (2.f * mu);
      if (damage) {
        const float q0 = (dg / ehn) * eh[0], q1 = (dg / ehn) * eh[1], q2 = (dg / ehn) * eh[2];
        ys_n = ys - P.softening * sqrtf(q0 * q0 + q1 * q1 + q2 * q2);
        ys_soft = ys_n;
        if (ys_n <= 0.f) {
          mu_n = 0.f;
          lam_n = 0.f;
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) g[i] = expf(e[i] - (dg / ehn) * eh[i]);
      if (P.hardening == 1.f) ys_n = ys_n + 2.f * mu_n * P.xi * dg;
      branch = RM_POWER;
      k = 1.f - dg / ehn;
    }
  }
  st.lo = 0.01f;
  st.use_abs = 0;
  st.branch = branch;
  st.k = k;
  st.ys_soft = ys_soft;
  st.se[0] = g[0];
  st.se[1] = g[1];
  st.se[2] = g[2];
  if (branch == RM_POWER) m_udvt(st.U, g, st.V, st.Fe);
  else m_copy(Ftr, st.Fe);
  mu = mu_n;
  lam = lam_n;
  ys = ys_n;
}

// viscoplasticity_return_mapping_with_StVK (mpm_utils.py:189-234)
EMPM_DEV void rm_visco(ConstitState& st, const float* Ftr, float mu, float ys, const MatParams& P, float dt) {
  // locals per branch, state written once (as in rm_von_mises)
  float sg[3], b[3], e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    sg[i] = fmaxf(st.s[i], 0.01f);
    b[i] = sg[i] * sg[i];
    e[i] = logf(sg[i]);
  }
  const float tr = e[0] + e[1] + e[2];
  float eh[3] = {e[0] - tr / 3.f, e[1] - tr / 3.f, e[2] - tr / 3.f};
  float s[3] = {2.f * mu * eh[0], 2.f * mu * eh[1], 2.f * mu * eh[2]};
  const float sn = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  const float y = sn - sqrtf(2.f / 3.f) * ys;
  float g[3] = {st.s[0], st.s[1], st.s[2]};
  float k = 1.f;
  int branch = RM_IDENTITY;
  if (y > 0.f) {
    const float muh = mu * (b[0] + b[1] + b[2]) / 3.f;
    const float snew = sn - y / (1.f + P.plastic_viscosity / (2.f * muh * dt));
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i] = expf(1.f / (2.f * mu) * ((snew / sn) * s[i]) + tr / 3.f);
    branch = RM_POWER;
    k = snew / sn;
  }
  st.lo = 0.01f;
  st.use_abs = 0;
  st.branch = branch;
  st.k = k;
  st.se[0] = g[0];
  st.se[1] = g[1];
  st.se[2] = g[2];
  if (branch == RM_POWER) m_udvt(st.U, g, st.V, st.Fe);
  else m_copy(Ftr, st.Fe);
}

// sand_return_mapping (mpm_utils.py:237-274)
EMPM_DEV void rm_sand(ConstitState& st, const float* Ftr, float mu, float lam, const MatParams& P) {
  // locals per branch, state written once (as in rm_von_mises)
  float e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) e[i] = logf(fmaxf(fabsf(st.s[i]), 1e-14f));
  const float tr = e[0] + e[1] + e[2];
  float eh[3] = {e[0] - tr / 3.f, e[1] - tr / 3.f, e[2] - tr / 3.f};
  const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
  const float dg = ehn + (3.f * lam + 2.f * mu) / (2.f * mu) * tr * P.alpha;
  float g[3] = {st.s[0], st.s[1], st.s[2]};
  float k = 1.f;
  int branch = RM_IDENTITY;
  if (!(dg <= 0.f)) {  // the reference's test is delta_gamma <= 0 for the identity
    if (tr > 0.f) {
      branch = RM_ROTATION;
      g[0] = g[1] = g[2] = 1.f;
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i) g[i] = expf(e[i] - eh[i] * (dg / ehn));
      branch = RM_POWER;
      k = 1.f - dg / ehn;
    }
  }
  st.lo = 1e-14f;
  st.use_abs = 1;
  st.branch = branch;
  st.k = k;
  st.se[0] = g[0];
  st.se[1] = g[1];
  st.se[2] = g[2];
  if (branch == RM_POWER) m_udvt(st.U, g, st.V, st.Fe);
  else if (branch == RM_ROTATION) m_mul_bt(st.U, st.V, st.Fe);
  else m_copy(Ftr, st.Fe);
}

// ------------------------- stresses (forward) -------------------------------
// The stress "t-vector": tau = U diag(t) V^T Fe^T for StVK, NH and DP
// (mpm_utils.py:42-47,61-66,83-84).
EMPM_DEV void stress_t(int m, const float* se, float J, float mu, float lam, float* t) {
  if (m == 1 || m == 3) {  // kirchoff_stress_StVK, mpm_utils.py:50-66
    float e[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) e[i] = logf(fmaxf(se[i], 0.01f));
    const float tr = e[0] + e[1] + e[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] = 2.f * mu * e[i] + lam * tr;
  } else if (m == 2) {  // kirchoff_stress_drucker_prager, mpm_utils.py:69-84
    const float L = logf(se[0]) + logf(se[1]) + logf(se[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] = 2.f * mu * logf(se[i]) * (1.f / se[i]) + lam * L * (1.f / se[i]);
  } else {  // 6, 7: kirchoff_stress_neoHookean, mpm_utils.py:18-47
    float b[3] = {se[0] * se[0], se[1] * se[1], se[2] * se[2]};
    const float mb = (b[0] + b[1] + b[2]) / 3.f;
    const float jp = hw_pow(J, -2.f / 3.f);
#pragma unroll
    for (int i = 0; i < 3; ++i) t[i] = mu * jp * (b[i] - mb) + lam / 2.f * (J * J - 1.f);
  }
}

// compute_stress_from_F_trial (mpm_utils.py:603-665): return map, stress,
// symmetrisation.  mu/lam/ys are updated in place by material 5 and by
// hardening, as the reference does to its model arrays.
EMPM_DEV void constitutive_fwd(ConstitState& st, const float* Ftr, float& mu, float& lam, float& ys,
                               const MatParams& P, float dt, float* tau_sym) {
  const int m = P.material;
  st.have_svd = false;
  st.branch = RM_IDENTITY;
  st.k = 1.f;
  st.lo = 0.01f;
  st.use_abs = 0;
  if (m == 4) {  // snow: no stress branch matches (mpm_utils.py:632-665)
    m_copy(Ftr, st.Fe);
    st.J = m_det(st.Fe);
    m_zero(tau_sym);
    return;
  }
  svd3(Ftr, st.U, st.s, st.V);
  st.have_svd = true;
  if (m == 1 || m == 7) rm_von_mises(st, Ftr, mu, lam, ys, P, false);
  else if (m == 5) rm_von_mises(st, Ftr, mu, lam, ys, P, true);
  else if (m == 2) rm_sand(st, Ftr, mu, lam, P);
  else if (m == 3) rm_visco(st, Ftr, mu, ys, P, dt);
  else {
    m_copy(Ftr, st.Fe);
    st.se[0] = st.s[0];
    st.se[1] = st.s[1];
    st.se[2] = st.s[2];
  }
  st.J = m_det(st.Fe);
  float tau[9];
  if (m == 0 || m == 5) {  // kirchoff_stress_FCR, mpm_utils.py:8-15
    float R[9], D[9];
    m_mul_bt(st.U, st.V, R);
#pragma unroll
    for (int i = 0; i < 9; ++i) D[i] = st.Fe[i] - R[i];
    m_mul_bt(D, st.Fe, tau);
    const float jj = lam * st.J * (st.J - 1.f);
#pragma unroll
    for (int i = 0; i < 9; ++i) tau[i] = 2.f * mu * tau[i] + ((i % 4 == 0) ? jj : 0.f);
  } else {
    float t[3], G[9];
    stress_t(m, st.se, st.J, mu, lam, t);
    m_udvt(st.U, t, st.V, G);
    m_mul_bt(G, st.Fe, tau);
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) tau_sym[r * 3 + c] = 0.5f * (tau[r * 3 + c] + tau[c * 3 + r]);
}

// ---------------------------------------------------------------------------
// Reverse mode.
//
// For an isotropic map G(F) = U diag(g(s)) V^T, with M = U^T Gbar V the
// adjoint in the SVD frame, the adjoint of F is U Fhat V^T with
//   Fhat_kk = sum_i M_ii dg_i/ds_k
//   Fhat_ij = a_ij M_ij + b_ij M_ji                     (i != j)
//   a_ij = (D_ij + P_ij)/2,  b_ij = (D_ij - P_ij)/2,
//   D_ij = (g_i - g_j)/(s_i - s_j),  P_ij = (g_i + g_j)/(s_i + s_j).
// (Derived from dF^ = Om_U S + dS - S Om_V with Om_U = U^T dU and
// Om_V = V^T dV skew.)  The divided differences D_ij are evaluated in closed,
// cancellation-free forms per map, so repeated singular values (F = I) are
// not a special case.
// ---------------------------------------------------------------------------
EMPM_DEV float safe_den(float d) {
  const float a = fabsf(d);
  return a < 1e-20f ? copysignf(1e-20f, d) : d;
}

// The divided differences below only enter derivatives.  They are functions
// of u = a / b through two forms that stay accurate to a few ulp with the
// hardware log2 / exp2 and need no switch to a series near u = 1:
// log(u) / (u - 1) (Goldberg; the quotient is insensitive to the rounding of
// u) and (e^y - 1) / y = (v - 1) / log(v), v = e^y (Kahan; likewise for the
// rounding of v).
constexpr float kLn2 = 0.693147180559945f, kLog2e = 1.44269504088896f;
// log(u) / (u - 1) for u > 0 (1 at u = 1)
EMPM_DEV float log_over(float u) { return u == 1.f ? 1.f : kLn2 * hw_log2(u) * hw_rcp(u - 1.f); }
// (e^y - 1) / y (1 at y = 0; -1/y once e^y underflows, +inf once it overflows)
EMPM_DEV float expm1_over(float y) {
  const float v = hw_exp2(y * kLog2e);
  if (v == 1.f) return 1.f;
  if (v == 0.f) return -hw_rcp(y);
  if (v == __builtin_inff()) return v;
  return (v - 1.f) * hw_rcp(kLn2 * hw_log2(v));
}
// (a^k - b^k) / (a - b) for a, b > 0: b^(k-1) (u^k - 1) / (u - 1), and with
// y = k log u,  (u^k - 1) / (u - 1) = k log_over(u) (e^y - 1) / y.
EMPM_DEV float pow_dd(float a, float b, float k) {
  const float u = a / b;
  return hw_pow(b, k - 1.f) * k * log_over(u) * expm1_over(k * kLn2 * hw_log2(u));
}
// (log a - log b) / (a - b) = log_over(a / b) / b for a, b > 0
EMPM_DEV float log_dd(float a, float b) { return log_over(a / b) / b; }

EMPM_DEV void iso_adj(const float* U, const float* V, const float* s, const float* g, const float* Dd, const float* Jg,
                      const float* Gbar, float* Fbar_out /* += */) {
  float T[9], M[9];
  m_mul_at(U, Gbar, T);  // U^T Gbar
  m_mul(T, V, M);        // U^T Gbar V
  float H[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) H[k * 4] = M[0] * Jg[0 * 3 + k] + M[4] * Jg[1 * 3 + k] + M[8] * Jg[2 * 3 + k];
  // pairs (0,1) (0,2) (1,2) -> Dd[0..2]
#pragma unroll
  for (int pr = 0; pr < 3; ++pr) {
    const int i = pr == 2 ? 1 : 0;
    const int j = pr == 0 ? 1 : 2;
    const float P = (g[i] + g[j]) / safe_den(s[i] + s[j]);
    const float a = 0.5f * (Dd[pr] + P), b = 0.5f * (Dd[pr] - P);
    H[i * 3 + j] = a * M[i * 3 + j] + b * M[j * 3 + i];
    H[j * 3 + i] = a * M[j * 3 + i] + b * M[i * 3 + j];
  }
  float X[9], Y[9];
  m_mul(U, H, X);
  m_mul_bt(X, V, Y);
#pragma unroll
  for (int i = 0; i < 9; ++i) Fbar_out[i] += Y[i];
}

// Derivative of the clamped log: d log(max(s, lo)) / ds  (|s| for sand)
EMPM_DEV float dlog_clamped(float s, float lo, int use_abs) {
  const float a = use_abs ? fabsf(s) : s;
  return a > lo ? hw_rcp(s) : 0.f;
}
EMPM_DEV float sbar_of(float s, float lo, int use_abs) {
  const float a = use_abs ? fabsf(s) : s;
  return fmaxf(a, lo);
}

// Adjoint of the stress stage (tau_raw -> Fe, mu, lam), in the (U, se, V)
// frame; Febar, mubar, lambar are accumulated.
EMPM_DEV void stress_adj(const ConstitState& st, int m, float mu, float lam, const float* taubar, float* Febar,
                         float& mubar, float& lambar) {
  const float* Fe = st.Fe;
  const float J = st.J;
  float cof[9];
  m_cof(Fe, cof);
  if (m == 4) return;
  if (m == 0 || m == 5) {
    // tau = 2 mu (Fe - R) Fe^T + lam J (J-1) I,  R = U V^T
    float R[9], D[9], FFt[9], RFt[9];
    m_mul_bt(st.U, st.V, R);
#pragma unroll
    for (int i = 0; i < 9; ++i) D[i] = Fe[i] - R[i];
    m_mul_bt(Fe, Fe, FFt);
    m_mul_bt(R, Fe, RFt);
    float tr_tb = taubar[0] + taubar[4] + taubar[8];
    float dot = 0.f;
#pragma unroll
    for (int i = 0; i < 9; ++i) dot += taubar[i] * 2.f * (FFt[i] - RFt[i]);
    mubar += dot;
    lambar += J * (J - 1.f) * tr_tb;
    // d(2 mu Fe Fe^T): Febar += 2 mu (taubar + taubar^T) Fe
    float S[9], T1[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) S[r * 3 + c] = taubar[r * 3 + c] + taubar[c * 3 + r];
    m_mul(S, Fe, T1);
#pragma unroll
    for (int i = 0; i < 9; ++i) Febar[i] += 2.f * mu * T1[i];
    // d(-2 mu R Fe^T): Rbar = -2 mu taubar Fe ; Febar += -2 mu taubar^T R
    float Rbar[9], T2[9];
    m_mul(taubar, Fe, Rbar);
#pragma unroll
    for (int i = 0; i < 9; ++i) Rbar[i] *= -2.f * mu;
    m_mul_at(taubar, R, T2);
#pragma unroll
    for (int i = 0; i < 9; ++i) Febar[i] += -2.f * mu * T2[i];
    // R = U diag(1) V^T : g = 1, D = 0, Jg = 0
    const float g1[3] = {1.f, 1.f, 1.f};
    const float Dz[3] = {0.f, 0.f, 0.f};
    const float Jz[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    iso_adj(st.U, st.V, st.se, g1, Dz, Jz, Rbar, Febar);
    // J term
    const float Jbar = lam * (2.f * J - 1.f) * tr_tb;
#pragma unroll
    for (int i = 0; i < 9; ++i) Febar[i] += Jbar * cof[i];
    return;
  }
  // tau = G Fe^T, G = U diag(t) V^T
  const float* s = st.se;
  float t[3];
  stress_t(m, s, J, mu, lam, t);
  float G[9];
  m_udvt(st.U, t, st.V, G);
  {  // direct: Febar += taubar^T G
    float T[9];
    m_mul_at(taubar, G, T);
#pragma unroll
    for (int i = 0; i < 9; ++i) Febar[i] += T[i];
  }
  float Gbar[9];
  m_mul(taubar, Fe, Gbar);
  // M diagonal needed for parameter grads: M = U^T Gbar V
  float T[9], M[9];
  m_mul_at(st.U, Gbar, T);
  m_mul(T, st.V, M);
  float Jg[9], Dd[3];
  float Jbar = 0.f;
  if (m == 1 || m == 3) {  // StVK
    float e[3], de[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      e[i] = logf(fmaxf(s[i], 0.01f));
      de[i] = s[i] > 0.01f ? hw_rcp(s[i]) : 0.f;
    }
    const float tr = e[0] + e[1] + e[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) Jg[i * 3 + k] = ((i == k ? 2.f * mu : 0.f) + lam) * de[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      mubar += M[i * 4] * 2.f * e[i];
      lambar += M[i * 4] * tr;
    }
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
      const int i = pr == 2 ? 1 : 0;
      const int j = pr == 0 ? 1 : 2;
      const bool ci = s[i] > 0.01f, cj = s[j] > 0.01f;
      if (ci && cj) Dd[pr] = 2.f * mu * log_dd(s[i], s[j]);
      else if (!ci && !cj) Dd[pr] = 0.f;
      else Dd[pr] = (t[i] - t[j]) / safe_den(s[i] - s[j]);
    }
  } else if (m == 2) {  // DP stress: t_i = (2 mu log s_i + lam L) / s_i
    float lg[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) lg[i] = logf(s[i]);
    const float L = lg[0] + lg[1] + lg[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float v = lam / (s[k] * s[i]);
        if (i == k) v += (2.f * mu - (2.f * mu * lg[i] + lam * L)) / (s[i] * s[i]);
        Jg[i * 3 + k] = v;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      mubar += M[i * 4] * 2.f * lg[i] / s[i];
      lambar += M[i * 4] * L / s[i];
    }
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
      const int i = pr == 2 ? 1 : 0;
      const int j = pr == 0 ? 1 : 2;
      // (lg_i/s_i - lg_j/s_j)/(s_i - s_j) = (log_dd * s_j - log s_j) / (s_i s_j)
      const float dphi = (log_dd(s[i], s[j]) * s[j] - lg[j]) / (s[i] * s[j]);
      Dd[pr] = 2.f * mu * dphi - lam * L / (s[i] * s[j]);
    }
  } else {  // NH (6, 7)
    const float jp = hw_pow(J, -2.f / 3.f);
    float b[3] = {s[0] * s[0], s[1] * s[1], s[2] * s[2]};
    const float mb = (b[0] + b[1] + b[2]) / 3.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) Jg[i * 3 + k] = mu * jp * ((i == k ? 2.f * s[k] : 0.f) - 2.f * s[k] / 3.f);
    const float djp = (-2.f / 3.f) * jp / J;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      mubar += M[i * 4] * jp * (b[i] - mb);
      lambar += M[i * 4] * 0.5f * (J * J - 1.f);
      Jbar += M[i * 4] * (mu * djp * (b[i] - mb) + lam * J);
    }
#pragma unroll
    for (int pr = 0; pr < 3; ++pr) {
      const int i = pr == 2 ? 1 : 0;
      const int j = pr == 0 ? 1 : 2;
      Dd[pr] = mu * jp * (s[i] + s[j]);
    }
  }
  iso_adj(st.U, st.V, s, t, Dd, Jg, Gbar, Febar);
  if (Jbar != 0.f) {
#pragma unroll
    for (int i = 0; i < 9; ++i) Febar[i] += Jbar * cof[i];
  }
}

// Adjoint of the return map (Fe -> F_trial, mu, lam, ys).  For the power map
// g_i = A sbar_i^k, log g_i = k ehat_i + tr/3 (ehat = deviator of log sbar),
//   d log g_i / d s_m = ehat_i dk/ds_m + l_m ((1-k)/3 + k [i==m])
//   d log g_i / d theta = ehat_i dk/dtheta,     l_m = d log sbar_m / d s_m.
EMPM_DEV void return_map_adj(const ConstitState& st, const float* Ftr, float mu, float lam, float ys,
                             const MatParams& P, float dt, const float* Febar, float* Ftrbar, float& mubar,
                             float& lambar, float& ysbar) {
  const int m = P.material;
  if (st.branch == RM_IDENTITY || st.branch == RM_DAMAGED || !is_plastic(m)) {
#pragma unroll
    for (int i = 0; i < 9; ++i) Ftrbar[i] += Febar[i];
    return;
  }
  const float* s = st.s;
  const float lo = st.lo;
  const int ua = st.use_abs;
  if (st.branch == RM_ROTATION) {
    const float g1[3] = {1.f, 1.f, 1.f};
    const float Dz[3] = {0.f, 0.f, 0.f};
    const float Jz[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    iso_adj(st.U, st.V, s, g1, Dz, Jz, Febar, Ftrbar);
    return;
  }
  // power map
  float sb[3], l[3], e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    sb[i] = sbar_of(s[i], lo, ua);
    l[i] = dlog_clamped(s[i], lo, ua);
    e[i] = logf(sb[i]);
  }
  const float tr = e[0] + e[1] + e[2];
  float eh[3] = {e[0] - tr / 3.f, e[1] - tr / 3.f, e[2] - tr / 3.f};
  const float ehl = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
  const float k = st.k;
  float dk_ds[3] = {0.f, 0.f, 0.f};
  float dk_dmu = 0.f, dk_dlam = 0.f, dk_dys = 0.f;
  if (m == 1 || m == 5 || m == 7) {
    // k = ys / (2 mu n), n = |ehat| + 1e-6
    const float n = ehl + 1e-6f;
    const float inv_ehl = ehl > 0.f ? hw_rcp(ehl) : 0.f;
#pragma unroll
    for (int q = 0; q < 3; ++q) dk_ds[q] = -ys / (2.f * mu * n * n) * (eh[q] * inv_ehl) * l[q];
    dk_dys = 1.f / (2.f * mu * n);
    dk_dmu = -ys / (2.f * mu * mu * n);
  } else if (m == 2) {
    // k = -c tr / |ehat|, c = (3 lam + 2 mu) alpha / (2 mu)
    const float c = (3.f * lam + 2.f * mu) / (2.f * mu) * P.alpha;
    const float inv = 1.f / ehl;
#pragma unroll
    for (int q = 0; q < 3; ++q) dk_ds[q] = -c * (l[q] * inv - tr * eh[q] * l[q] * inv * inv * inv);
    dk_dmu = -tr * inv * (-3.f * lam * P.alpha / (2.f * mu * mu));
    dk_dlam = -tr * inv * (3.f * P.alpha / (2.f * mu));
  } else if (m == 3) {
    // k = (1-q) + q sqrt(2/3) ys / |s|,  |s| = 2 mu |ehat|,
    // q = 1 / (1 + eta / (2 muh dt)), muh = mu sum(sbar^2)/3
    const float sn = 2.f * mu * ehl;
    float sb2 = sb[0] * sb[0] + sb[1] * sb[1] + sb[2] * sb[2];
    const float muh = mu * sb2 / 3.f;
    const float q = 1.f / (1.f + P.plastic_viscosity / (2.f * muh * dt));
    const float c23 = sqrtf(2.f / 3.f);
    const float dk_dsn = -q * c23 * ys / (sn * sn);
    const float dk_dq = -1.f + c23 * ys / sn;
    const float dq_dmuh = q * q * P.plastic_viscosity / (2.f * muh * muh * dt);
    const float inv_ehl = ehl > 0.f ? hw_rcp(ehl) : 0.f;
#pragma unroll
    for (int qq = 0; qq < 3; ++qq) {
      const float dsn = 2.f * mu * eh[qq] * inv_ehl * l[qq];
      const float dmuh = (sb[qq] > lo || ua) ? mu * 2.f * sb[qq] / 3.f * (s[qq] > lo ? 1.f : 0.f) : 0.f;
      dk_ds[qq] = dk_dsn * dsn + dk_dq * dq_dmuh * dmuh;
    }
    dk_dmu = dk_dsn * 2.f * ehl + dk_dq * dq_dmuh * (sb2 / 3.f);
    dk_dys = q * c23 / sn;
  }
  const float* g = st.se;
  float Jg[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int q = 0; q < 3; ++q)
      Jg[i * 3 + q] = g[i] * (eh[i] * dk_ds[q] + l[q] * ((1.f - k) / 3.f + (i == q ? k : 0.f)));
  // M diagonal for parameter grads
  float T[9], M[9];
  m_mul_at(st.U, Febar, T);
  m_mul(T, st.V, M);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float w = M[i * 4] * g[i] * eh[i];
    mubar += w * dk_dmu;
    lambar += w * dk_dlam;
    ysbar += w * dk_dys;
  }
  // divided differences of g_i = A sbar_i^k, A = exp((1-k) tr / 3)
  const float A = hw_exp2((1.f - k) * tr * (kLog2e / 3.f));
  float Dd[3];
#pragma unroll
  for (int pr = 0; pr < 3; ++pr) {
    const int i = pr == 2 ? 1 : 0;
    const int j = pr == 0 ? 1 : 2;
    const float ai = ua ? fabsf(s[i]) : s[i], aj = ua ? fabsf(s[j]) : s[j];
    const bool ci = ai > lo, cj = aj > lo;
    const bool same_sign = !ua || ((s[i] >= 0.f) == (s[j] >= 0.f));
    if (ci && cj && same_sign) {
      // (sbar_i^k - sbar_j^k)/(s_i - s_j); with |.| and a common sign the
      // denominator is sign * (|s_i| - |s_j|)
      const float sg = (ua && s[i] < 0.f) ? -1.f : 1.f;
      Dd[pr] = sg * A * pow_dd(ai, aj, k);
    } else if (!ci && !cj) {
      Dd[pr] = 0.f;
    } else {
      Dd[pr] = (g[i] - g[j]) / safe_den(s[i] - s[j]);
    }
  }
  iso_adj(st.U, st.V, s, g, Dd, Jg, Febar, Ftrbar);
}

// Adjoint of the constitutive state update of the von Mises return maps on
// their yielding branch: material 5's softening (mpm_utils.py:162-167:
// Y' = Y - softening |(dgamma / n) ehat|, and mu' = lam' = 0 once Y' <= 0)
// and hardening (:124-127, :178-181: Y'' = Y' + 2 mu' xi dgamma), with
// n = |ehat| + 1e-6, dgamma = n - Y / (2 mu), ehat the deviator of
// log max(s, 0.01).  Other branches leave (Y, mu, lam) unchanged.  Given the
// cotangents of the updated values (ysb1, mub1, lamb1), accumulates those of
// the values before the update (ysb0, mub0, lamb0) and of F_trial (through
// its singular values: d h(s) / dF = U diag(dh/ds) V^T).
EMPM_DEV void vm_state_adj(const ConstitState& st, float mu0, float ys0, float mu1, const MatParams& P,
                           bool damage, float ysb1, float mub1, float lamb1, float* Ftrbar, float& ysb0,
                           float& mub0, float& lamb0) {
  if (st.branch != RM_POWER) {
    ysb0 += ysb1;
    mub0 += mub1;
    lamb0 += lamb1;
    return;
  }
  float e[3], l[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e[i] = logf(fmaxf(st.s[i], 0.01f));
    l[i] = st.s[i] > 0.01f ? hw_rcp(st.s[i]) : 0.f;
  }
  const float m3 = (e[0] + e[1] + e[2]) / 3.f;
  const float eh[3] = {e[0] - m3, e[1] - m3, e[2] - m3};
  const float L = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
  const float n = L + 1e-6f;
  const float dg = n - ys0 / (2.f * mu0);
  float Yb = ysb1;      // cotangent of Y' (after softening)
  float mupb = mub1;    // cotangent of mu'
  float dgb = 0.f, Lb = 0.f;
  if (P.hardening == 1.f) {
    mupb += ysb1 * 2.f * P.xi * dg;
    dgb += ysb1 * 2.f * mu1 * P.xi;
  }
  ysb0 += Yb;
  if (damage) {
    // Y' = Y - softening dgamma L / n;  d(L/n)/dL = 1e-6 / n^2
    dgb -= Yb * P.softening * L / n;
    Lb -= Yb * P.softening * dg * (1e-6f / (n * n));
    if (st.ys_soft > 0.f) {  // mu' = mu, lam' = lam; else both are 0 whatever they were
      mub0 += mupb;
      lamb0 += lamb1;
    }
  } else {
    mub0 += mupb;
    lamb0 += lamb1;
  }
  // dgamma = L + 1e-6 - Y / (2 mu)
  Lb += dgb;
  ysb0 -= dgb / (2.f * mu0);
  mub0 += dgb * ys0 / (2.f * mu0 * mu0);
  // L = |ehat|: dL/ds_m = (ehat_m / L) d log sbar_m / ds_m
  const float iL = L > 0.f ? hw_rcp(L) : 0.f;
  const float sb[3] = {Lb * eh[0] * iL * l[0], Lb * eh[1] * iL * l[1], Lb * eh[2] * iL * l[2]};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      Ftrbar[r * 3 + c] += st.U[r * 3 + 0] * sb[0] * st.V[c * 3 + 0] + st.U[r * 3 + 1] * sb[1] * st.V[c * 3 + 1] +
                           st.U[r * 3 + 2] * sb[2] * st.V[c * 3 + 2];
}

}  // namespace empm
