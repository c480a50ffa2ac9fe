"""NumPy restatement (float64) of the linear-blend-skinning motion
interpolation of Gaussians on simulated particles.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's CPU baseline,
as the checker of empm_amd.gaussian_splatting, never as the product.

Reference: third_party/gaussian_splatting/dynamic_utils.py
- get_topk_indices (:248-265): the K + 1 nearest points of each point, the
  first (itself) dropped;
- knn_weights_sparse (:295-302) and calc_weights_vals_from_indices
  (:282-292): inverse-distance weights 1 / (d + 1e-6) of the K nearest bones,
  normalised per row;
- interpolate_motions_speedup (:305-493): per bone, F = sum_j n_j o_j^T over
  the adjacent bones (W = I, :325-329), a rotation from the SVD of F with two
  determinant corrections (:343-363), translation = motion; per Gaussian the
  weighted sum of R_j (x - b_j) + m_j + b_j, and (with quaternions) the
  weighted sum of the bones' unit quaternions, normalised, times the
  Gaussian's quaternion (:452-485).

The quaternion of a rotation matrix is kornia 0.8.1's
rotation_matrix_to_quaternion (kornia is a dependency of the reference,
env_setup/requirements.txt:29, not installed here): Shepperd's method, branch
trace > 0, else the largest diagonal entry, with eps = 1e-8 under the square
root.

Ranks (torch.linalg.matrix_rank, :331: singular values above
3 * eps_float32 * s_max count) split the bones (:333-338).  The reference
computes one rotation per rank-2/3 bone (:342-363) and, when any bone has
rank 1, replaces that set by one rotation per rank-1 bone (:365-397); the
set is then assigned to all bones (:399-405), which works only when it holds
every bone or exactly one (a broadcast of that rotation), and otherwise
raises into the identity for every bone.  The rank-1 rotation takes the
x axis to the bone's direction (the left singular vector `axis` of F) about
axis x e_x (the identity when that is shorter than 1e-6); it is the identity
for all when a rank-1 F has a second singular value above 1e-8 (the
allclose assert, :375).  The sign of `axis` is torch.svd's on the
reference's device; this restatement (and the HIP path) takes the one whose
largest-magnitude component is positive: parity unpinned for that sign.
"""
from __future__ import annotations

import numpy as np

EPS32 = float(np.finfo(np.float32).eps)


def knn(q, r, k, skip_first=0, chunk=4096):
    """(idx, dist) of the k nearest rows of r for each row of q, nearest
    first, ties to the smaller index; skip_first drops the first entry."""
    q = np.asarray(q, np.float64)
    r = np.asarray(r, np.float64)
    kk = k + skip_first
    idx = np.empty((q.shape[0], k), np.int64)
    dist = np.empty((q.shape[0], k), np.float64)
    for s in range(0, q.shape[0], chunk):
        d = ((q[s:s + chunk, None, :] - r[None, :, :]) ** 2).sum(-1)
        o = np.argsort(d, axis=1, kind="stable")[:, :kk][:, skip_first:]
        idx[s:s + chunk] = o
        dist[s:s + chunk] = np.sqrt(np.take_along_axis(d, o, axis=1))
    return idx, dist


def inv_dist_weights(xyz, bones, idx):
    """calc_weights_vals_from_indices (dynamic_utils.py:282-292)."""
    xyz = np.asarray(xyz, np.float64)
    bones = np.asarray(bones, np.float64)
    d = np.linalg.norm(xyz[:, None, :] - bones[idx], axis=2)
    w = 1.0 / (d + 1e-6)
    return w / w.sum(axis=1, keepdims=True)


def quat_from_rot(R):
    """kornia 0.8.1 rotation_matrix_to_quaternion, (w, x, y, z), for R (..., 3, 3)."""
    R = np.asarray(R, np.float64)
    m00, m01, m02 = R[..., 0, 0], R[..., 0, 1], R[..., 0, 2]
    m10, m11, m12 = R[..., 1, 0], R[..., 1, 1], R[..., 1, 2]
    m20, m21, m22 = R[..., 2, 0], R[..., 2, 1], R[..., 2, 2]
    tr = m00 + m11 + m22
    eps = 1e-8
    with np.errstate(invalid="ignore", divide="ignore"):
        s0 = np.sqrt(tr + 1.0 + eps) * 2.0
        q0 = np.stack([0.25 * s0, (m21 - m12) / s0, (m02 - m20) / s0, (m10 - m01) / s0], -1)
        s1 = np.sqrt(1.0 + m00 - m11 - m22 + eps) * 2.0
        q1 = np.stack([(m21 - m12) / s1, 0.25 * s1, (m01 + m10) / s1, (m02 + m20) / s1], -1)
        s2 = np.sqrt(1.0 + m11 - m00 - m22 + eps) * 2.0
        q2 = np.stack([(m02 - m20) / s2, (m01 + m10) / s2, 0.25 * s2, (m12 + m21) / s2], -1)
        s3 = np.sqrt(1.0 + m22 - m00 - m11 + eps) * 2.0
        q3 = np.stack([(m10 - m01) / s3, (m02 + m20) / s3, (m12 + m21) / s3, 0.25 * s3], -1)
    c1 = ((m00 > m11) & (m00 > m22))[..., None]
    c2 = (m11 > m22)[..., None]
    return np.where((tr > 0.0)[..., None], q0, np.where(c1, q1, np.where(c2, q2, q3)))


def quat_to_rot(q):
    """quat2mat (dynamic_utils.py:5-22): rotation of a (w, x, y, z)
    quaternion, normalised first."""
    q = np.asarray(q, np.float64)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    r, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = 1 - 2 * (y * y + z * z)
    R[..., 0, 1] = 2 * (x * y - r * z)
    R[..., 0, 2] = 2 * (x * z + r * y)
    R[..., 1, 0] = 2 * (x * y + r * z)
    R[..., 1, 1] = 1 - 2 * (x * x + z * z)
    R[..., 1, 2] = 2 * (y * z - r * x)
    R[..., 2, 0] = 2 * (x * z - r * y)
    R[..., 2, 1] = 2 * (y * z + r * x)
    R[..., 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def quat_mul(a, b):
    """Hamilton product (dynamic_utils.py:452-460), (w, x, y, z)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    w1, x1, y1, z1 = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    w2, x2, y2, z2 = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], -1)


def normalize(v, eps=1e-12):
    """torch.nn.functional.normalize along the last axis."""
    return v / np.maximum(np.linalg.norm(v, axis=-1, keepdims=True), eps)


def bone_covariance(bones, motions, relations):
    """F_b = sum_j n_j o_j^T (dynamic_utils.py:321-329)."""
    bones = np.asarray(bones, np.float64)
    motions = np.asarray(motions, np.float64)
    new = bones + motions
    o = bones[relations] - bones[:, None]
    n = new[relations] - new[:, None]
    return np.einsum("bji,bjk->bik", n, o)


def rank1_rotation(F):
    """The rotation of a rank-1 bone (dynamic_utils.py:373-397) from the
    direction of its F's range: x = e_x, perp = axis x x normalised,
    R = axis x^T + perp perp^T + (axis x perp)(x x perp)^T."""
    cols = np.linalg.norm(F, axis=0)
    axis = F[:, int(np.argmax(cols))]
    axis = axis / np.linalg.norm(axis)
    if axis[int(np.argmax(np.abs(axis)))] < 0:
        axis = -axis
    x = np.array([1.0, 0.0, 0.0])
    perp = np.cross(axis, x)
    if np.linalg.norm(perp) < 1e-6:
        return np.eye(3)
    perp = perp / np.linalg.norm(perp)
    third = np.cross(x, perp)
    third_after = np.cross(axis, perp)
    X = np.stack([x, perp, third], -1)
    Y = np.stack([axis, perp, third_after], -1)
    return Y @ X.T


def bone_rotations(bones, motions, relations):
    """(R (n_bones, 3, 3) as the blend uses it, number of bones of rank < 2),
    dynamic_utils.py:321-416 (see the module notes)."""
    F = bone_covariance(bones, motions, relations)
    nb = F.shape[0]
    s_all = np.linalg.svd(F, compute_uv=False)
    rank = (s_all > 3 * EPS32 * s_all[:, :1]).sum(axis=1)
    n_lt2 = int((rank < 2).sum())
    eye = np.broadcast_to(np.eye(3), F.shape).copy()
    one = rank == 1
    if one.any():
        if (s_all[one, 1:] > 1e-8).any():      # the allclose assert fails: identity rows
            return eye, n_lt2
        R1 = np.stack([rank1_rotation(f) for f in F[one]])
        if R1.shape[0] == nb:
            return R1, n_lt2
        if R1.shape[0] == 1:
            return np.broadcast_to(R1[0], F.shape).copy(), n_lt2
        return eye, n_lt2
    full = rank >= 2
    if not full.any():
        return eye, n_lt2
    Ff = F[full]
    U, _, Vh = np.linalg.svd(Ff)
    S = np.broadcast_to(np.eye(3), Ff.shape).copy()
    S[np.linalg.det(Ff) < 0, 2, 2] = -1.0
    R = U @ S @ Vh
    flip = np.abs(np.linalg.det(R) + 1.0) < 1e-3
    S[flip, 2, 2] *= -1.0
    R = U @ S @ Vh
    if R.shape[0] == nb:
        return R, n_lt2
    if R.shape[0] == 1:
        return np.broadcast_to(R[0], F.shape).copy(), n_lt2
    return eye, n_lt2


def interpolate(bones, motions, relations, xyz, quat, weights, idx):
    """interpolate_motions_speedup (dynamic_utils.py:305-493): (xyz', quat')
    (quat' is None when quat is None)."""
    bones = np.asarray(bones, np.float64)
    motions = np.asarray(motions, np.float64)
    xyz = np.asarray(xyz, np.float64)
    if weights is None:
        weights = inv_dist_weights(xyz, bones, idx)
    weights = np.asarray(weights, np.float64)
    R, _ = bone_rotations(bones, motions, relations)
    Rs = R[idx]                                   # (n, k, 3, 3)
    b = bones[idx]
    local = xyz[:, None, :] - b
    pts = np.einsum("nkij,nkj->nki", Rs, local) + motions[idx] + b
    out = (pts * weights[:, :, None]).sum(axis=1)
    if quat is None:
        return out, None
    bq = normalize(quat_from_rot(R))[idx]         # (n, k, 4)
    q = normalize((bq * weights[:, :, None]).sum(axis=1))
    return out, quat_mul(q, np.asarray(quat, np.float64))
