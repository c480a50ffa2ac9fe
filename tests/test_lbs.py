"""Linear-blend-skinning interpolation of Gaussians on the CPU: the float64
restatement (oracle/lbs.py) pinned by library routines (scipy's k-d tree,
rotations and Kabsch alignment) and closed forms (rigid motions, pure
translations, the identity fallback), and the host-side argument handling of
the GPU mirror (empm_amd.gaussian_splatting)."""
import numpy as np
import pytest
import torch
from scipy.spatial import cKDTree
from scipy.spatial.transform import Rotation

from empm_amd import gaussian_splatting as gs
from oracle import lbs


def _wxyz(r: Rotation):
    x, y, z, w = r.as_quat().T
    return np.stack([w, x, y, z], -1)


def test_knn_matches_kdtree():
    rng = np.random.default_rng(0)
    r = rng.random((500, 3))
    q = rng.random((300, 3))
    for k in (1, 5, 16):
        idx, dist = lbs.knn(q, r, k)
        d_ref, i_ref = cKDTree(r).query(q, k=k)
        d_ref = d_ref.reshape(len(q), k)
        i_ref = i_ref.reshape(len(q), k)
        np.testing.assert_allclose(dist, d_ref, rtol=1e-12, atol=1e-14)
        np.testing.assert_array_equal(idx, i_ref)
    # get_topk_indices: K + 1 nearest of the set itself, the point dropped
    idx, _ = lbs.knn(r, r, 16, skip_first=1)
    _, i_ref = cKDTree(r).query(r, k=17)
    np.testing.assert_array_equal(idx, i_ref[:, 1:])
    assert not np.any(idx == np.arange(len(r))[:, None])


def test_knn_ties_keep_smaller_index():
    r = np.array([[1.0, 0, 0], [-1.0, 0, 0], [0, 1.0, 0], [0, 0, 2.0]])
    idx, dist = lbs.knn(np.zeros((1, 3)), r, 3)
    assert idx.tolist() == [[0, 1, 2]]
    assert np.allclose(dist, 1.0)


def test_quaternion_of_rotation_all_branches():
    rots = [Rotation.from_rotvec(v) for v in (
        [0.1, -0.2, 0.3],               # trace > 0
        [np.pi * 0.95, 0.1, 0.05],      # m00 largest
        [0.05, np.pi * 0.95, 0.1],      # m11 largest
        [0.1, 0.05, np.pi * 0.95],      # m22 largest
    )] + list(Rotation.random(200, random_state=3))
    R = np.stack([r.as_matrix() for r in rots])
    q = lbs.quat_from_rot(R)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-7)
    # the quaternion's rotation is R (quat2mat, dynamic_utils.py:5-22)
    np.testing.assert_allclose(lbs.quat_to_rot(q), R, atol=1e-7)
    # same rotation as scipy's, up to the sign of the quaternion
    qs = np.stack([_wxyz(r) for r in rots])
    s = np.sign(np.sum(q * qs, axis=1))
    np.testing.assert_allclose(q, qs * s[:, None], atol=1e-7)
    # the branch's own component is the positive one
    tr = np.trace(R, axis1=1, axis2=2)
    d = np.stack([R[:, 0, 0], R[:, 1, 1], R[:, 2, 2]], 1)
    big = np.where(tr > 0, 0, 1 + np.where((d[:, 0] > d[:, 1]) & (d[:, 0] > d[:, 2]), 0,
                                           np.where(d[:, 1] > d[:, 2], 1, 2)))
    assert np.all(q[np.arange(len(q)), big] > 0)
    for b in range(4):
        assert np.any(big == b)


def test_quaternion_product_is_rotation_product():
    a = Rotation.random(50, random_state=1)
    b = Rotation.random(50, random_state=2)
    qa, qb = _wxyz(a), _wxyz(b)
    qab = lbs.quat_mul(qa, qb)
    np.testing.assert_allclose(lbs.quat_to_rot(qab), a.as_matrix() @ b.as_matrix(), atol=1e-12)
    s = np.sign(np.sum(qab * _wxyz(a * b), axis=1))
    np.testing.assert_allclose(qab, _wxyz(a * b) * s[:, None], atol=1e-12)


def _cloud(n, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)) * 0.2


def test_bone_rotations_are_the_kabsch_rotation():
    bones = _cloud(300, 4)
    rng = np.random.default_rng(5)
    motions = 0.01 * rng.standard_normal(bones.shape)
    rel, _ = lbs.knn(bones, bones, 16, skip_first=1)
    R, degen = lbs.bone_rotations(bones, motions, rel)
    assert degen == 0
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=1e-12)
    new = bones + motions
    for b in range(0, 300, 7):
        o = bones[rel[b]] - bones[b]
        n = new[rel[b]] - new[b]
        ref, _ = Rotation.align_vectors(n, o)
        np.testing.assert_allclose(R[b], ref.as_matrix(), atol=1e-9)


def test_rigid_motion_moves_every_gaussian_rigidly():
    bones = _cloud(400, 6)
    xyz = _cloud(1000, 7) + 0.01
    Q = Rotation.from_rotvec([0.3, -0.5, 0.2])
    c = np.array([0.1, 0.1, 0.1])
    t = np.array([0.02, -0.01, 0.03])
    moved = (bones - c) @ Q.as_matrix().T + c + t
    rel, _ = lbs.knn(bones, bones, 16, skip_first=1)
    idx, _ = lbs.knn(xyz, bones, 16)
    quat = _wxyz(Rotation.random(1000, random_state=8)) * 1.7    # raw, not unit
    out, qo = lbs.interpolate(bones, moved - bones, rel, xyz, quat, None, idx)
    np.testing.assert_allclose(out, (xyz - c) @ Q.as_matrix().T + c + t, atol=1e-12)
    qq = lbs.quat_from_rot(Q.as_matrix())
    np.testing.assert_allclose(qo, lbs.quat_mul(np.broadcast_to(qq, quat.shape), quat), atol=1e-12)
    # the weights never matter for a rigid motion
    w = np.random.default_rng(9).random(idx.shape)
    out2, _ = lbs.interpolate(bones, moved - bones, rel, xyz, None, w / w.sum(1, keepdims=True), idx)
    np.testing.assert_allclose(out2, out, atol=1e-12)


def test_translation_and_weights_closed_forms():
    bones = _cloud(200, 10)
    xyz = _cloud(300, 11)
    rel, _ = lbs.knn(bones, bones, 16, skip_first=1)
    idx, dist = lbs.knn(xyz, bones, 5)
    w = lbs.inv_dist_weights(xyz, bones, idx)
    np.testing.assert_allclose(w.sum(1), 1.0, atol=1e-14)
    # w_j (d_j + 1e-6) is the same for every j of a row: 1 / sum_j 1 / (d_j + 1e-6)
    np.testing.assert_allclose(w * (dist + 1e-6),
                               np.broadcast_to(1.0 / (1.0 / (dist + 1e-6)).sum(1, keepdims=True), w.shape), rtol=1e-12)
    m = np.tile([0.01, 0.02, -0.03], (len(bones), 1))
    out, _ = lbs.interpolate(bones, m, rel, xyz, None, None, idx)
    np.testing.assert_allclose(out, xyz + m[0], atol=1e-14)


def _rot_z(t):
    c, s_ = np.cos(t), np.sin(t)
    return np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])


def _generic_bones(n, seed):
    rng = np.random.default_rng(seed)
    bones = rng.uniform(0, 1, (n, 3))
    rel, _ = lbs.knn(bones, bones, 4, skip_first=1)
    motions = 0.01 * rng.standard_normal((n, 3))
    return bones, rel, motions


def test_collinear_bones_take_their_rank1_rotation():
    # bones on a line along x sheared in y: every F = (1, 0.1, 0) (x) (1, 0, 0)
    # times a sum of squared offsets has rank 1, so every bone keeps its own
    # rank-1 rotation (dynamic_utils.py:365-397 with all bones rank 1), which
    # takes e_x to the sheared direction about e_z: a rotation by atan(0.1)
    bones = np.stack([np.linspace(0, 1, 40), np.zeros(40), np.zeros(40)], 1)
    rel, _ = lbs.knn(bones, bones, 4, skip_first=1)
    motions = np.zeros_like(bones)
    motions[:, 1] = 0.1 * bones[:, 0]
    R, lt2 = lbs.bone_rotations(bones, motions, rel)
    assert lt2 == 40
    np.testing.assert_allclose(R, np.broadcast_to(_rot_z(np.arctan(0.1)), R.shape), atol=1e-12)
    # and the blend rotates a point's offset from its bones by that rotation
    xyz = np.array([[0.5, 0.01, 0.0]])
    idx, _ = lbs.knn(xyz, bones, 4)
    out, q = lbs.interpolate(bones, motions, rel, xyz, np.array([[1.0, 0, 0, 0]]), None, idx)
    w = lbs.inv_dist_weights(xyz, bones, idx)
    want = (w[0][:, None] * ((xyz[0] - bones[idx[0]]) @ _rot_z(np.arctan(0.1)).T + bones[idx[0]]
                             + motions[idx[0]])).sum(0)
    np.testing.assert_allclose(out[0], want, atol=1e-14)
    t = np.arctan(0.1) / 2
    np.testing.assert_allclose(q, [[np.cos(t), 0, 0, np.sin(t)]], atol=1e-12)


def test_one_rank1_bone_is_broadcast_and_two_give_identity():
    bones, rel, motions = _generic_bones(30, 41)
    # bone 0's adjacent bones on the line through it along e_y, moved along e_z:
    # F_0 = e_z-ish (x) e_y, rank 1
    bones[1] = bones[0] + [0.0, 0.05, 0.0]
    bones[2] = bones[0] + [0.0, -0.03, 0.0]
    rel[0] = [1, 2, 1, 2]
    motions[1] = motions[0] + [0.0, 0.05, 0.004]
    motions[2] = motions[0] + [0.0, -0.03, -0.0024]
    R, lt2 = lbs.bone_rotations(bones, motions, rel)
    assert lt2 == 1
    F0 = lbs.bone_covariance(bones, motions, rel)[0]
    R0 = lbs.rank1_rotation(F0)
    np.testing.assert_allclose(R, np.broadcast_to(R0, R.shape), atol=1e-14)
    d = F0[:, 1] / np.linalg.norm(F0[:, 1])
    np.testing.assert_allclose(R0 @ [1.0, 0, 0], d if d[np.argmax(np.abs(d))] > 0 else -d, atol=1e-12)
    assert abs(np.linalg.det(R0) - 1) < 1e-12
    # a second rank-1 bone: the set of two is not assignable to all bones
    bones[4] = bones[3] + [0.04, 0.0, 0.0]
    rel[3] = [4, 4, 4, 4]
    R, lt2 = lbs.bone_rotations(bones, motions, rel)
    assert lt2 == 2
    np.testing.assert_array_equal(R, np.broadcast_to(np.eye(3), R.shape))


def test_rank0_bones_and_one_full_bone_broadcast():
    # bones whose relations point at themselves have F = 0 (rank 0); with a
    # single rank-3 bone left its rotation is broadcast to all, with two the
    # set is not assignable and every rotation is the identity
    bones, rel, motions = _generic_bones(12, 43)
    rel[2:] = np.arange(2, 12)[:, None]
    R, lt2 = lbs.bone_rotations(bones, motions, rel)
    assert lt2 == 10
    np.testing.assert_array_equal(R, np.broadcast_to(np.eye(3), R.shape))
    rel[1] = 1
    R, lt2 = lbs.bone_rotations(bones, motions, rel)
    assert lt2 == 11
    R_all, _ = lbs.bone_rotations(*_generic_bones(12, 43)[:1], motions, _generic_bones(12, 43)[1])
    np.testing.assert_allclose(R, np.broadcast_to(R_all[0], R.shape), atol=1e-14)


def test_rank1_with_a_second_singular_value_above_1e8_gives_identity():
    # rank 1 for matrix_rank (s1 <= 3 eps_f32 s0) but s1 > 1e-8: the
    # reference's allclose assert fails and every rotation is the identity
    k = np.arange(5.0)
    bones = np.stack([10 * k, 5e-4 * k ** 2, np.zeros(5)], 1)  # a slightly bent line
    rel, _ = lbs.knn(bones, bones, 2, skip_first=1)
    motions = np.zeros_like(bones)
    motions[:, 1] = 0.1 * bones[:, 0]
    F = lbs.bone_covariance(bones, motions, rel)
    s = np.linalg.svd(F, compute_uv=False)
    assert np.all(s[:, 1] <= 3 * lbs.EPS32 * s[:, 0]) and np.any(s[:, 1] > 1e-8)
    R, lt2 = lbs.bone_rotations(bones, motions, rel)
    assert lt2 == 5
    np.testing.assert_array_equal(R, np.broadcast_to(np.eye(3), R.shape))

def test_quat2mat_and_mat2quat():
    R = Rotation.random(300, random_state=21)
    rs = [Rotation.from_rotvec(v) for v in ([np.pi * 0.999, 0, 0], [0, np.pi * 0.999, 0], [0, 0, np.pi * 0.999],
                                             [np.pi, 0, 0], [0.2, 0.1, -0.3])]
    M = np.concatenate([R.as_matrix(), np.stack([r.as_matrix() for r in rs])])
    q = gs.mat2quat(torch.tensor(M)).numpy()
    # every branch of the reference's rule is reached and gives a unit
    # quaternion of the matrix (quat2mat restated in the oracle)
    tr = np.trace(M, axis1=1, axis2=2)
    assert np.any(tr <= -1 + 1e-9) or np.any(tr < -0.99)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-9)
    np.testing.assert_allclose(lbs.quat_to_rot(q), M, atol=1e-9)
    # quat2mat normalises and matches scipy
    raw = _wxyz(R) * 3.0
    np.testing.assert_allclose(gs.quat2mat(torch.tensor(raw)).numpy(), R.as_matrix(), atol=1e-12)


def test_mirror_refuses_cpu_tensors():
    b = torch.rand(20, 3)
    with pytest.raises(RuntimeError, match="GPU"):
        gs.get_topk_indices(b, K=4)
    with pytest.raises(RuntimeError, match="GPU"):
        gs.knn_weights_sparse(b, torch.rand(5, 3), K=4)
    with pytest.raises(RuntimeError, match="GPU"):
        gs.interpolate_motions_speedup(b, b, torch.zeros(20, 4, dtype=torch.long), torch.rand(5, 3),
                                       weights_indices=torch.zeros(5, 4, dtype=torch.long))
