"""Linear-blend-skinning interpolation of Gaussians on the GPU
(empm_amd.gaussian_splatting over empm_knn / empm_lbs_*) against the float64
restatement (oracle/lbs.py), and at the full interactive size against the
closed form of a rigid motion.

Tolerances: positions live in a 0.2 m box, so 2e-5 absolute is ~1e-4
relative to the extent and ~100 float32 ulps of a coordinate; quaternions are
unit, 2e-5 absolute.  Neighbour indices are integers: the GPU's k nearest must
be the oracle's up to exact float32 ties."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from empm_amd import gaussian_splatting as gs
from empm_amd._lib import EmpmError
from oracle import lbs

pytestmark = pytest.mark.gpu

POS_TOL = 2e-5
QUAT_TOL = 2e-5


def _f32(a):
    """float64 arrays holding float32 values: the GPU and the oracle then
    read the same inputs (a weight 1 / (d + 1e-6) of a bone 1e-4 away has a
    relative sensitivity of 1e-4 to a coordinate's float32 rounding)."""
    return np.asarray(a, np.float32).astype(np.float64)


def _scene(nb, ng, seed):
    rng = np.random.default_rng(seed)
    bones = _f32(rng.random((nb, 3)) * 0.2)
    xyz = _f32(bones[rng.integers(0, nb, ng)] + 0.004 * rng.standard_normal((ng, 3)))
    quat = _f32(rng.standard_normal((ng, 4)))
    return bones, xyz, quat


def _deform(bones, seed, angle=0.15):
    """A smooth non-rigid motion: a rotation about the centre plus a swirl."""
    rng = np.random.default_rng(seed)
    c = bones.mean(0)
    Q = Rotation.from_rotvec(angle * rng.standard_normal(3)).as_matrix()
    x, y, z = (bones - c).T
    swirl = 0.01 * np.stack([np.sin(20 * y), np.sin(20 * z), np.sin(20 * x)], 1)
    return _f32((bones - c) @ Q.T + c + swirl - bones)


def _t(a, dev, dtype=torch.float32):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def _check_knn(idx_g, dist_g, q, r, k, skip):
    idx_o, dist_o = lbs.knn(q, r, k, skip)
    np.testing.assert_allclose(dist_g, dist_o, rtol=1e-5, atol=1e-7)
    # every GPU pick is at the distance the oracle ranks there
    d_pick = np.linalg.norm(q[:, None, :] - r[idx_g], axis=2)
    np.testing.assert_allclose(d_pick, dist_o, rtol=1e-5, atol=1e-7)
    assert np.mean(idx_g == idx_o) > 0.999
    assert len(np.unique(idx_g[0])) == k


@pytest.mark.parametrize("k,skip", [(1, 0), (5, 0), (16, 0), (16, 1), (31, 1), (32, 0)])
def test_knn_matches_oracle(dev, k, skip):
    print(f"[lbs] knn k={k} skip={skip}", flush=True)
    rng = np.random.default_rng(k * 10 + skip)
    r = rng.random((1500, 3)) * 0.2
    q = r if skip else rng.random((3000, 3)) * 0.2
    tq, tr = _t(q, dev), _t(r, dev)
    idx, dist = gs.dynamic_utils._knn(tq, tr, k, skip)
    torch.cuda.synchronize()
    _check_knn(idx.cpu().numpy(), dist.cpu().numpy(), q, r, k, skip)


def test_knn_ragged_tiles_and_edges(dev):
    print("[lbs] knn ragged", flush=True)
    rng = np.random.default_rng(3)
    # reference count not a multiple of the 2048-point LDS tile, queries not a
    # multiple of the 256-thread block
    r = rng.random((4099, 3))
    q = rng.random((777, 3))
    idx, dist = gs.dynamic_utils._knn(_t(q, dev), _t(r, dev), 16, 0)
    _check_knn(idx.cpu().numpy(), dist.cpu().numpy(), q, r, 16, 0)
    # no queries
    idx, dist = gs.dynamic_utils._knn(_t(np.zeros((0, 3)), dev), _t(r, dev), 4, 0)
    assert idx.shape == (0, 4)
    # too few reference points, K too large
    with pytest.raises(ValueError):
        gs.dynamic_utils._knn(_t(q, dev), _t(r[:3], dev), 4, 0)
    with pytest.raises(ValueError):
        gs.get_topk_indices(_t(r, dev), K=32)


def test_mirror_setup_functions(dev):
    print("[lbs] setup functions", flush=True)
    bones, xyz, _ = _scene(800, 2000, 5)
    tb, tx = _t(bones, dev), _t(xyz, dev)
    rel = gs.get_topk_indices(tb, K=16).cpu().numpy()
    rel_o, _ = lbs.knn(bones, bones, 16, skip_first=1)
    assert np.mean(rel == rel_o) > 0.999
    assert not np.any(rel == np.arange(800)[:, None])
    w, idx = gs.knn_weights_sparse(tb, tx, K=16)
    idx = idx.cpu().numpy()
    np.testing.assert_allclose(w.cpu().numpy(), lbs.inv_dist_weights(xyz, bones, idx), rtol=1e-5, atol=1e-7)
    dense = gs.knn_weights(tb, tx, K=16).cpu().numpy()
    np.testing.assert_allclose(dense.sum(1), 1.0, atol=1e-5)
    assert np.all((dense > 0).sum(1) == 16)
    rm = gs.create_relation_matrix(tb[:300], K=5)
    assert rm.dtype == torch.int32 and rm.device.type == "cpu"
    assert torch.all(rm.sum(1) == 5) and torch.all(rm.diagonal() == 0)


@pytest.mark.parametrize("idx_dtype", [torch.int32, torch.int64])
def test_calc_weights_matches_oracle(dev, idx_dtype):
    print(f"[lbs] weights {idx_dtype}", flush=True)
    bones, xyz, _ = _scene(1000, 5000, 7)
    idx_o, _ = lbs.knn(xyz, bones, 16)
    w = gs.calc_weights_vals_from_indices(_t(bones, dev), _t(xyz, dev), _t(idx_o, dev, idx_dtype))
    np.testing.assert_allclose(w.cpu().numpy(), lbs.inv_dist_weights(xyz, bones, idx_o), rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("given_weights,idx_dtype", [(True, torch.int64), (False, torch.int32), (False, torch.int64)])
def test_interpolate_matches_oracle(dev, given_weights, idx_dtype):
    print(f"[lbs] interpolate weights={given_weights} {idx_dtype}", flush=True)
    bones, xyz, quat = _scene(2000, 20000, 11)
    motions = _deform(bones, 12)
    rel, _ = lbs.knn(bones, bones, 16, skip_first=1)
    idx, _ = lbs.knn(xyz, bones, 16)
    w = lbs.inv_dist_weights(xyz, bones, idx)
    out_o, q_o = lbs.interpolate(bones, motions, rel, xyz, quat, w, idx)
    out, q, (w_ret, i_ret) = gs.interpolate_motions_speedup(
        bones=_t(bones, dev), motions=_t(motions, dev), relations=_t(rel, dev, idx_dtype), xyz=_t(xyz, dev),
        quat=_t(quat, dev), weights=_t(w, dev) if given_weights else None, weights_indices=_t(idx, dev, idx_dtype))
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), out_o, atol=POS_TOL)
    np.testing.assert_allclose(q.cpu().numpy(), q_o, atol=QUAT_TOL * np.abs(quat).max())
    assert i_ret is not None and (w_ret is None) == (not given_weights)


def _block_rotations(bones, seed):
    """Motions that turn each of 3x3x3 blocks of the box rigidly about its
    centre, alternating 0.6 and 2.8 rad about axes near x, y, z: the blocks'
    rotations fall in every branch of the quaternion conversion."""
    rng = np.random.default_rng(seed)
    cell = np.minimum((bones / (0.2 / 3)).astype(int), 2)
    g = cell[:, 0] + 3 * cell[:, 1] + 9 * cell[:, 2]
    out = np.empty_like(bones)
    for k in range(27):
        axis = np.eye(3)[k % 3] + 0.1 * rng.standard_normal(3)
        axis /= np.linalg.norm(axis)
        Q = Rotation.from_rotvec(axis * (2.8 if k % 2 else 0.6)).as_matrix()
        c = (np.array([k % 3, (k // 3) % 3, k // 9]) + 0.5) * 0.2 / 3
        m = g == k
        out[m] = (bones[m] - c) @ Q.T + c - bones[m]
    return _f32(out)


def test_default_k_and_no_quaternions_match_oracle(dev):
    """The signatures' default K = 5 (the blend's generic path) and a blend
    without quaternions."""
    print("[lbs] K=5, no quaternions", flush=True)
    bones, xyz, _ = _scene(1200, 8000, 25)
    motions = _deform(bones, 26)
    tb = _t(bones, dev)
    rel = gs.get_topk_indices(tb, K=5)
    w, idx = gs.knn_weights_sparse(tb, _t(xyz, dev), K=5)
    out, rot, _ = gs.interpolate_motions_speedup(bones=tb, motions=_t(motions, dev), relations=rel,
                                                 xyz=_t(xyz, dev), weights=w, weights_indices=idx)
    torch.cuda.synchronize()
    assert rot is None
    rel_np, idx_np = rel.long().cpu().numpy(), idx.long().cpu().numpy()
    out_o, _ = lbs.interpolate(bones, motions, rel_np, xyz, None, w.double().cpu().numpy(), idx_np)
    np.testing.assert_allclose(out.cpu().numpy(), out_o, atol=POS_TOL)


def test_bone_records_match_oracle(dev):
    print("[lbs] bone records", flush=True)
    bones, _, _ = _scene(3000, 1, 13)
    motions = _block_rotations(bones, 14)
    rel, _ = lbs.knn(bones, bones, 16, skip_first=1)
    rec, degen = gs.LBSInterpolator().bones(_t(bones, dev), _t(motions, dev), _t(rel, dev, torch.int32))
    rec = rec.cpu().numpy()
    assert int(degen[0].item()) == 0
    R_o, _ = lbs.bone_rotations(bones, motions, rel)
    R = lbs.quat_to_rot(rec[:, 0:4])
    # the closest rotation moves by ~eps / gap under a perturbation of F, with
    # gap = (s2 + s3) / s1, or (s2 - s3) / s1 when det F < 0 flips the third
    # direction; bones on block faces mix two rotations and can come close
    F = lbs.bone_covariance(bones, motions, rel)
    sv = np.linalg.svd(F, compute_uv=False)
    gap = np.where(np.linalg.det(F) < 0, sv[:, 1] - sv[:, 2], sv[:, 1] + sv[:, 2]) / sv[:, 0]
    good = gap > 1e-2
    assert good.mean() > 0.99
    np.testing.assert_allclose(R[good], R_o[good], atol=1e-5)
    np.testing.assert_allclose(R[~good], R_o[~good], atol=1e-3)
    np.testing.assert_allclose(rec[:, 4:7], bones.astype(np.float32), atol=0)
    np.testing.assert_allclose(rec[:, 8:11], motions.astype(np.float32), atol=0)
    np.testing.assert_allclose(np.linalg.norm(rec[:, 0:4], axis=1), 1.0, atol=1e-6)
    q_o = lbs.normalize(lbs.quat_from_rot(R_o))
    # the kornia branch (and with it the sign) is decided by float32
    # comparisons of R's diagonal: compare where the decision has a margin
    tr = np.trace(R_o, axis1=1, axis2=2)
    d = np.stack([R_o[:, 0, 0], R_o[:, 1, 1], R_o[:, 2, 2]], 1)
    ds = np.sort(d, 1)
    clear = (np.abs(tr) > 1e-4) & (ds[:, 2] - ds[:, 1] > 1e-4) & good
    assert clear.mean() > 0.99
    np.testing.assert_allclose(rec[clear, 0:4], q_o[clear], atol=QUAT_TOL)
    big = np.where(tr > 0, 0, 1 + np.argmax(d, 1))
    for b in range(4):
        assert np.any(big[clear] == b)


def _blend_vs_oracle(dev, bones, motions, rel, seed, n=500):
    rng = np.random.default_rng(seed)
    xyz = _f32(bones[rng.integers(0, len(bones), n)] + 0.002 * rng.standard_normal((n, 3)))
    quat = _f32(rng.standard_normal((n, 4)))
    idx, _ = lbs.knn(xyz, bones, 4)
    out_o, q_o = lbs.interpolate(bones, motions, rel, xyz, quat, None, idx)
    interp = gs.LBSInterpolator()
    out, q = interp(_t(bones, dev), _t(motions, dev), _t(rel, dev), _t(xyz, dev), quat=_t(quat, dev),
                    weights_indices=_t(idx, dev))
    np.testing.assert_allclose(out.cpu().numpy(), out_o, atol=POS_TOL)
    np.testing.assert_allclose(q.cpu().numpy(), q_o, atol=QUAT_TOL * np.abs(quat).max())
    return interp._degen.cpu().numpy()


def test_collinear_bones_use_their_rank1_rotations(dev):
    # every bone of rank 1: each keeps its own rank-1 rotation (here a turn by
    # atan(0.1) about z for all)
    print("[lbs] rank 1, all bones", flush=True)
    nb = 64
    bones = _f32(np.stack([np.linspace(0, 0.2, nb), np.zeros(nb), np.zeros(nb)], 1))
    motions = np.zeros_like(bones)
    motions[:, 1] = 0.1 * bones[:, 0]
    motions = _f32(motions)
    rel, _ = lbs.knn(bones, bones, 4, skip_first=1)
    st = _blend_vs_oracle(dev, bones, motions, rel, 15)
    assert list(st[:6]) == [nb, nb, 0, nb - 1, -1, 0]


def _generic(nb, seed):
    rng = np.random.default_rng(seed)
    bones = _f32(rng.uniform(0, 0.2, (nb, 3)))
    rel, _ = lbs.knn(bones, bones, 4, skip_first=1)
    motions = _f32(0.002 * rng.standard_normal((nb, 3)))
    return bones, motions, rel.astype(np.int64)


def test_one_rank1_bone_is_broadcast_two_give_identity(dev):
    print("[lbs] rank 1, one and two bones", flush=True)
    bones, motions, rel = _generic(200, 16)
    bones[1] = bones[0] + _f32([0.0, 0.01, 0.0])
    bones[2] = bones[0] + _f32([0.0, -0.006, 0.0])
    motions[1] = motions[0] + _f32([0.0, 0.01, 0.0008])
    motions[2] = motions[0] + _f32([0.0, -0.006, -0.00048])
    rel[0] = [1, 2, 1, 2]
    st = _blend_vs_oracle(dev, bones, motions, rel, 17)
    assert list(st[:6]) == [1, 1, 199, 0, 199, 0]
    bones[4] = bones[3] + _f32([0.008, 0.0, 0.0])
    rel[3] = [4, 4, 4, 4]
    st = _blend_vs_oracle(dev, bones, motions, rel, 18)
    assert list(st[:3]) == [2, 2, 198]


def test_rank0_bones_with_one_full_bone_broadcast(dev):
    print("[lbs] rank 0 with one full bone", flush=True)
    bones, motions, rel = _generic(100, 19)
    rel[1:] = np.arange(1, 100)[:, None]
    st = _blend_vs_oracle(dev, bones, motions, rel, 20)
    assert list(st[:6]) == [99, 0, 1, -1, 0, 0]


def test_reference_frame_loop_matches_oracle(dev):
    """Three frames of the interactive loop (scripts/interactive_viser.py:
    259-280): relations and indices from the first frame, weights recomputed
    every frame from the previous particle positions."""
    print("[lbs] frame loop", flush=True)
    bones0, xyz0, quat0 = _scene(1500, 10000, 17)
    frames = [bones0]
    for f in range(3):
        frames.append(frames[-1] + _deform(frames[-1], 20 + f, angle=0.05))
    frames = [_f32(f) for f in frames]
    # the reference's call sequence on the mirror
    gs_pos, gs_rot = _t(xyz0, dev), _t(quat0, dev)
    relations = weights = weights_indices = None
    for f in range(1, 4):
        prev_x, x = _t(frames[f - 1], dev), _t(frames[f], dev)
        if relations is None:
            relations = gs.get_topk_indices(prev_x, K=16)
        if weights is None:
            weights, weights_indices = gs.knn_weights_sparse(prev_x, gs_pos, K=16)
        weights = gs.calc_weights_vals_from_indices(prev_x, gs_pos, weights_indices)
        gs_pos, gs_rot, _ = gs.interpolate_motions_speedup(bones=prev_x, motions=x - prev_x, relations=relations,
                                                           weights=weights, weights_indices=weights_indices,
                                                           xyz=gs_pos, quat=gs_rot)
    torch.cuda.synchronize()
    rel_o, _ = lbs.knn(frames[0], frames[0], 16, skip_first=1)
    idx_o, _ = lbs.knn(xyz0, frames[0], 16)
    assert np.mean(relations.cpu().numpy() == rel_o) > 0.999
    assert np.mean(weights_indices.cpu().numpy() == idx_o) > 0.999
    # the oracle's frames on the GPU's neighbour lists (a float32 near-tie at
    # the 16th neighbour may pick the other of two equally near bones)
    rel_o = relations.long().cpu().numpy()
    idx_o = weights_indices.long().cpu().numpy()
    x_o, q_o = xyz0, quat0
    for f in range(1, 4):
        w_o = lbs.inv_dist_weights(x_o, frames[f - 1], idx_o)
        x_o, q_o = lbs.interpolate(frames[f - 1], frames[f] - frames[f - 1], rel_o, x_o, q_o, w_o, idx_o)
    np.testing.assert_allclose(gs_pos.cpu().numpy(), x_o, atol=POS_TOL)
    np.testing.assert_allclose(gs_rot.cpu().numpy(), q_o, atol=QUAT_TOL * np.abs(quat0).max())


def test_interactive_loop_on_simulated_frames(dev):
    """Both halves of the interactive loop (scripts/interactive_viser.py:
    240-291) together: the simulator mirror advances a spinning jelly
    ellipsoid frame by frame, and its particles drive the Gaussians through
    the reference's LBS call sequence; the oracle replays the same float32
    frames on the same neighbour lists."""
    print("[lbs] loop on simulated frames", flush=True)
    from empm_amd import scenes
    from empm_amd.diff_simulator import WarpMPMWrapper
    from empm_amd.diff_simulator.warp_mpm.warp_mpm_wrapper import material_gallery
    sc = scenes.small_scene(n_particles=800, n_grid=16, material=0, spin=6.0)
    mpm = WarpMPMWrapper(grid_size=sc.n_grid, grid_lim=1.0, frame_dt=sc.dt * 40, sim_dt=sc.dt, device=dev)
    mpm.material_params = dict(material_gallery["jelly"])
    mpm.add_particles(sc.x, sc.params["v0"], sc.vol.astype(np.float32), sc.E.astype(np.float32),
                      sc.nu.astype(np.float32), sc.rho.astype(np.float32))
    for c in sc.colliders:
        mpm.surface_colliders.append([c["point"], c["normal"], "sticky", 0.0])
    mpm.initialize()
    rng = np.random.default_rng(29)
    xyz0 = _f32(sc.x[rng.integers(0, sc.n, 4000)] + 0.01 * rng.standard_normal((4000, 3)))
    quat0 = _f32(rng.standard_normal((4000, 4)))
    gs_pos, gs_rot = _t(xyz0, dev), _t(quat0, dev)
    relations = weights = weights_indices = prev_x = None
    frames = []
    for f in range(5):
        world_pts = mpm.advance_to(n_frames=1)[-1]
        frames.append(world_pts.double().cpu().numpy())
        if prev_x is not None:
            if relations is None:
                relations = gs.get_topk_indices(prev_x, K=16)
            if weights is None:
                weights, weights_indices = gs.knn_weights_sparse(prev_x, gs_pos, K=16)
            weights = gs.calc_weights_vals_from_indices(prev_x, gs_pos, weights_indices)
            gs_pos, gs_rot, _ = gs.interpolate_motions_speedup(bones=prev_x, motions=world_pts - prev_x,
                                                               relations=relations, weights=weights,
                                                               weights_indices=weights_indices, xyz=gs_pos,
                                                               quat=gs_rot)
        prev_x = world_pts.clone()
    torch.cuda.synchronize()
    # the particles moved: ~1 mm a frame, rotating about z
    step = np.linalg.norm(frames[-1] - frames[0], axis=1)
    assert step.mean() > 2e-3
    rel_o, _ = lbs.knn(frames[0], frames[0], 16, skip_first=1)
    idx_o, _ = lbs.knn(xyz0, frames[0], 16)
    assert np.mean(relations.cpu().numpy() == rel_o) > 0.999
    assert np.mean(weights_indices.cpu().numpy() == idx_o) > 0.999
    rel_o = relations.long().cpu().numpy()
    idx_o = weights_indices.long().cpu().numpy()
    x_o, q_o = xyz0, quat0
    for f in range(1, 5):
        w_o = lbs.inv_dist_weights(x_o, frames[f - 1], idx_o)
        x_o, q_o = lbs.interpolate(frames[f - 1], frames[f] - frames[f - 1], rel_o, x_o, q_o, w_o, idx_o)
    np.testing.assert_allclose(gs_pos.cpu().numpy(), x_o, atol=POS_TOL)
    np.testing.assert_allclose(gs_rot.cpu().numpy(), q_o, atol=QUAT_TOL * np.abs(quat0).max())
    # the Gaussians followed the body: their mean moved as the particles' did
    np.testing.assert_allclose(x_o.mean(0) - xyz0.mean(0), frames[-1].mean(0) - frames[0].mean(0), atol=2e-3)


def test_rigid_motion_full_size(dev):
    """50k bones and 1M Gaussians, the interpolation's benchmark size: a
    rigid motion moves every Gaussian rigidly whatever its weights."""
    print("[lbs] rigid full size", flush=True)
    bones, xyz, quat = _scene(50_000, 1_000_000, 19)
    c = np.array([0.1, 0.1, 0.1])
    Q = Rotation.from_rotvec([0.4, -0.3, 0.5]).as_matrix()
    t = np.array([0.01, -0.02, 0.005])
    moved = (bones - c) @ Q.T + c + t
    tb = _t(bones, dev)
    rel = gs.get_topk_indices(tb, K=16)
    w, idx = gs.knn_weights_sparse(tb, _t(xyz, dev), K=16)
    out, q, _ = gs.interpolate_motions_speedup(bones=tb, motions=_t(moved, dev) - tb, relations=rel,
                                               xyz=_t(xyz, dev), quat=_t(quat, dev), weights=w, weights_indices=idx)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), (xyz - c) @ Q.T + c + t, atol=POS_TOL)
    qq = lbs.quat_from_rot(Q)
    np.testing.assert_allclose(q.cpu().numpy(), lbs.quat_mul(np.broadcast_to(qq, quat.shape), quat),
                               atol=QUAT_TOL * np.abs(quat).max())


def test_quat_none_and_argument_errors(dev):
    print("[lbs] errors", flush=True)
    bones, xyz, _ = _scene(200, 300, 23)
    tb, tx = _t(bones, dev), _t(xyz, dev)
    rel = gs.get_topk_indices(tb, K=8)
    w, idx = gs.knn_weights_sparse(tb, tx, K=8)
    sentinel = object()
    out, rot, _ = gs.interpolate_motions_speedup(tb, torch.zeros_like(tb), rel, tx, rot=sentinel, weights=w,
                                                 weights_indices=idx)
    assert rot is sentinel
    torch.testing.assert_close(out, tx, atol=1e-6, rtol=0)      # zero motion: identity
    with pytest.raises(ValueError):
        gs.interpolate_motions_speedup(tb, torch.zeros_like(tb), rel, tx, weights=w)
    with pytest.raises(ValueError):
        gs.interpolate_motions_speedup(tb, torch.zeros_like(tb), rel, tx, weights=w[:, :4], weights_indices=idx)
    with pytest.raises(ValueError):
        gs.interpolate_motions_speedup(tb, torch.zeros_like(tb), rel, tx, quat=torch.zeros(5, 4, device=dev),
                                       weights=w, weights_indices=idx)
    # the C-ABI rejects what the mirror would not send
    from empm_amd import _lib
    with pytest.raises(EmpmError):
        _lib.check(_lib.load().empm_knn(4, tx.data_ptr(), 200, tb.data_ptr(), 40, 0, idx.data_ptr(), w.data_ptr(),
                                        None))
    with pytest.raises(EmpmError):
        _lib.check(_lib.load().empm_lbs_points(300, tx.data_ptr(), None, 8, idx.data_ptr(), 2, None, 200, None,
                                               None, tx.data_ptr(), None, None))
