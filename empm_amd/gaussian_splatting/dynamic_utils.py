"""Motion interpolation of Gaussians on the simulated particles (linear blend
skinning), with the functions, argument order and return values of the
reference's third_party/gaussian_splatting/dynamic_utils.py, computed by the
HIP kernels behind empm_knn / empm_lbs_* (include/empm.h).

The per-frame call sequence of the reference's interactive loop
(scripts/interactive_viser.py:259-280) and rendering loop
(phys_sim/engine/trainer_warp_mpm.py:1630-1656) runs unchanged:

    relations = get_topk_indices(prev_x, K=16)                 # first frame
    weights, weights_indices = knn_weights_sparse(prev_x, gs_pos, K=16)
    weights = calc_weights_vals_from_indices(prev_x, gs_pos, weights_indices)
    gs_pos, gs_rot, _ = interpolate_motions_speedup(bones=prev_x, motions=x - prev_x,
        relations=relations, weights=weights, weights_indices=weights_indices,
        xyz=gs_pos, quat=gs_rot)

Differences a caller can see:
- index tensors come back as int32 (the kernels also take int64, torch.topk's
  dtype, without a copy);
- weights=None in interpolate_motions_speedup computes the inverse-distance
  weights of calc_weights_vals_from_indices(bones, xyz, weights_indices) inside
  the blend kernel (the two calls of the loop in one pass);
- nearest-neighbour searches run on the GPU (knn_weights_sparse computes its
  distance matrix on the CPU in the reference);
- no diagnostics are printed (and no debugger is entered on a bad
  determinant);
- the rank-1 rotation (dynamic_utils.py:373-397) turns e_x to the direction
  of the bone's F with that direction's largest-magnitude component
  positive, where the reference takes torch.svd's sign; which rotations are
  kept and broadcast (:331-405) is the reference's.
"""
from __future__ import annotations

import torch

from .. import _lib
from ..runtime import stream_handle


def quat2mat(q):
    """Rotation matrices (n, 3, 3) of quaternions (n, 4) in (w, x, y, z)
    order, normalised first (dynamic_utils.py:5-22).  Elementwise torch on
    the quaternions' device."""
    q = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = q.unbind(1)
    return torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1).reshape(-1, 3, 3)


def mat2quat(rot):
    """Quaternions (n, 4), (w, x, y, z), of rotation matrices (n, 3, 3) by
    the reference's own branch rule (dynamic_utils.py:25-64): the trace
    branch whenever the trace (clamped at -1) exceeds -1, otherwise the
    largest diagonal entry, ties going to the earlier axis for x, then y.
    In the three diagonal branches the reference sets the branch's own
    component to 0.25 / sqrt(1 + 2 r_aa - tr) instead of
    0.5 * sqrt(1 + 2 r_aa - tr) (:38-64; the function is its fallback when
    kornia is missing); here every branch returns the quaternion of the
    matrix."""
    m = rot
    d0, d1, d2 = m[:, 0, 0], m[:, 1, 1], m[:, 2, 2]
    t = torch.clamp(d0 + d1 + d2, min=-1)
    q = torch.zeros((m.shape[0], 4), dtype=m.dtype, device=m.device)
    b0 = t > -1
    b1 = ~b0 & (d0 >= d1) & (d0 >= d2)
    b2 = ~b0 & (d1 >= d2) & (d1 > d0)
    b3 = ~b0 & (d2 > d0) & (d2 > d1)
    if bool(b0.any()):
        r = torch.sqrt(t[b0] + 1)
        h = 0.5 / r
        mm = m[b0]
        q[b0] = torch.stack([0.5 * r, (mm[:, 2, 1] - mm[:, 1, 2]) * h, (mm[:, 0, 2] - mm[:, 2, 0]) * h,
                             (mm[:, 1, 0] - mm[:, 0, 1]) * h], 1)
    for sel, a, b, c in ((b1, 0, 1, 2), (b2, 1, 2, 0), (b3, 2, 0, 1)):
        if bool(sel.any()):
            mm = m[sel]
            h = 0.5 / torch.sqrt(1 + mm[:, a, a] - mm[:, b, b] - mm[:, c, c])
            out = torch.empty((mm.shape[0], 4), dtype=m.dtype, device=m.device)
            out[:, 0] = (mm[:, c, b] - mm[:, b, c]) * h
            out[:, 1 + a] = 0.25 / h
            out[:, 1 + b] = (mm[:, b, a] + mm[:, a, b]) * h
            out[:, 1 + c] = (mm[:, c, a] + mm[:, a, c]) * h
            q[sel] = out
    return q


def _dev(t: torch.Tensor, what: str) -> None:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise RuntimeError(f"{what} must be a tensor on a GPU (the interpolation runs only on the GPU)")


def _points(t: torch.Tensor, what: str) -> torch.Tensor:
    _dev(t, what)
    if t.dim() != 2 or t.shape[1] != 3:
        raise ValueError(f"{what} must be (n, 3); got {tuple(t.shape)}")
    return t.detach().to(torch.float32).contiguous()


def _aligned(t: torch.Tensor) -> torch.Tensor:
    t = t.contiguous()
    return t if t.data_ptr() % 16 == 0 else t.clone()


def _index(t: torch.Tensor, what: str, rows: int) -> tuple[torch.Tensor, int]:
    _dev(t, what)
    if t.dim() != 2 or t.shape[0] != rows:
        raise ValueError(f"{what} must be ({rows}, k); got {tuple(t.shape)}")
    if t.dtype not in (torch.int32, torch.int64):
        t = t.long()
    return t.contiguous(), t.element_size()


def _knn(q: torch.Tensor, r: torch.Tensor, k: int, skip_first: int):
    n, m = q.shape[0], r.shape[0]
    if not 1 <= k + skip_first <= 32:
        raise ValueError(f"K={k} outside the supported 1..{32 - skip_first}")
    if m < k + skip_first:
        raise ValueError(f"{m} reference points for {k + skip_first} neighbours")
    idx = torch.empty((n, k), dtype=torch.int32, device=q.device)
    dist = torch.empty((n, k), dtype=torch.float32, device=q.device)
    _lib.check(_lib.load().empm_knn(n, q.data_ptr(), m, r.data_ptr(), k, skip_first, idx.data_ptr(),
                                    dist.data_ptr(), stream_handle(q.device)))
    return idx, dist


def get_topk_indices(points, K=5):
    """Indices (N, K) of the K nearest other points of every point
    (dynamic_utils.py:248-265: the K + 1 nearest, the first dropped)."""
    p = _points(points, "points")
    return _knn(p, p, K, 1)[0]


def create_relation_matrix(points, K=5):
    """N x N int matrix with 1 at each point's K nearest other points
    (dynamic_utils.py:219-245); on the CPU, as the reference returns it."""
    idx = get_topk_indices(points, K).long().cpu()
    N = idx.shape[0]
    rel = torch.zeros((N, N), dtype=torch.int)
    rel[torch.arange(N).unsqueeze(1).expand(-1, K), idx] = 1
    return rel


def knn_weights_sparse(bones, pts, K=5):
    """(weights (n, K), indices (n, K)): the K nearest bones of every point
    and their normalised inverse-distance weights (dynamic_utils.py:295-302)."""
    b = _points(bones, "bones")
    p = _points(pts, "pts")
    idx, dist = _knn(p, b, K, 0)
    w = 1.0 / (dist + 1e-6)
    return w / w.sum(dim=-1, keepdim=True), idx


def knn_weights(bones, pts, K=5):
    """Dense (n_pts, n_bones) weights, zero except at each point's K nearest
    bones (dynamic_utils.py:268-279)."""
    w, idx = knn_weights_sparse(bones, pts, K)
    out = torch.zeros((pts.shape[0], bones.shape[0]), dtype=torch.float32, device=pts.device)
    out.scatter_(1, idx.long(), w)
    return out


def calc_weights_vals_from_indices(bones, pts, indices):
    """Normalised inverse-distance weights of the given bones per point
    (dynamic_utils.py:282-292)."""
    b = _points(bones, "bones")
    p = _points(pts, "pts")
    idx, ib = _index(indices, "indices", p.shape[0])
    k = idx.shape[1]
    w = torch.empty((p.shape[0], k), dtype=torch.float32, device=p.device)
    if p.shape[0] and b.shape[0] == 0:
        raise ValueError("no bones")
    _lib.check(_lib.load().empm_lbs_weights(p.shape[0], p.data_ptr(), k, idx.data_ptr(), ib, b.shape[0],
                                            b.data_ptr(), w.data_ptr(), stream_handle(p.device)))
    return w


class LBSInterpolator:
    """Per-frame interpolation with persistent device buffers: the bone
    records (64 bytes per bone: unit quaternion, position, motion) and the
    degenerate-bone counter are reused across frames."""

    def __init__(self, device=None):
        self.device = device
        self._rec = None
        self._degen = None

    def _buffers(self, n_bones: int, device):
        if self._rec is None or self._rec.shape[0] < n_bones or self._rec.device != device:
            self._rec = torch.empty((max(n_bones, 1), 16), dtype=torch.float32, device=device)
            if self._rec.data_ptr() % 64:
                raise RuntimeError("bone records must be 64-byte aligned")
            self._degen = torch.empty(8, dtype=torch.int32, device=device)  # the rank state
        return self._rec, self._degen

    def bones(self, bones, motions, relations):
        """Bone records of one frame; returns (records, rank state: 8 device
        ints, [0] the number of bones of rank < 2)."""
        b = _points(bones, "bones")
        m = _points(motions, "motions")
        if m.shape != b.shape:
            raise ValueError(f"motions {tuple(m.shape)} do not match bones {tuple(b.shape)}")
        rel, rb = _index(relations, "relations", b.shape[0])
        rec, degen = self._buffers(b.shape[0], b.device)
        _lib.check(_lib.load().empm_lbs_bones(b.shape[0], b.data_ptr(), m.data_ptr(), rel.shape[1], rel.data_ptr(),
                                              rb, rec.data_ptr(), degen.data_ptr(), stream_handle(b.device)))
        return rec, degen

    def __call__(self, bones, motions, relations, xyz, quat=None, weights=None, weights_indices=None,
                 xyz_out=None, quat_out=None):
        """(xyz', quat') of one frame; quat' is None when quat is None."""
        x = _points(xyz, "xyz")
        if weights_indices is None:
            raise ValueError("weights_indices is required (the K nearest bones of every point)")
        idx, ib = _index(weights_indices, "weights_indices", x.shape[0])
        k = idx.shape[1]
        if not 1 <= k <= 32:
            raise ValueError(f"K={k} outside 1..32")
        w = None
        if weights is not None:
            _dev(weights, "weights")
            if tuple(weights.shape) != tuple(idx.shape):
                raise ValueError(f"weights {tuple(weights.shape)} do not match indices {tuple(idx.shape)}")
            w = weights.detach().to(torch.float32).contiguous()
        q = None
        if quat is not None:
            _dev(quat, "quat")
            if tuple(quat.shape) != (x.shape[0], 4):
                raise ValueError(f"quat must be ({x.shape[0]}, 4); got {tuple(quat.shape)}")
            q = _aligned(quat.detach().to(torch.float32))
        rec, degen = self.bones(bones, motions, relations)
        n_bones = bones.shape[0]
        if x.shape[0] and n_bones == 0:
            raise ValueError("no bones")
        if xyz_out is None:
            xyz_out = torch.empty_like(x)
        if q is not None and quat_out is None:
            quat_out = torch.empty_like(q)
        _lib.check(_lib.load().empm_lbs_points(
            x.shape[0], x.data_ptr(), q.data_ptr() if q is not None else None, k, idx.data_ptr(), ib,
            w.data_ptr() if w is not None else None, n_bones, rec.data_ptr(), degen.data_ptr(), xyz_out.data_ptr(),
            quat_out.data_ptr() if q is not None else None, stream_handle(x.device)))
        return xyz_out, (quat_out if q is not None else None)


def interpolate_motions_speedup(bones, motions, relations, xyz, rot=None, quat=None, weights=None,
                                weights_indices=None, device="cuda", step="n/a"):
    """(xyz_transformed (n, 3), rot, (weights, weights_indices)) as
    dynamic_utils.py:305-493: rot is the blended bone quaternion times quat
    when quat is given, else the rot argument unchanged."""
    out, q = LBSInterpolator()(bones, motions, relations, xyz, quat=quat, weights=weights,
                               weights_indices=weights_indices)
    return out, (q if quat is not None else rot), (weights, weights_indices)
