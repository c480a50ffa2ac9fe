"""The training step's loss on the simulated frames
(phys_sim/engine/trainer_warp_mpm.py:727-766): per frame, the Chamfer
distance between the predicted and observed points that are visible in that
frame, plus the tracking term |pred - target| / n over the points with a
valid motion; both summed over the frames and scaled by 1e3.

`chamfer_distance` takes the arguments the trainer passes to
pytorch3d.loss.chamfer_distance (a dependency of the reference, not used
here) and computes what that call computes: squared Euclidean distance to
the nearest point of the other cloud, averaged over each cloud's points, the
two directions summed, averaged over the batch.  The nearest-neighbour
search is the HIP kernel behind `empm_nn_sqdist`; the gradient is written
out (nearest-neighbour assignments are piecewise constant):

    dL/dx_i = 2 (x_i - y_nn(i)) / N + sum_{j : nn(j) = i} 2 (x_i - y_j) / M

and symmetrically for y.

The frame gradients flow into the simulation's reverse sweep as per-frame
seeds (each frame's dL/dx enters when the sweep reaches that frame).
"""
from __future__ import annotations

import torch

from . import _lib
from .runtime import _ptr, stream_handle


def nn_sqdist(a: torch.Tensor, b: torch.Tensor):
    """(dist, idx): for every row of a (n, 3), the squared distance to its
    nearest row of b (m, 3) and that row's index (the smallest on ties)."""
    if a.device.type != "cuda" or b.device.type != "cuda":
        raise RuntimeError("nn_sqdist runs on the GPU; got tensors on " + f"{a.device} and {b.device}")
    a = a.detach().float().contiguous()
    b = b.detach().float().contiguous()
    if a.dim() != 2 or a.shape[1] != 3 or b.dim() != 2 or b.shape[1] != 3:
        raise ValueError(f"point clouds must be (n, 3); got {tuple(a.shape)} and {tuple(b.shape)}")
    n, m = int(a.shape[0]), int(b.shape[0])
    if n > 0 and m == 0:
        raise ValueError("nearest neighbour in an empty cloud")
    dist = torch.empty(n, dtype=torch.float32, device=a.device)
    idx = torch.empty(n, dtype=torch.int32, device=a.device)
    lib = _lib.load()
    ws_bytes = int(lib.empm_nn_sqdist_workspace(n, m))
    ws = torch.empty(max(ws_bytes // 4, 1), dtype=torch.int32, device=a.device)
    _lib.check(lib.empm_nn_sqdist(n, _ptr(a), m, _ptr(b), _ptr(dist), _ptr(idx), ws.data_ptr(), ws_bytes,
                                  stream_handle(a.device)))
    return dist, idx.long()


def nn_sqdist_segments(a: torch.Tensor, b: torch.Tensor, a_off: torch.Tensor, b_off: torch.Tensor, max_a: int,
                       max_b: int):
    """nn_sqdist within segments in one launch: rows [a_off[g], a_off[g+1])
    of a against rows [b_off[g], b_off[g+1]) of b; (dist, idx) with idx
    global into b.  a_off, b_off: int32 device tensors of n_seg + 1 offsets;
    max_a, max_b: the longest segments."""
    if a.device.type != "cuda" or b.device.type != "cuda":
        raise RuntimeError("nn_sqdist_segments runs on the GPU; got tensors on " + f"{a.device} and {b.device}")
    a = a.detach().float().contiguous()
    b = b.detach().float().contiguous()
    a_off = a_off.to(device=a.device, dtype=torch.int32).contiguous()
    b_off = b_off.to(device=a.device, dtype=torch.int32).contiguous()
    n_seg, n = int(a_off.shape[0]) - 1, int(a.shape[0])
    dist = torch.empty(n, dtype=torch.float32, device=a.device)
    idx = torch.empty(n, dtype=torch.int32, device=a.device)
    lib = _lib.load()
    ws_bytes = int(lib.empm_nn_sqdist_segments_workspace(n_seg, n, max_a, max_b))
    ws = torch.empty(max(ws_bytes // 4, 1), dtype=torch.int32, device=a.device)
    _lib.check(lib.empm_nn_sqdist_segments(n_seg, a_off.data_ptr(), b_off.data_ptr(), n, max_a, max_b,
                                           a.data_ptr(), b.data_ptr(), dist.data_ptr(), idx.data_ptr(),
                                           ws.data_ptr(), ws_bytes, stream_handle(a.device)))
    return dist, idx.long()


class _SegmentChamfer(torch.autograd.Function):
    """Sum over segments of the Chamfer distance of x_g against y_g (equal
    counts c_g per segment): sum_i w_i (d(x_i, y) + d(y_i, x)) with
    w_i = 1 / c_g of the row's segment."""

    @staticmethod
    def forward(ctx, x, y, off, w, max_len):
        dx, ix = nn_sqdist_segments(x, y, off, off, max_len, max_len)
        dy, iy = nn_sqdist_segments(y, x, off, off, max_len, max_len)
        ctx.save_for_backward(x, y, ix, iy, w)
        return (w * (dx + dy)).sum()

    @staticmethod
    def backward(ctx, g):
        x, y, ix, iy, w = ctx.saved_tensors
        rx = x.float() - y.float()[ix]   # x_i - y_nn(i)
        ry = y.float() - x.float()[iy]   # y_j - x_nn(j)
        gx = (2.0 * w)[:, None] * rx
        gx = gx.index_add(0, iy, (-2.0 * w)[:, None] * ry)
        gy = (2.0 * w)[:, None] * ry
        gy = gy.index_add(0, ix, (-2.0 * w)[:, None] * rx)
        return (g * gx).to(x.dtype), (g * gy).to(y.dtype), None, None, None


class _Chamfer(torch.autograd.Function):
    """Chamfer distance of one pair of clouds x (N, 3), y (M, 3)."""

    @staticmethod
    def forward(ctx, x, y):
        dx, ix = nn_sqdist(x, y)
        dy, iy = nn_sqdist(y, x)
        ctx.save_for_backward(x, y, ix, iy)
        return dx.mean() + dy.mean()

    @staticmethod
    def backward(ctx, g):
        x, y, ix, iy = ctx.saved_tensors
        N, M = x.shape[0], y.shape[0]
        gx = gy = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            rx = x.float() - y.float()[ix]   # x_i - y_nn(i)
            ry = y.float() - x.float()[iy]   # y_j - x_nn(j)
            gx = (2.0 / N) * rx
            gx = gx.index_add(0, iy, (-2.0 / M) * ry)
            gy = (2.0 / M) * ry
            gy = gy.index_add(0, ix, (-2.0 / N) * rx)
            gx = (g * gx).to(x.dtype)
            gy = (g * gy).to(y.dtype)
        return (gx if ctx.needs_input_grad[0] else None), (gy if ctx.needs_input_grad[1] else None)


def chamfer_distance(x, y, x_lengths=None, y_lengths=None, x_normals=None, y_normals=None, weights=None,
                     batch_reduction: str | None = "mean", point_reduction: str | None = "mean", norm: int = 2,
                     single_directional: bool = False, abs_cosine: bool = True):
    """(loss, None) for clouds x (B, N, 3) and y (B, M, 3) (or (N, 3) and
    (M, 3)), as pytorch3d.loss.chamfer_distance with its defaults: squared
    Euclidean nearest distances, mean over points, both directions, mean
    over the batch.  Variable lengths, normals, weights and other reductions
    are not supported."""
    if x_lengths is not None or y_lengths is not None or x_normals is not None or y_normals is not None:
        raise NotImplementedError("chamfer_distance: per-cloud lengths and normals are not supported")
    if weights is not None or batch_reduction != "mean" or point_reduction != "mean" or norm != 2:
        raise NotImplementedError("chamfer_distance: only the default weights, mean reductions and norm=2")
    if single_directional:
        raise NotImplementedError("chamfer_distance: single_directional is not supported")
    if x.dim() == 2:
        x, y = x.unsqueeze(0), y.unsqueeze(0)
    if x.dim() != 3 or y.dim() != 3 or x.shape[0] != y.shape[0] or x.shape[-1] != 3 or y.shape[-1] != 3:
        raise ValueError(f"chamfer_distance: expected (B, N, 3) and (B, M, 3), got {tuple(x.shape)}, "
                         f"{tuple(y.shape)}")
    loss = sum(_Chamfer.apply(x[b], y[b]) for b in range(x.shape[0])) / x.shape[0]
    return loss, None


def tracking_loss(pred: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """|pred - target| (Frobenius norm) / number of points
    (trainer_warp_mpm.py:755-762)."""
    return (pred - target).norm() / pred.shape[0]


def frame_loss(points, target_points, visibilities, motions_valid, n_track: int | None = None):
    """The trainer's loss_func (trainer_warp_mpm.py:727-766) over frames:
    points and target_points (T, N, 3); visibilities and motions_valid (T, N)
    boolean masks; the first n_track points are the tracked ones.  Returns
    1e3 * (sum of per-frame Chamfer distances of the visible points + sum of
    per-frame tracking terms of the points with a valid motion).  The frames
    are evaluated together: one segmented nearest-neighbour launch per
    direction, one host synchronisation for the per-frame counts."""
    if n_track is not None:
        points = points[:, :n_track, :]
    T, N = points.shape[0], points.shape[1]
    if T == 0:
        return points.sum() * 0.0
    vis = visibilities.to(device=points.device, dtype=torch.bool)
    valid = motions_valid.to(device=points.device, dtype=torch.bool)
    if tuple(vis.shape) != (T, N) or tuple(valid.shape) != (T, N) or target_points.shape[0] != T:
        raise ValueError(f"frame_loss: points {tuple(points.shape)}, targets {tuple(target_points.shape)}, masks "
                         f"{tuple(vis.shape)} and {tuple(valid.shape)} do not match")
    # all frames at once: the visible rows of every frame, frame after frame
    counts = vis.sum(1)
    cnt = counts.tolist()
    tf, ti = torch.nonzero(vis, as_tuple=True)
    x = points[tf, ti]
    y = target_points[tf, ti]
    off = torch.zeros(T + 1, dtype=torch.int32, device=points.device)
    off[1:] = torch.cumsum(counts, 0)
    w = torch.repeat_interleave(1.0 / counts.clamp(min=1).float(), counts, output_size=int(sum(cnt)))
    chamfer = _SegmentChamfer.apply(x, y, off, w, max(cnt))
    if min(cnt) == 0:
        chamfer = chamfer + float("nan")   # an empty frame: the reference's mean over no points
    # tracking: |pred - target| over the valid rows of each frame / their number
    diff = (points - target_points[:, :N]) * valid[..., None]
    tracking = (torch.linalg.vector_norm(diff, dim=(1, 2)) / valid.sum(1)).sum()
    return chamfer * 1e3 + tracking * 1e3
