"""torch.autograd wrapper of the photometric plane cost.

``ncc_cost(est, depth, ab, views)`` returns the per-pixel cost (mean over the
selected source views of 1 - NCC, the stride-1 bilateral NCC of the
PatchMatch cost on the plane-induced local affine warp) of the planes
given by ``depth`` (H, W) and the normal slopes ``ab`` (H, W, 2) of the
camera-facing normal (a, b, -1).  The forward pass runs the HIP kernel,
which also produces d cost / d(depth, a, b) for every pixel; pixels are
independent, so the backward pass scales that gradient by the incoming one.
Tensors live on the estimator's device (cuda:0 by default); the estimator
must have its views uploaded.
"""
from __future__ import annotations

import torch


class NCCCost(torch.autograd.Function):
    @staticmethod
    def forward(ctx, est, depth, ab, views):
        if not (depth.is_cuda and ab.is_cuda and views.is_cuda):
            raise ValueError("ncc_cost: tensors must be on the estimator's GPU")
        depth_c = depth.detach().contiguous().to(torch.float32)
        ab_c = ab.detach().contiguous().to(torch.float32)
        views_c = views.detach().contiguous().to(torch.int32)
        if ab_c.shape != depth_c.shape + (2,) or views_c.shape != depth_c.shape:
            raise ValueError("ncc_cost: expected depth (H, W), ab (H, W, 2), views (H, W)")
        cost = torch.empty_like(depth_c)
        grad = torch.empty(depth_c.shape + (3,), dtype=torch.float32, device=depth_c.device)
        # the library runs on its own stream: the inputs must be complete first
        torch.cuda.current_stream(depth_c.device).synchronize()
        est.ncc_cost_grad_device(depth_c.data_ptr(), ab_c.data_ptr(), views_c.data_ptr(), cost.data_ptr(),
                                 grad.data_ptr())
        ctx.save_for_backward(grad)
        return cost

    @staticmethod
    def backward(ctx, grad_cost):
        (grad,) = ctx.saved_tensors
        g = grad_cost.unsqueeze(-1) * grad
        return None, g[..., 0], g[..., 1:3], None


def ncc_cost(est, depth: torch.Tensor, ab: torch.Tensor, views: torch.Tensor) -> torch.Tensor:
    return NCCCost.apply(est, depth, ab, views)


def planes_to_depth_ab(planes: torch.Tensor, K) -> tuple:
    """(n, d) planes (H, W, 4) of the reference camera -> depth (H, W) and slopes (H, W, 2)."""
    h, w = planes.shape[:2]
    dev = planes.device
    ys, xs = torch.meshgrid(torch.arange(h, device=dev, dtype=torch.float32),
                            torch.arange(w, device=dev, dtype=torch.float32), indexing="ij")
    rx = (xs - float(K[0][2])) / float(K[0][0])
    ry = (ys - float(K[1][2])) / float(K[1][1])
    n = planes[..., :3]
    denom = n[..., 0] * rx + n[..., 1] * ry + n[..., 2]
    depth = -planes[..., 3] / denom
    ab = n[..., :2] / (-n[..., 2:3])
    return depth, ab
