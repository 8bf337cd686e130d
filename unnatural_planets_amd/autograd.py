"""torch.autograd wrapper of the field and its hand-written adjoint.

    theta = torch.tensor([...], device="cuda", requires_grad=True)
    fn = FieldParams(spec, [("amp", 0, 0), ("freq", 0, 0), ("weight", 1, 0), ("shape", 0, 0)])
    f = fn(theta, xyz)          # upf_field_eval_points with theta written into the spec
    f.square().mean().backward()  # upf_field_adjoint: d/dtheta and d/dxyz

Forward writes the current theta values into a host copy of the spec (one device-to-host
read of len(theta) floats), evaluates the field on the GPU and keeps the spec for
backward; backward runs the fused forward + adjoint kernel once and returns
dL/dtheta (gathered from the 76-entry upf_theta_grad vector) and dL/dxyz.
"""
from __future__ import annotations

import torch

from . import field_adjoint, field_eval_points
from .spec import FieldSpec, copy_spec, theta_index


def set_param(s: FieldSpec, kind: str, a: int, b: int, value: float) -> None:
    if kind == "shape":
        s.shape_params[a] = value
    elif kind == "weight":
        s.layers[a].weight = value
    elif kind == "freq":
        s.layers[a].freq[b] = value
    elif kind == "amp":
        s.layers[a].amp[b] = value
    else:
        raise KeyError(kind)


def get_param(s: FieldSpec, kind: str, a: int, b: int) -> float:
    if kind == "shape":
        return float(s.shape_params[a])
    if kind == "weight":
        return float(s.layers[a].weight)
    if kind == "freq":
        return float(s.layers[a].freq[b])
    if kind == "amp":
        return float(s.layers[a].amp[b])
    raise KeyError(kind)


class _FieldFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, xyz, base: FieldSpec, params):
        s = copy_spec(base)
        for (kind, a, b), v in zip(params, theta.detach().cpu().tolist()):
            set_param(s, kind, a, b, v)
        f = field_eval_points(s, xyz.detach().contiguous())
        ctx.spec = s
        ctx.params = params
        ctx.save_for_backward(xyz)
        return f

    @staticmethod
    def backward(ctx, dLdf):
        (xyz,) = ctx.saved_tensors
        dLdf = dLdf.contiguous().to(torch.float32)
        _, dfdx, dth = field_adjoint(ctx.spec, xyz.detach().contiguous(), dLdf, want_f=False,
                                     want_dfdx=ctx.needs_input_grad[1])
        idx = torch.tensor([theta_index(k, a, b) for (k, a, b) in ctx.params], device=dth.device)
        g_theta = dth.index_select(0, idx)
        g_xyz = dfdx * dLdf[:, None] if ctx.needs_input_grad[1] else None
        return g_theta, g_xyz, None, None


class FieldParams:
    """The field of `spec` as a differentiable function of the listed parameters.
    Each parameter is (kind, a, b): ("shape", i, 0), ("weight", layer, 0),
    ("freq", layer, octave) or ("amp", layer, octave)."""

    def __init__(self, spec: FieldSpec, params):
        self.spec = copy_spec(spec)
        self.params = [tuple(p) for p in params]

    def initial(self, device="cuda") -> torch.Tensor:
        return torch.tensor([get_param(self.spec, *p) for p in self.params], dtype=torch.float32, device=device)

    def __call__(self, theta: torch.Tensor, xyz: torch.Tensor) -> torch.Tensor:
        return _FieldFn.apply(theta, xyz, self.spec, self.params)
