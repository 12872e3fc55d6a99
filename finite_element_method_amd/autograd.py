"""torch.autograd bridge: the global stiffness values as a differentiable
function of element parameters and node coordinates.

    from finite_element_method_amd import FEM, PLATE
    from finite_element_method_amd.autograd import assemble
    val = assemble(model, plate_params=p, xyz=x)   # CSR values (nnz,), float64 CUDA
    loss = (val * w).sum(); loss.backward()        # p.grad, x.grad from fem_backward

Forward replaces the model's parameters / coordinates on the device
(fem_update_params / fem_update_nodes) and assembles into a fresh tensor;
backward calls the hand-written adjoint kernels (fem_backward), which recompute
the element stiffness instead of storing it.  The sparsity pattern is the
model's (fem_csr_info); a parameter change that alters a frame's zero pattern
changes it, and the caller then sees a new nnz.
"""

import torch

from . import BEAM, PLATE, TRUSS


class AssembleStiffness(torch.autograd.Function):
    @staticmethod
    def forward(ctx, model, truss_params, beam_params, plate_params, xyz):
        stream = torch.cuda.current_stream()
        for t, prm in ((TRUSS, truss_params), (BEAM, beam_params), (PLATE, plate_params)):
            if prm is not None:
                model.update_params(t, prm, stream=stream)
        if xyz is not None:
            model.update_nodes(xyz, stream=stream)
        model.build_pattern()
        _, nnz = model.csr_info()
        dev = next((x.device for x in (truss_params, beam_params, plate_params, xyz) if x is not None),
                   torch.device("cuda", torch.cuda.current_device()))
        val = torch.empty(nnz, dtype=torch.float64, device=dev)
        model.assemble_into(val, stream=stream)
        ctx.model = model
        ctx.version = getattr(model, "_version", 0)
        return val

    @staticmethod
    def backward(ctx, grad):
        model = ctx.model
        if getattr(model, "_version", 0) != ctx.version:
            raise RuntimeError("the model's parameters changed between forward and backward")
        need = ctx.needs_input_grad
        grads, dxyz = model.backward(grad.contiguous(), params=any(need[1:4]), xyz=need[4])
        out = [None]
        for i, t in ((1, TRUSS), (2, BEAM), (3, PLATE)):
            out.append(grads.get(t) if need[i] else None)
        out.append(dxyz if need[4] else None)
        return tuple(out)


def assemble(model, truss_params=None, beam_params=None, plate_params=None, xyz=None):
    """Global stiffness values (CSR order) as a differentiable function of the
    given parameter tensors (add_* layouts, float64 CUDA) and node coordinates."""
    return AssembleStiffness.apply(model, truss_params, beam_params, plate_params, xyz)
