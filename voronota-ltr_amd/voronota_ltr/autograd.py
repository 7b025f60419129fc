"""torch.autograd.Function over the tessellation energy's adjoint.

    E = sum_contacts W[t_a][t_b] * area + sas_weight * sum SAS + volume_weight * sum volume

as a differentiable function of the ball centres (a float64 tensor on the GPU).
The forward call runs the tessellation and the adjoint on the library's own
stream, on the tensors' device memory, and keeps the gradient for backward.
torch is plumbing here (device memory and streams); the compute is the HIP
library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native


class TessellationEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, centers, radii, types, W, probe, sas_weight=1.0, volume_weight=1.0):
        if centers.dtype != torch.float64 or not centers.is_cuda:
            raise TypeError("centers must be a float64 tensor on the GPU")
        n = centers.shape[0]
        balls = torch.cat([centers.detach(), radii.to(centers).reshape(n, 1)], dim=1).contiguous()
        t32 = types.to(device=centers.device, dtype=torch.int32).contiguous()
        grad = torch.empty((n, 3), dtype=torch.float64, device=centers.device)
        w = np.ascontiguousarray(np.asarray(W.detach().cpu() if torch.is_tensor(W) else W, dtype=np.float64)
                                 .reshape(16))
        lib = _native.load_library()
        # the library works on its own stream: inputs must be ready
        torch.cuda.current_stream(centers.device).synchronize()
        e = C.c_double()
        rc = lib.vltr_energy_gradient_device(_native.context(centers.device.index or 0), C.c_void_p(balls.data_ptr()),
                                             n, float(probe), None, C.c_void_p(t32.data_ptr()),
                                             w.ctypes.data_as(C.POINTER(C.c_double)), float(sas_weight),
                                             float(volume_weight), C.c_void_p(grad.data_ptr()), C.byref(e))
        _native.check(rc, "vltr_energy_gradient_device")
        ctx.save_for_backward(grad)
        return torch.tensor(e.value, dtype=torch.float64, device=centers.device)

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None, None, None, None


def tessellation_energy(centers, radii, types, W, probe, sas_weight=1.0, volume_weight=1.0):
    return TessellationEnergy.apply(centers, radii, types, W, probe, sas_weight, volume_weight)
