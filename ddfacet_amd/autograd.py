"""Differentiable forward model: model image -> predicted visibilities, with the hand-written
adjoint as its backward.

The forward is the degridding half of a major cycle (GiveModelTessel + centred FFT per facet,
then the batched all-facet degridder, ClassImToGrid.py:138-264, ClassDDEGridMachine.py:922-1108):

    A(x) = sum_f D_f F M_f x            x: float32 [nband, npol, m, m] model image

with M_f the facet cut / normalisation, F the centred forward FFT and D_f the degridder of facet f.
The backward applies A^H = sum_f M_f^H F^H D_f^H exactly (ddf_plan_degrid_adjoint,
ddf_model_to_grid_adjoint): conjugate phasors and CF stamps, J0^H V J1 Jones, the transposed
Stokes map, the adjoint centred FFT and the transposed cut.  For a real image and a real loss,
torch's vector-Jacobian product is Re(A^H g).
"""
import ctypes as C

import torch

from . import imaging
from ._lib import check, lib
from .device import stream_handle


class DegridOperator:
    def __init__(self, cycle):
        self.cycle = cycle
        fs = cycle.fs
        self.shape = tuple(fs.t["vis"].shape)
        self.dev = fs.dev

    def forward(self, image):
        """Model visibilities A(image), complex64 [nrow, nchan, ncorr]."""
        vis = torch.zeros(self.shape, dtype=torch.complex64, device=self.dev)
        self.cycle.predict(image.contiguous(), vis=vis)
        return -vis

    def adjoint(self, vis_bar):
        """Re(A^H vis_bar) as a float32 image [nband, npol, m, m]."""
        cyc = self.cycle
        fs = cyc.fs
        vb = vis_bar.to(torch.complex64).contiguous()
        fs.grids.zero_()
        check(lib().ddf_plan_degrid_adjoint(fs._plan, C.c_void_p(vb.data_ptr()), C.c_void_p(stream_handle())))
        out = torch.empty_like(cyc.image)
        g = cyc.geom
        arr = imaging._facets_struct(g, cyc.sphe, None, cyc.sw, model=fs.grids)
        check(lib().ddf_model_to_grid_adjoint(arr, g.nf, g.n, int(out.shape[0]), int(out.shape[1]), g.m,
                                              C.c_void_p(out.data_ptr()), C.c_void_p(cyc.norm.data_ptr()),
                                              cyc.overs, C.c_void_p(stream_handle())))
        return out


class _Degrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, op):
        ctx.op = op
        return op.forward(image)

    @staticmethod
    def backward(ctx, grad_vis):
        return ctx.op.adjoint(grad_vis), None


def degrid(image, op):
    """Differentiable A(image) (see module docstring)."""
    return _Degrid.apply(image, op)


def chi2(image, op, data, weights):
    """sum w |A(image) - data|^2 over every visibility and correlation (weights [nrow, nchan])."""
    r = degrid(image, op) - data
    return (weights[..., None] * (r.real.double() ** 2 + r.imag.double() ** 2)).sum()
