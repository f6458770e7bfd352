"""The imaging major cycle on one GPU, all facets batched.

Device counterpart of the reference's per-chunk facet loop (ClassImagerDeconv.main ->
ClassFacetMachine.putChunkInBackground / FacetsToIm / getChunkInBackground,
DDFacet/Imager/ClassDeconvMachine.py:1347-1395, ClassFacetMachine.py:912-1249, 1859-1943):

    residual_image(vis)   grid every facet (batched plan), GridToIm (rocFFT, x OverS^2), stitch
                          with FacetNorm -> residual / dirty (or PSF) image on the main grid
    predict(model, vis)   GiveModelTessel + forward FFT per facet, then vis -= sum_f degrid_f
    cycle(model, vis)     predict then residual_image: one major cycle

Everything stays in HBM; the deconvolution (minor cycle) that turns the residual into the next
model is the reference CPU code and is outside this module.
"""
import numpy as np
import torch

from . import imaging
from .facets import DeviceFacetSet


class MajorCycle:
    def __init__(self, P, geom, device=None, precision="mixed", tile=0, chunk=0, workgroups=0, stream=None):
        self.P = P
        self.geom = geom
        self.fs = DeviceFacetSet(P, device=device)
        dev = self.fs.dev
        self.overs = int(P.facets[0].winfos[3])
        sphe = np.stack([np.float32(f.sphe) for f in P.facets])
        self.sphe = torch.from_numpy(sphe).to(dev)
        self.inv_sphe = torch.from_numpy(np.float32(1.0 / sphe.astype(np.float64))).to(dev)
        self.sw = torch.from_numpy(geom.spatial_weights()).to(dev)
        self.norm = torch.zeros((geom.m, geom.m), dtype=torch.float32, device=dev)
        imaging.facet_norm(geom, self.sw, self.norm, stream)
        self.plan_info = self.fs.plan_batched(tile=tile, chunk=chunk, precision=precision,
                                              workgroups=workgroups, stream=stream)
        self.image = torch.zeros((P.nband, P.npol, geom.m, geom.m), dtype=torch.float32, device=dev)

    def residual_image(self, vis=None, dopsf=0, stream=None):
        fs = self.fs
        fs.zero_grids()
        fs.grid_batched(dopsf=dopsf, stream=stream, vis=vis)
        imaging.grid_to_image(fs.grids, self.overs, stream)
        sj = fs.sum_jones if fs.jones is not None else None
        imaging.stitch(self.geom, self.sphe, self.inv_sphe, self.sw, fs.grids, fs.sumwt, sj, self.norm,
                       self.image, stream)
        return self.image

    def predict(self, model_image, vis=None, stream=None):
        imaging.model_to_grid(self.geom, self.sphe, self.sw, model_image, self.norm, self.fs.models, self.overs,
                              stream)
        self.fs.degrid_batched(stream=stream, vis=vis)

    def cycle(self, model_image, vis=None, stream=None):
        self.predict(model_image, vis, stream)
        return self.residual_image(vis, 0, stream)
