"""ClassDDEGridMachine on the GPU: the per-facet grid machine surface the reference Imager calls.

Mirrors DDFacet/Imager/ClassDDEGridMachine.py:308-1190 (constructor, reinitGrid, setDecorr,
GiveParamJonesList, put, get, CheckTypes, GridToIm), with the native calls routed to the HIP
library through ddfacet_amd._pyGridderSmearPols (same argument lists as the reference module).
``GD`` is the reference parset dictionary (sections CF, Image, RIME, Facets, Comp,
DDESolutions); only the keys the reference reads here are used.
"""
import numpy as np

from . import _pyGridderSmearPols as _gridder
from .configs import estimate_npix
from .wterm import ClassWTermModified


class ClassDDEGridMachine:
    def __init__(self, GD, ChanFreq, Npix, lmShift=(0., 0.), IDFacet=0, SpheNorm=True, NFreqBands=1,
                 DataCorrelationFormat=[5, 6, 7, 8], ExpectedOutputStokes=[1], ListSemaphores=None,
                 cf_dict=None, compute_cf=False, wmax=None, bda_grid=None, bda_degrid=None, Padding=None,
                 Sups=None):
        self.GD = GD
        self.IDFacet = IDFacet
        self.SpheNorm = SpheNorm
        self._bda_grid = bda_grid
        self._bda_degrid = bda_degrid
        prec = GD["RIME"]["Precision"]
        self.dtype = np.complex64 if prec == "S" else np.complex128
        if Padding is None:
            Padding = GD["Facets"]["Padding"]
        self.NonPaddedNpix, Npix = estimate_npix(Npix, Padding)
        self.Padding = Npix / float(self.NonPaddedNpix)
        self.LSmear = []
        self.PolMode = ExpectedOutputStokes
        self.npol = len(ExpectedOutputStokes)
        self.SkyType = 1
        self.PolMap = np.array([0, 5, 5, 0], np.int32)
        self.PolModeID = 0
        self.Npix = Npix
        self.NFreqBands = NFreqBands
        self.NonPaddedShape = (NFreqBands, self.npol, self.NonPaddedNpix, self.NonPaddedNpix)
        self.GridShape = (NFreqBands, self.npol, Npix, Npix)
        x0 = (Npix - self.NonPaddedNpix) // 2
        self.PaddingInnerCoord = (x0, x0 + self.NonPaddedNpix)
        self.OverS = GD["CF"]["OverS"]
        self.Sup = GD["CF"]["Support"]
        self.Nw = GD["CF"]["Nw"]
        self.Cell = GD["Image"]["Cell"]
        self.incr = (np.array([-self.Cell, self.Cell], dtype=np.float64) / 3600.) * (np.pi / 180)
        self.JonesType = {"Scalar": 0, "Diag": 1, "Full": 2}.get(GD["DDESolutions"]["JonesMode"], 2)
        self.ChanFreq = np.asarray(ChanFreq, np.float64)
        self.lmShift = lmShift
        if cf_dict is not None or compute_cf:
            self.InitCF(cf_dict, compute_cf, wmax, Sups)
        self.reinitGrid()
        self.DataCorrelationFormat = DataCorrelationFormat
        self.ExpectedOutputStokes = ExpectedOutputStokes

    def InitCF(self, cf_dict, compute_cf, wmax, sups=None):
        if compute_cf or cf_dict is None or "Wplanes" not in cf_dict:
            self.WTerm = ClassWTermModified(Cell=self.Cell, Sup=self.Sup, Npix=self.Npix, Freqs=self.ChanFreq,
                                            wmax=wmax, Nw=self.Nw, OverS=self.OverS, lmShift=self.lmShift,
                                            sups=sups, IDFacet=self.IDFacet)
            if cf_dict is not None:
                cf_dict.update(Wplanes=self.WTerm, Sphe=self.WTerm.Sphe, InvSphe=self.WTerm.InvSphe,
                               CuCv=np.array([self.WTerm.Cu, self.WTerm.Cv]), wmax=self.WTerm.wmax)
        else:
            self.WTerm = cf_dict["Wplanes"]
        self.ifzfCF = self.WTerm.ifzfCF

    def reinitGrid(self):
        self.NChan, self.npol, _, _ = self.GridShape
        self.SumWeigths = np.zeros((self.NChan, self.npol), np.float64)
        self.SumJones = np.zeros((2, self.NChan), np.float64)

    def setDecorr(self, uvw_dt, DT, Dnu, SmearMode="FT", lm_min=None, lm_PhaseCenter=None):
        fs = 1 if "F" in SmearMode else 0
        ts = 1 if "T" in SmearMode else 0
        lmin, mmin = self.lmShift
        if lm_min is not None:
            lmin, mmin = lm_min
        if lm_PhaseCenter is not None:
            lmin -= lm_PhaseCenter[0]
            mmin -= lm_PhaseCenter[1]
        if uvw_dt.dtype != np.float64:
            raise NameError("uvw_dt.dtype %s %s" % (str(uvw_dt.dtype), str(np.float64)))
        self.LSmear = [uvw_dt, DT, Dnu, ts, fs, lmin, mmin]

    def GiveParamJonesList(self, DicoJonesMatrices, times, A0, A1, uvw, gridder=False, degridder=False):
        """LJones[0:13] (ClassDDEGridMachine.py:571-679)."""
        apply_kms = "DicoJones_killMS" in DicoJonesMatrices
        apply_beam = "DicoJones_Beam" in DicoJonesMatrices
        l0, m0 = self.lmShift
        idir_kms = 0
        w_kms = np.array([], np.float32)
        interp = self.GD["DDESolutions"]["Type"]
        d0 = self.GD["DDESolutions"]["Scale"] * np.pi / 180
        gamma = self.GD["DDESolutions"]["gamma"]
        if apply_kms:
            dirs = DicoJonesMatrices["DicoJones_killMS"]["Dirs"]
            d = np.sqrt((l0 - dirs["l"]) ** 2 + (m0 - dirs["m"]) ** 2)
            idir_kms = np.argmin(d)
            w = dirs["I"] / (1. + d / d0) ** gamma
            w /= np.sum(w)
            w[w < (0.2 * w.max())] = 0
            ind = np.argsort(w)[::-1]
            w[ind[3::]] = 0
            w /= np.sum(w)
            w_kms = w
        idir_beam = 0
        if apply_beam:
            dirs = DicoJonesMatrices["DicoJones_Beam"]["Dirs"]
            idir_beam = np.argmin(np.sqrt((l0 - dirs["l"]) ** 2 + (m0 - dirs["m"]) ** 2))
        interp = {"Nearest": 0, "Krigging": 1}.get(interp, interp)
        mode = self.GD["Comp"]["BDAJones"]
        if gridder:
            bda_mode = 2 if mode in ("grid", "both") else 1
        else:
            bda_mode = 2 if mode == "both" else 1
        if A0.size != uvw.shape[0]:
            raise RuntimeError("Antenna array is expected to have the same number of rows as the uvw array")
        empty4 = np.array([], np.complex64).reshape((0, 0, 0, 0))
        jb, mb, cb = empty4, np.array([], np.int32), np.array([], np.int32)
        jk, mk, ck = empty4, np.array([], np.int32), np.array([], np.int32)
        alpha = np.array([], np.float32).reshape((0, 0))
        if apply_beam:
            D = DicoJonesMatrices["DicoJones_Beam"]
            jb, mb, cb = D["Jones"]["Jones"], D["TimeMapping"], np.int32(D["Jones"]["VisToJonesChanMapping"])
            self.CheckTypes(A0=A0, A1=A1, Jones=jb)
        if apply_kms:
            D = DicoJonesMatrices["DicoJones_killMS"]
            jk, mk, ck = D["Jones"]["Jones"], D["TimeMapping"], np.int32(D["Jones"]["VisToJonesChanMapping"])
            if D.get("AlphaReg") is not None:
                alpha = D["AlphaReg"]
            self.CheckTypes(A0=A0, A1=A1, Jones=jk)
        return [jk, mk, jb, mb, A0, A1, np.array([idir_kms], np.int32), np.float32(w_kms),
                np.array([idir_beam], np.int32), np.array([interp, bda_mode], np.int32), ck, cb, alpha]

    def _jones_list(self, DicoJonesMatrices, times, A0, A1, uvw, side):
        if DicoJonesMatrices is None:
            return []
        key = "DDModeGrid" if side == "grid" else "DDModeDeGrid"
        scale_key = "ScaleAmpGrid" if side == "grid" else "ScaleAmpDeGrid"
        amp = 1 if "A" in self.GD["DDESolutions"][key] else 0
        phase = 1 if "P" in self.GD["DDESolutions"][key] else 0
        scale, calib = 0, 0.
        if self.GD["DDESolutions"][scale_key]:
            scale = 1
            calib = (self.GD["DDESolutions"]["CalibErr"] / 3600.) * np.pi / 180
        lst = self.GiveParamJonesList(DicoJonesMatrices, times, A0, A1, uvw, gridder=(side == "grid"),
                                      degridder=(side != "grid"))
        return lst + [amp, phase, scale, calib, self.SumJones, self.SumJonesChan,
                      np.float32(self.GD["DDESolutions"]["ReWeightSNR"])]

    @staticmethod
    def _equidistant(freqs):
        if freqs.size > 1:
            df = freqs[1:] - freqs[:-1]
            return int(np.max(np.abs(df - np.mean(df))) < 1.)
        return 0

    def _winfos(self):
        return np.array([self.WTerm.RefWave, self.WTerm.wmax, len(self.WTerm.Wplanes), self.WTerm.OverS],
                        dtype=np.float64)

    def put(self, times, uvw, visIn, flag, A0A1, W=None, PointingID=0, DoNormWeights=True,
            DicoJonesMatrices=None, freqs=None, DoPSF=0, ChanMapping=None, ResidueGrid=None,
            sparsification=None):
        """Grid visIn onto ResidueGrid (ClassDDEGridMachine.py:681-861)."""
        vis = visIn
        self.DoNormWeights = DoNormWeights
        if not self.DoNormWeights:
            self.reinitGrid()
        if ChanMapping is None:
            ChanMapping = np.zeros((visIn.shape[1],), np.int64)
        self.ChanMappingGrid = ChanMapping
        Grid = ResidueGrid
        if Grid.dtype != self.dtype:
            raise TypeError("Grid must be of type " + str(self.dtype))
        A0, A1 = A0A1
        self.SumJonesChan = np.zeros((2, vis.shape[1]), np.float64)
        if W is None:
            W = np.ones((uvw.shape[0], vis.shape[1]), dtype=np.float64)
        if vis.shape != flag.shape:
            raise Exception("vis[%s] and flag[%s] should have the same shape" % (str(vis.shape), str(flag.shape)))
        l0, m0 = self.lmShift
        FacetInfos = np.float64(np.array([self.WTerm.Cu, self.WTerm.Cv, l0, m0, self.IDFacet]))
        self.CheckTypes(Grid=Grid, vis=vis, uvw=uvw, flag=flag, ListWTerm=self.WTerm.Wplanes, W=W)
        LJones = self._jones_list(DicoJonesMatrices, times, A0, A1, uvw, "grid")
        if self.GD["RIME"]["BackwardMode"] != "BDA-grid":
            raise ValueError("unknown --RIME-BackwardMode %s" % self.GD["RIME"]["BackwardMode"])
        opt = [self.JonesType, self._equidistant(freqs), self.SkyType, self.PolModeID]
        fn = _gridder.pyGridderWPol32 if self.dtype == np.complex64 else _gridder.pyGridderWPol64
        fn(Grid, vis, uvw, flag, W, self.SumWeigths, DoPSF, self.WTerm.Wplanes, self.WTerm.WplanesConj,
           self._winfos(), self.incr.astype(np.float64), freqs, [self.PolMap, FacetInfos], LJones, self._bda_grid,
           sparsification if sparsification is not None else np.array([], dtype=bool), opt, self.LSmear,
           np.int32(ChanMapping), np.array(self.DataCorrelationFormat).astype(np.uint16),
           np.array(self.ExpectedOutputStokes).astype(np.uint16))

    def CheckTypes(self, Grid=None, vis=None, uvw=None, flag=None, ListWTerm=None, W=None, A0=None, A1=None,
                   Jones=None):
        """dtype / contiguity checks of ClassDDEGridMachine.CheckTypes (:863-920)."""
        def need(x, dtype, name):
            if x is None:
                return
            if x.dtype != dtype:
                raise NameError("%s.dtype %s" % (name, str(x.dtype)))
            if not x.flags.c_contiguous:
                raise NameError("%s has to be contiguous" % name)
        need(Grid, np.dtype(self.dtype), "Grid")
        need(vis, np.complex64, "vis")
        need(uvw, np.float64, "uvw")
        need(flag, np.bool_, "flag")
        need(W, np.float32, "W")
        need(A0, np.int32, "A0")
        need(A1, np.int32, "A1")
        need(Jones, np.complex64, "Jones")
        if ListWTerm is not None and ListWTerm[0].dtype != np.complex64:
            raise NameError("ListWTerm.dtype %s" % str(ListWTerm[0].dtype))

    def get(self, times, uvw, visIn, flag, A0A1, ModelImage, PointingID=0, Row0Row1=(0, -1),
            DicoJonesMatrices=None, freqs=None, ImToGrid=True, TranformModelInput="", ChanMapping=None,
            sparsification=None):
        """vis -= degridded model (ClassDDEGridMachine.py:922-1152, ForwardMode BDA-degrid)."""
        vis = visIn.view()
        A0, A1 = A0A1
        if ImToGrid:
            if np.max(np.abs(ModelImage)) == 0:
                return vis
            Grid = self.dtype(self.setModelIm(ModelImage))
        else:
            Grid = ModelImage
        if ChanMapping is None:
            ChanMapping = np.zeros((visIn.shape[1],), np.int32)
        if TranformModelInput == "FT":
            if np.max(np.abs(ModelImage)) == 0:
                return vis
            Grid = np.complex64(self.FT(np.complex64(ModelImage)))
        self.ChanMappingDegrid = np.int32(ChanMapping)
        self.SumJonesChan = np.zeros((2, visIn.shape[1]), np.float64)
        if vis.shape != flag.shape:
            raise Exception("vis[%s] and flag[%s] should have the same shape" % (str(vis.shape), str(flag.shape)))
        l0, m0 = self.lmShift
        FacetInfos = np.float64(np.array([self.WTerm.Cu, self.WTerm.Cv, l0, m0, self.IDFacet]))
        Row0, Row1 = Row0Row1
        if Row1 == -1:
            Row1 = uvw.shape[0]
        RowInfos = np.array([Row0, Row1]).astype(np.int32)
        self.CheckTypes(Grid=Grid, vis=vis, uvw=uvw, flag=flag, ListWTerm=self.WTerm.Wplanes)
        LJones = self._jones_list(DicoJonesMatrices, times, A0, A1, uvw, "degrid")
        if self.GD["RIME"]["ForwardMode"] != "BDA-degrid":
            raise ValueError("unknown --RIME-ForwardMode %s" % self.GD["RIME"]["ForwardMode"])
        opt = [self.JonesType, self._equidistant(freqs), self.SkyType, self.PolModeID]
        return _gridder.pyDeGridderWPol(
            Grid, vis, uvw, flag, self.SumWeigths, 0, self.WTerm.WplanesConj, self.WTerm.Wplanes, self._winfos(),
            self.incr.astype(np.float64), freqs, [self.PolMap, FacetInfos, RowInfos], LJones, self._bda_degrid,
            sparsification if sparsification is not None else np.array([], dtype=bool), opt, self.LSmear,
            np.int32(ChanMapping), np.array(self.DataCorrelationFormat).astype(np.uint16),
            np.array(self.ExpectedOutputStokes).astype(np.uint16))

    # ---- FFT glue (ClassDDEGridMachine.py:1158-1190, ModFFTW.FFTW_2Donly) ------------------------
    def _fft_dev(self, a, inverse, scale):
        import torch
        from .imaging import fft2c
        t = torch.from_numpy(np.ascontiguousarray(a, np.complex64)).cuda()
        fft2c(t, inverse, scale)
        return t.cpu().numpy()

    def FT(self, A):
        return self._fft_dev(A, False, 1.0)

    def setModelIm(self, ModelIm):
        _, _, n, n = ModelIm.shape
        x0, x1 = self.PaddingInnerCoord
        padded = np.zeros(self.GridShape, dtype=self.dtype)
        padded[:, :, x0:x1, x0:x1] = ModelIm
        return self.ImToGrid(padded) * n ** 2

    def ImToGrid(self, ModelIm):
        corr = ModelIm * (self.OverS * self.Padding) ** 2
        nchan, npol, _, _ = corr.shape
        for ichan in range(nchan):
            for ipol in range(npol):
                corr[ichan, ipol][:, :] = corr[ichan, ipol][:, :].real / self.ifzfCF.real
        return self.FT(corr)

    def GridToIm(self, Grid):
        """Grid *= OverS^2 then the unnormalised centred inverse FFT (ClassDDEGridMachine.py:1186-1190)."""
        out = self._fft_dev(Grid, True, float(self.WTerm.OverS) ** 2)
        Grid[...] = out
        return Grid
