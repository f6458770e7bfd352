"""TEST INFRASTRUCTURE: numpy restatement of the facet image side of the major cycle.

    give_edges        DDFacet/ToolsDir/GiveEdges.py:29-82 (Parity=None)
    fft2c / ifft2c    ModFFTW.FFTW_2Donly.fft / .ifft (DDFacet/ToolsDir/ModFFTW.py:557-600)
    grid_to_im        ClassDDEGridMachine.GridToIm (ClassDDEGridMachine.py:1186-1190)
    facet_norm        ClassFacetMachine.BuildFacetNormImage (ClassFacetMachine.py:1289-1325)
    sum_jones_norm    ClassFacetMachine.py:1065-1074
    facets_to_im      ClassFacetMachine.FacetsToIm_Channel, kind Dirty / PSF (:1331-1416)
    give_model_tessel ClassImToGrid.GiveModelTessel(ToGrid=False, ApplyNorm=True) (:138-264)
    model_grid        GiveModelTessel + get(TranformModelInput="FT") (ClassDDEGridMachine.py:951-958)
The transforms are evaluated in FP64 (numpy) and cast to complex64; the reference uses complex64
FFTW, so parity with the device rocFFT path is to FP32 FFT accuracy.
"""
import numpy as np


def give_edges(xc0, yc0, n0, xc1, yc1, n1):
    mx0 = xc0 - n1 // 2
    x0main = max(0, mx0)
    x0facet = x0main - mx0
    mx1 = xc0 + n1 // 2
    x1main = min(n0 - 1, mx1)
    x1facet = n1 - (mx1 - x1main)
    x1main += 1
    my0 = yc0 - n1 // 2
    y0main = max(0, my0)
    y0facet = y0main - my0
    my1 = yc0 + n1 // 2
    y1main = min(n0 - 1, my1)
    y1facet = n1 - (my1 - y1main)
    y1main += 1
    return [x0main, x1main, y0main, y1main], [x0facet, x1facet, y0facet, y1facet]


def fft2c(a):
    ax = (-2, -1)
    n = a.shape[-1] * a.shape[-2]
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(a, axes=ax), axes=ax), axes=ax) / n


def ifft2c(a):
    ax = (-2, -1)
    n = a.shape[-1] * a.shape[-2]
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(a, axes=ax), axes=ax), axes=ax) * n


def grid_to_im(grid, overs):
    return ifft2c(np.asarray(grid, np.complex128) * overs ** 2).astype(np.complex64)


def facet_norm(facets, m):
    """facets: list of dict(xc, yc, sw [n, n] float32)."""
    norm = np.zeros((m, m), np.float32)
    for f in facets:
        n = f["sw"].shape[0]
        (x0d, x1d, y0d, y1d), (x0p, x1p, y0p, y1p) = give_edges(f["xc"], f["yc"], m, n // 2, n // 2, n)
        sw_grid = f["sw"].T[::-1, :]
        sw = sw_grid[::-1, :].T[x0p:x1p, y0p:y1p]
        norm[x0d:x1d, y0d:y1d] += np.real(sw)
    return norm


def sum_jones_norm(sum_jones, nband):
    out = np.ones(nband)
    if sum_jones is None:
        return out
    for ch in range(nband):
        ssq = sum_jones[1][ch]
        if ssq == 0:
            ssq = 1.
        sj = sum_jones[0][ch] / ssq
        if sj == 0:
            sj = 1.
        out[ch] = sj
    return out


def facets_to_im(facets, nband, npol, m, norm):
    """facets: list of dict(xc, yc, sphe, inv_sphe, sw, dirty [nband, npol, n, n] complex64,
    sumwt [nband, npol], sum_jones [2, nband] or None)."""
    image = np.zeros((nband, npol, m, m), np.float32)
    for f in facets:
        n = f["dirty"].shape[-1]
        sphe, inv_sphe = f["sphe"], f["inv_sphe"]
        sw_grid = f["sw"].T[::-1, :]
        (x0m, x1m, y0m, y1m), (x0f, x1f, y0f, y1f) = give_edges(f["xc"], f["yc"], m, n // 2, n // 2, n)
        sjn = sum_jones_norm(f.get("sum_jones"), nband)
        for ch in range(nband):
            for pol in range(npol):
                im = f["dirty"][ch][pol].real.copy()
                w = f["sumwt"][ch][pol] * np.sqrt(sjn[ch])
                w = w if w else 1.0
                im[...] = (im.astype(np.float64) * inv_sphe.astype(np.float64) * sw_grid.astype(np.float64) / w)
                im[sphe < 1e-3] = 0
                im = im[::-1, :].T[x0f:x1f, y0f:y1f]
                image[ch, pol, x0m:x1m, y0m:y1m] += im
    if (norm != 0.0).any():
        with np.errstate(divide="ignore", invalid="ignore"):
            image /= norm[None, None]
    return image


def give_model_tessel(image, f, norm, overs):
    """Facet model (grid orientation, before the FFT) of facet dict f from the main image."""
    nch, npol, m, _ = image.shape
    n = f["sw"].shape[0]
    (x0d, x1d, y0d, y1d), (x0p, x1p, y0p, y1p) = give_edges(f["xc"], f["yc"], m, n // 2, n // 2, n)
    out = np.zeros((nch, npol, n, n), np.float32)
    for ch in range(nch):
        for pol in range(npol):
            mi = np.zeros((n, n), np.float32)
            mi[x0p:x1p, y0p:y1p] = image[ch, pol][x0d:x1d, y0d:y1d].real
            if (mi == 0).all():
                out[ch, pol] = mi
                continue
            with np.errstate(divide="ignore", invalid="ignore"):
                mi[x0p:x1p, y0p:y1p] /= norm[x0d:x1d, y0d:y1d].real
                mi[x0p:x1p, y0p:y1p] *= f["sw"][x0p:x1p, y0p:y1p]
                mi[x0p:x1p, y0p:y1p] /= f["sphe"][x0p:x1p, y0p:y1p].real
            mi[f["sphe"] < 1e-3] = 0
            out[ch, pol] = mi.T[::-1, :]
    out *= (overs * n) ** 2
    return out


def model_grid(image, f, norm, overs):
    return fft2c(give_model_tessel(image, f, norm, overs).astype(np.complex128)).astype(np.complex64)
