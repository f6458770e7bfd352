"""TEST INFRASTRUCTURE: FP64 direct Fourier transforms (the physics oracle).

predict      V(row, chan) = sum_s S_s exp(+2 pi i nu/c (u l_s + v m_s + w (n_s - 1))) on XX and YY,
             the sign and (l, m) convention of ClassDDEGridMachine.testGrid
             (DDFacet/Imager/ClassDDEGridMachine.py:73-83, 229-271).
dirty        I(l, m) = sum w Re[V exp(-2 pi i nu/c (u l + v m + w (n - 1)))] / sum w over unflagged
             visibilities, Stokes I = (XX + YY) / 2: what gridding + FFT + taper correction
             approximates at a pixel.
image_pixel_lm  (l, m) of main-image pixel (X, Y): ((X - M//2) cell, (Y - M//2) cell).
Pure numpy, chunked; meant for small cases (seconds).
"""
import numpy as np

C_LIGHT = 299792458.0


def image_pixel_lm(X, Y, m, cell):
    return (np.asarray(X) - m // 2) * cell, (np.asarray(Y) - m // 2) * cell


def predict(uvw, freqs, flux, l, m, chunk=4096):
    l = np.asarray(l, np.float64)
    m = np.asarray(m, np.float64)
    n1 = np.sqrt(1 - l * l - m * m) - 1
    R = uvw.shape[0]
    out = np.zeros((R, freqs.size), np.complex128)
    for r0 in range(0, R, chunk):
        q = uvw[r0:r0 + chunk]
        geo = q[:, 0:1] * l[None] + q[:, 1:2] * m[None] + q[:, 2:3] * n1[None]
        ph = (2 * np.pi / C_LIGHT) * geo[:, None, :] * freqs[None, :, None]
        out[r0:r0 + chunk] = (np.exp(1j * ph) * np.asarray(flux)[None, None, :]).sum(-1)
    return out


def dirty(uvw, freqs, vis_i, weights, flags, l, m, chunk=2048):
    """Direct dirty image (Stokes I) at the points (l, m)."""
    l = np.asarray(l, np.float64).ravel()
    m = np.asarray(m, np.float64).ravel()
    n1 = np.sqrt(1 - l * l - m * m) - 1
    w = np.where(flags, 0.0, weights.astype(np.float64))
    acc = np.zeros(l.size)
    for r0 in range(0, uvw.shape[0], chunk):
        q = uvw[r0:r0 + chunk]
        geo = q[:, 0:1] * l[None] + q[:, 1:2] * m[None] + q[:, 2:3] * n1[None]   # [r, p]
        ph = (-2 * np.pi / C_LIGHT) * geo[:, None, :] * freqs[None, :, None]      # [r, c, p]
        acc += np.einsum("rc,rcp->p", w[r0:r0 + chunk] * 1.0, np.real(vis_i[r0:r0 + chunk, :, None] * np.exp(1j * ph)))
    return acc / w.sum()
