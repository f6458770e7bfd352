"""ddf_weights_briggs_noalloc (the robust transform with its partial sums in a library-owned
device array) against ddf_weights_briggs (caller workspace) through the C ABI: identical grids on
two calls in a row, the second with fewer bands and cells (the array keeps the first call's
contents), and the band limit rejected with an error."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _grid(seed, nbands, npix):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0, 50, (nbands, npix))
    g[rng.random((nbands, npix)) < 0.4] = 0
    return torch.from_numpy(g).cuda()


def test_noalloc_briggs_matches_workspace_entry_point():
    from ddfacet_amd._lib import check, lib
    L = lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for seed, nbands, npix, robust in ((1, 5, 200003, 0.5), (2, 2, 1001, -1.0)):
        a = _grid(seed, nbands, npix)
        b = a.clone()
        work = torch.empty(int(L.ddf_weights_briggs_work_size(nbands)), dtype=torch.float64, device="cuda")
        check(L.ddf_weights_briggs(C.c_void_p(a.data_ptr()), nbands, npix, robust, C.c_void_p(work.data_ptr()),
                                   work.numel(), st))
        check(L.ddf_weights_briggs_noalloc(C.c_void_p(b.data_ptr()), nbands, npix, robust, st))
        torch.cuda.synchronize()
        assert torch.equal(a, b), (seed, nbands, npix)
        assert not torch.equal(b, _grid(seed, nbands, npix))
    g = _grid(3, 65, 10)
    with pytest.raises(ValueError):
        check(L.ddf_weights_briggs_noalloc(C.c_void_p(g.data_ptr()), 65, 10, 0.0, st))
