"""Chunk staging on the device (ddfacet_amd/chunk.py) against the oracle (oracle/chunk.py,
pinned by tests/test_chunk_oracle.py): the baseline-time sort index identical to
sorted(zip(A0, A1, TIME, rows)) (ties in time and repeated pairs included), the row gather of
every chunk array in that order, and the UpdateFlags selections (flags and per-antenna
fractions) bit for bit; the double-buffered upload pipeline delivers every chunk intact."""
import numpy as np
import pytest
import torch

from oracle import chunk as ochunk

pytestmark = pytest.mark.gpu


def _chunk(seed, nrow=20000, na=20, nchan=8, ncorr=4):
    rng = np.random.default_rng(seed)
    a0 = rng.integers(0, na, nrow).astype(np.int32)
    a1 = rng.integers(0, na, nrow).astype(np.int32)
    # MS times (MJD seconds) on a coarse grid so that equal times are frequent
    times = 5.0e9 + 8.0 * rng.integers(0, 50, nrow).astype(np.float64)
    uvw = rng.normal(0, 2000, (nrow, 3))
    flags = rng.random((nrow, nchan, ncorr)) < 0.05
    bad = rng.integers(0, na, 2)
    for b in bad:  # two antennas mostly flagged
        rows = (a0 == b) | (a1 == b)
        flags[rows] |= rng.random((int(rows.sum()), nchan, ncorr)) < 0.97
    data = (rng.standard_normal((nrow, nchan, ncorr)) + 1j * rng.standard_normal((nrow, nchan, ncorr))).astype(np.complex64)
    data[rng.random((nrow, nchan, ncorr)) < 0.001] = np.nan
    return a0, a1, times, uvw, flags, data


@pytest.mark.parametrize("seed", [1, 2])
def test_sort_index_and_gather(seed):
    from ddfacet_amd import chunk
    a0, a1, times, uvw, flags, data = _chunk(seed)
    want = ochunk.sort_index(a0, a1, times)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()  # noqa: E731
    idx = chunk.sort_baseline_time(d(a0), d(a1), d(times))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(idx.cpu().numpy(), want)
    out = chunk.apply_sort(idx, {"uvw": d(uvw), "flags": d(flags.view(np.uint8)), "data": d(data), "a0": d(a0)})
    np.testing.assert_array_equal(out["uvw"].cpu().numpy(), uvw[want])
    np.testing.assert_array_equal(out["flags"].cpu().numpy().view(bool), flags[want])
    np.testing.assert_array_equal(out["data"].cpu().numpy().view(np.uint64), data[want].view(np.uint64))
    np.testing.assert_array_equal(out["a0"].cpu().numpy(), a0[want])


@pytest.mark.parametrize("uv_range,time_range,named", [(None, None, ()), ((0.5, 3.0), None, ()),
                                                        ((0.2, 2.5), (1493283200.0 + 80, 1493283200.0 + 300), (3,))])
def test_update_flags_match_oracle(uv_range, time_range, named):
    from ddfacet_amd import chunk
    a0, a1, times, uvw, flags, data = _chunk(7)
    na = 20
    f_ref, d_ref = flags.copy(), data.copy()
    frac_ref = ochunk.update_flags(f_ref, uvw, d_ref, a0, a1, na, times_unix=times - chunk.MJD_UNIX_OFFSET_S,
                                   uv_range_km=uv_range, time_range=time_range, named_antennas=named)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()  # noqa: E731
    tf, td = d(flags.view(np.uint8)), d(data)
    frac, flagged = chunk.update_flags(tf, d(uvw), td, d(a0), d(a1), times=d(times), uv_range_km=uv_range,
                                       time_range_unix=time_range, named_antennas=named or None, na=na)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(frac.cpu().numpy(), frac_ref)
    assert (frac_ref > 0.9).sum() >= 1
    np.testing.assert_array_equal(tf.cpu().numpy().view(bool), f_ref)
    np.testing.assert_array_equal(td.cpu().numpy().view(np.uint64), d_ref.view(np.uint64))
    print("flagged antennas", np.nonzero(flagged.cpu().numpy())[0].tolist(), "flagged cells %.3f" % f_ref.mean())


def test_pipeline_delivers_every_chunk():
    from ddfacet_amd.chunk import ChunkPipeline
    rng = np.random.default_rng(3)
    chunks = [{"vis": rng.standard_normal((5000, 16, 4)).astype(np.float32), "uvw": rng.standard_normal((5000, 3))}
              for _ in range(5)]
    pipe = ChunkPipeline()
    host = [ChunkPipeline.pinned(c) for c in chunks]
    pipe.submit(0, host[0])
    sums = []
    for i in range(len(chunks)):
        if i + 1 < len(chunks):
            pipe.submit(i + 1, host[i + 1])
        dev = pipe.wait(i)
        sums.append((dev["vis"].double().sum().item(), dev["uvw"].sum().item()))
        pipe.release(i)
    torch.cuda.synchronize()
    for (sv, su), c in zip(sums, chunks):
        assert sv == pytest.approx(c["vis"].astype(np.float64).sum(), rel=1e-12)
        assert su == pytest.approx(c["uvw"].sum(), rel=1e-12)
