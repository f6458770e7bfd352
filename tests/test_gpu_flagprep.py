"""The chunk flag preparation on the device (flagprep.prepare_flags_device, ddf_prepare_flags)
against the host restatement (oracle/flagprep.py, pinned in tests/test_flagprep.py): flags and
data identical, bit for bit, with autocorrelation rows, NaN visibilities (real or imaginary part)
and partially flagged correlations, for 4 and 2 correlations and without data."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _chunk(seed, nrow=700, nchan=9, ncorr=4):
    rng = np.random.default_rng(seed)
    a0 = rng.integers(0, 6, nrow).astype(np.int32)
    a1 = rng.integers(0, 6, nrow).astype(np.int32)
    flags = rng.random((nrow, nchan, ncorr)) < 0.05
    data = (rng.standard_normal((nrow, nchan, ncorr)) + 1j * rng.standard_normal((nrow, nchan, ncorr))).astype(
        np.complex64)
    nan = rng.random((nrow, nchan, ncorr))
    data.real[nan < 0.02] = np.nan
    data.imag[(nan > 0.5) & (nan < 0.52)] = np.nan
    return flags, data, a0, a1


@pytest.mark.parametrize("ncorr,with_data", [(4, True), (2, True), (4, False)])
def test_prepare_flags_device_matches_host(ncorr, with_data):
    from ddfacet_amd.flagprep import prepare_flags_device
    from oracle.flagprep import prepare_flags
    print("case", ncorr, with_data, flush=True)
    flags, data, a0, a1 = _chunk(3 + ncorr, ncorr=ncorr)
    dev = torch.device("cuda")
    tf = torch.from_numpy(flags.astype(np.uint8)).to(dev)
    td = torch.from_numpy(data).to(dev) if with_data else None
    prepare_flags_device(tf, torch.from_numpy(a0).to(dev), torch.from_numpy(a1).to(dev), td)
    torch.cuda.synchronize()
    want_f = prepare_flags(flags.copy(), a0, a1, data if with_data else None)
    assert np.array_equal(tf.cpu().numpy().astype(bool), want_f)
    if with_data:
        got = td.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), data.view(np.uint32))


def test_prepare_flags_device_on_unaligned_views():
    """Flags and values that start one (row, channel) into a buffer (values 8-byte but not 16-byte
    aligned): the word / 16-byte path is not taken, the result is the same."""
    from ddfacet_amd.flagprep import prepare_flags_device
    from oracle.flagprep import prepare_flags
    flags, data, a0, a1 = _chunk(11, nrow=300, nchan=5, ncorr=4)
    dev = torch.device("cuda")
    bf = torch.zeros(flags.size + 4, dtype=torch.uint8, device=dev)
    bd = torch.zeros(data.size + 1, dtype=torch.complex64, device=dev)
    tf = bf[1:1 + flags.size].view(flags.shape)
    td = bd[1:1 + data.size].view(data.shape)
    tf.copy_(torch.from_numpy(flags.astype(np.uint8)))
    td.copy_(torch.from_numpy(data))
    prepare_flags_device(tf, torch.from_numpy(a0).to(dev), torch.from_numpy(a1).to(dev), td)
    torch.cuda.synchronize()
    want_f = prepare_flags(flags.copy(), a0, a1, data)
    assert np.array_equal(tf.cpu().numpy().astype(bool), want_f)
    assert np.array_equal(td.cpu().numpy().view(np.uint32), data.view(np.uint32))
