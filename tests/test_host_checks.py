"""Host-side argument checks of the drop-ins (no GPU needed).

Every kernel indexes its inputs and outputs by the dimensions alone, so an
undersized dip, mask, spray volume or out= buffer must be refused with a
ValueError before the library is called (on this CPU-only host the library
call itself would raise a SeistrError, so reaching it fails these tests)."""
import numpy as np
import pytest
import torch

import pyseistr_amd as ps
from pyseistr_amd import _lib, bandpass, batch, dip, filters, interp, smooth, warmstart

N1, N2, N3 = 16, 6, 5
V3 = np.zeros((N1, N2, N3), np.float32)
V2 = np.zeros((N1, N2), np.float32)
SHORT3 = np.zeros(N1 * N2 * N3 - 1, np.float32)
SHORT2 = np.zeros(N1 * N2 - 1, np.float32)


def _raises(fn, *a, **k):
    with pytest.raises(ValueError):
        fn(*a, **k)


def test_spray_and_filters_refuse_short_dips():
    _raises(ps.pwspray2d, V2, SHORT2, 2, 2, 0.01)
    _raises(filters.pwspray2d_adjoint, np.zeros(N1 * N2 * 5), SHORT2, 2, 2, 0.01, (N1, N2))
    _raises(filters.pwspray2d_adjoint, np.zeros(N1 * N2 * 5 - 1), V2, 2, 2, 0.01, (N1, N2))
    _raises(ps.pwspray3d, V3, SHORT3, V3, 1, 1, 2, 0.01)
    _raises(ps.pwspray3d, V3, V3, SHORT3, 1, 1, 2, 0.01)
    _raises(ps.somean2dc, V2, SHORT2, 2, 2, 0.01)
    _raises(ps.somf2dc, V2, SHORT2, 2, 2, 0.01)
    _raises(ps.somean3dc, V3, SHORT3, V3, 1, 1, 0.01, 2)
    _raises(ps.somf3dc, V3, V3, SHORT3, 1, 1, 0.01, 2)


def test_device_calls_refuse_short_inputs_and_outputs():
    flat3 = torch.zeros(N1 * N2 * N3)
    short = torch.zeros(N1 * N2 * N3 - 1)
    _raises(filters.somf3d_device, flat3, flat3, short, (N1, N2, N3), 1, 1)
    _raises(filters.somf3d_device, flat3, flat3, flat3, (N1, N2, N3), 1, 1, out=short)
    _raises(filters.somean3d_device, short, flat3, flat3, (N1, N2, N3), 1, 1)
    _raises(filters.somean3d_device, flat3, flat3, flat3, (N1, N2, N3), 1, 1, out=short)
    flat2 = torch.zeros(N1 * N2)
    _raises(filters.somf2d_device, flat2, torch.zeros(N1 * N2 - 1), (N1, N2, 1), 2)
    _raises(filters.somf2d_device, flat2, flat2, (N1, N2, 1), 2, out=torch.zeros(3))
    _raises(dip.dip_device, short, (N1, N2, N3))
    _raises(dip.dip_device, flat3, (N1, N2, N3), out=flat3)  # 3-D dips need 2N
    _raises(dip.dip_device, flat3, (N1, N2, N3), init=flat3)
    _raises(dip.dip_device, flat3, (N1, N2, N3), mask=short)
    _raises(smooth.smooth_device, short, (N1, N2, N3))
    _raises(smooth.smooth_device, flat3, (N1, N2, N3), out=short)
    _raises(bandpass.bandpass_device, short, (N1, N2, N3), 0.004)
    _raises(bandpass.bandpass_device, flat3, (N1, N2, N3), 0.004, out=short)
    par = torch.zeros(warmstart.NPARAM)
    _raises(warmstart.cnn_forward, par, short, (N1, N2, N3), 1.0)
    _raises(warmstart.cnn_forward, par[:-1], flat3, (N1, N2, N3), 1.0)
    _raises(warmstart.cnn_forward, par, flat3, (N1, N2, N3), 1.0, out=flat3)
    _raises(warmstart.cnn_grad, par, flat3, (N1, N2, N3), 1.0, flat3)  # target needs 2N
    _raises(warmstart.cnn_grad, par, flat3, (N1, N2, N3), 1.0, torch.zeros(2 * flat3.numel()), wt=short)


def test_dips_and_interpolation_refuse_mismatched_arrays():
    _raises(ps.dip2dc, V2, mask=SHORT2)
    _raises(ps.dip3dc, V3, mask=SHORT3)
    _raises(ps.dip3dc, V3, init=(V3, SHORT3))  # the pair is stacked: 2N - 1 samples
    _raises(interp.soint3dc, V3, SHORT3, V3, V3)
    _raises(interp.soint3dc, V3, V3, V3, SHORT3)
    _raises(interp.soint2dc, V2, V2, SHORT2)
    _raises(interp.soint2dc, V2, SHORT2, V2, prec=1)
    _raises(interp.sint3dc, V3, V3, SHORT3, V3)
    _raises(interp.sint2dc, V2, SHORT2, V2)


def test_batched_calls_refuse_short_dips_and_outputs():
    b = batch.Batch(shapes=[(16, 8), (32, 8)], flat=torch.zeros(16 * 8 + 32 * 8))
    _raises(batch.somf2d_batched, b, torch.zeros(b.ntot - 1))
    _raises(batch.somf2d_batched, b, torch.zeros(b.ntot), out=torch.zeros(b.ntot + 1))
    _raises(batch.dip2d_batched, b, init=torch.zeros(b.ntot - 1), tile=0)
    _raises(batch.dip2d_batched, b, out=torch.zeros(5), tile=0)


def test_correct_sizes_reach_the_library():
    """With matching sizes the checks pass and the call goes on to the
    library, which refuses here only because no GPU is visible."""
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.SeistrError):
        ps.somf3dc(V3, V3, V3, 1, 1, 0.01, 2)
    with pytest.raises(_lib.SeistrError):
        ps.dip3dc(V3, mask=V3, init=(V3, V3))
