"""sb_predict_noalloc, the prediction entry point that allocates nothing,
against sb_predict on the same inputs: two calls in a row on one device
(the second with fewer individuals and SNPs, so that it reuses the first
call's scratch contents), each within 1e-12 relative of sb_predict; sizes
beyond the scratch limit are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from sodabayes_amd import _lib, synth

pytestmark = pytest.mark.gpu


def host_predict(n, p, bed, freq, eff, mu):
    out = np.zeros(n)
    dp = C.POINTER(C.c_double)
    _lib.check(_lib.load().sb_predict(n, p, np.ascontiguousarray(bed).ctypes.data_as(C.c_void_p), 0,
                                      freq.ctypes.data_as(dp), eff.ctypes.data_as(dp), mu, out.ctypes.data_as(dp)))
    return out


def test_noalloc_predict_equals_sb_predict_on_two_calls():
    import torch
    L = _lib.load()
    for k, (n, p) in enumerate([(6001, 4097), (1203, 777)]):
        bed, _, _ = synth.dataset(n, p, seed=90 + k, nqtl=10)
        x = synth.genotypes(n, p, 90 + k)[0]
        freq = x.sum(0) / (2.0 * n)
        freq[3] = 0.0                    # monomorphic: z = 0
        eff = np.random.default_rng(k).standard_normal(p) * 0.01
        mu = 100.0 + k
        want = host_predict(n, p, bed, freq, eff, mu)
        dbed = torch.from_numpy(np.ascontiguousarray(bed)).cuda()
        dfreq = torch.from_numpy(freq).cuda()
        deff = torch.from_numpy(eff).cuda()
        dpred = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(L.sb_predict_noalloc(n, p, dbed.data_ptr(), dfreq.data_ptr(), deff.data_ptr(), mu,
                                        dpred.data_ptr(), None))
        torch.cuda.synchronize()
        got = dpred.cpu().numpy()
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f"call {k + 1}: n={n} p={p} max relative difference from sb_predict {err:.2e}")
        assert err <= 1e-12


def test_noalloc_predict_refuses_sizes_beyond_its_scratch():
    L = _lib.load()
    big = 2097152 + 1
    rc = L.sb_predict_noalloc(big, 10, 1, 1, 1, 0.0, 1, None)   # refused before any pointer is used
    assert rc == _lib.SB_ERR_ARG
    assert "scratch limit" in L.sb_last_error().decode()
