"""afx_true_residual_device_noalloc against afx_true_residual_device through the C ABI: two calls
in a row on one device, the second with a smaller input (fewer columns, a smaller system), each
equal to the allocating entry point's result; the size limit is refused with an error."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cols(s, b, k, seed):
    import torch
    rng = np.random.default_rng(seed)
    B = np.tile(b[:, None], (1, k)) * (1.0 + 0.1 * rng.standard_normal(k))
    X = (rng.standard_normal((s.n, k)) + 1j * rng.standard_normal((s.n, k))) * 1e-3
    return (torch.from_numpy(np.ascontiguousarray(B)).cuda(), torch.from_numpy(np.ascontiguousarray(X)).cuda())


def test_noalloc_residual_matches_allocating_entry_point():
    from actfemtide_amd import afx, mesh
    w = 2 * math.pi / mesh.PERIODS["M2"]
    big, b_big, _ = afx.forward_system(mesh.tide_case(8, 8, 6, seed=1001), w)
    small, b_small, _ = afx.forward_system(mesh.tide_case(4, 4, 6, seed=11), w)
    calls = [(big, b_big, 5, np.array([1.0, 1.03, 0.98, 0.52, 0.48])), (small, b_small, 2, None)]
    for q, (s, b, k, sc) in enumerate(calls):
        dB, dX = _cols(s, b, k, q)
        ref = s.true_residual_device(dB, dX, omega_scale=sc)
        got = s.true_residual_device_noalloc(dB, dX, omega_scale=sc)
        print(f"call {q}: n={s.n} k={k} noalloc {got} allocating {ref}", flush=True)
        assert np.all(ref > 0)
        assert np.max(np.abs(got - ref) / ref) <= 1e-14
    # beyond the scratch's 64 columns: a clear error, nothing launched
    import torch
    dB = torch.zeros((small.n, 65), dtype=torch.complex128, device="cuda")
    res = np.zeros(65)
    rc = afx.lib().afx_true_residual_device_noalloc(small.h, 65, C.c_void_p(dB.data_ptr()), C.c_void_p(dB.data_ptr()),
                                                    None, res.ctypes.data_as(C.c_void_p))
    assert rc == afx.AFX_ERR_ARG and b"64" in afx.lib().afx_last_error()
    big.close()
    small.close()
