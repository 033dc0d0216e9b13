"""FBP filter stage on the CPU (no GPU needed).

The reference filters with ``conv2d(sino * weight, taps, padding=(0, D-1))``
(deeprecon/deeprecon/torch/projector/projector.py:207; conv3d with
padding (0, 0, D-1) at :385, :560).  ``pridediff_amd.fbp.filter_sino`` runs
it as a GEMM with the Toeplitz matrix of the taps; these tests pin it
against a direct loop restatement of a zero-padded correlation and against
torch's own conv2d / conv3d, with asymmetric random taps so a transposed or
shifted Toeplitz matrix fails, and check that gradients with respect to the
taps (trainable filters) agree with conv2d's.
"""
import math

import numpy as np
import pytest
import torch

from pridediff_amd.fbp import angular_scale, filter_sino, toeplitz


def _direct_correlation(rows, taps):
    """out[k] = sum_m taps[m] * pad[k + m], pad = D-1 zeros | row | D-1 zeros."""
    d = rows.shape[-1]
    flat = rows.reshape(-1, d)
    out = np.zeros_like(flat)
    for i, row in enumerate(flat):
        pad = np.concatenate([np.zeros(d - 1), row, np.zeros(d - 1)])
        for k in range(d):
            out[i, k] = sum(taps[m] * pad[k + m] for m in range(2 * d - 1))
    return out.reshape(rows.shape)


def test_toeplitz_layout():
    taps = torch.arange(7, dtype=torch.float64)  # D = 4
    t = toeplitz(taps, 4)
    for j in range(4):
        for k in range(4):
            assert t[j, k] == taps[j - k + 3]


@pytest.mark.parametrize("d", [1, 2, 9, 24])
def test_filter_equals_direct_correlation(d):
    rng = np.random.default_rng(d)
    taps = rng.normal(size=2 * d - 1)
    w = rng.uniform(0.5, 1.0, d)
    sino = rng.normal(size=(2, 1, 3, d))
    got = filter_sino(torch.from_numpy(sino), torch.from_numpy(w),
                      torch.from_numpy(taps)[None, None, None, :]).numpy()
    want = _direct_correlation(sino * w, taps)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_filter_equals_torch_conv2d_and_conv3d():
    rng = np.random.default_rng(5)
    d = 31
    taps = torch.from_numpy(rng.normal(size=2 * d - 1))
    w = torch.from_numpy(rng.uniform(0.5, 1.0, d))
    s = torch.from_numpy(rng.normal(size=(3, 1, 7, d)))
    ref = torch.nn.functional.conv2d(s * w, taps[None, None, None, :], padding=(0, d - 1))
    assert torch.allclose(filter_sino(s, w, taps[None, None, None, :]), ref, rtol=1e-12, atol=1e-12)
    w3 = torch.from_numpy(rng.uniform(0.5, 1.0, (5, d)))
    s3 = torch.from_numpy(rng.normal(size=(2, 1, 4, 5, d)))
    ref3 = torch.nn.functional.conv3d(s3 * w3, taps.reshape(1, 1, 1, 1, -1), padding=(0, 0, d - 1))
    got3 = filter_sino(s3, w3, taps.reshape(1, 1, 1, 1, -1))
    assert torch.allclose(got3, ref3, rtol=1e-12, atol=1e-12)


def test_taps_gradient_matches_conv2d():
    rng = np.random.default_rng(6)
    d = 17
    s = torch.from_numpy(rng.normal(size=(2, 1, 5, d)))
    w = torch.from_numpy(rng.uniform(0.5, 1.0, d))
    g = torch.from_numpy(rng.normal(size=(2, 1, 5, d)))
    t1 = torch.from_numpy(rng.normal(size=(1, 1, 1, 2 * d - 1))).requires_grad_(True)
    t2 = t1.detach().clone().requires_grad_(True)
    (filter_sino(s, w, t1) * g).sum().backward()
    (torch.nn.functional.conv2d(s * w, t2, padding=(0, d - 1)) * g).sum().backward()
    assert torch.allclose(t1.grad, t2.grad, rtol=1e-12, atol=1e-12)


def test_ramp_filter_removes_the_mean():
    """Ram-Lak taps sum to 1/(4 D^2) - (2/(pi^2 D^2)) sum_{odd n<N} 1/n^2; on a
    constant row the centre output is that partial sum, whose limit is 0."""
    from pridediff_amd.deeprecon.torch.projector.filter import filter_generate
    d, det = 101, 0.9
    taps = filter_generate(d, det, "ramp", torch.float64)[None, None, None, :]
    out = filter_sino(torch.ones(1, 1, 1, d, dtype=torch.float64), torch.ones(d, dtype=torch.float64),
                      taps)[0, 0, 0, d // 2].item()
    odd = np.arange(1, d // 2 + 1, 2, dtype=np.float64)
    want = 1 / (4 * det * det) - 2 / (math.pi ** 2 * det * det) * np.sum(1.0 / odd ** 2)
    assert out == pytest.approx(want, rel=1e-12)
    assert abs(out) < 0.01 / (det * det)


def test_angular_scale():
    a = torch.arange(720, dtype=torch.float64) * (2 * math.pi / 720)
    assert angular_scale(a, True).item() == pytest.approx(math.pi / 720 * (719 / 719), rel=1e-12)
    assert angular_scale(a, False).item() == pytest.approx(2 * math.pi / 720, rel=1e-12)
