"""FBP filter taps (deeprecon/deeprecon/torch/projector/filter.py:8-30).

Spatial-domain kernels of length 2*num_det - 1 centred on tap num_det - 1:
ramp (Ram-Lak): 1/(4 d^2) at the centre, -1/(pi^2 d^2 n^2) at odd offsets n,
0 at even offsets; Shepp-Logan: -2/(pi^2 d^2 (4 n^2 - 1)).
"""
import math

import torch

FILTERS = ("ramp", "sheeplogan")


def ramp(num_det, det_size, dtype=None):
    n = torch.arange(-(num_det - 1), num_det, dtype=torch.float64)
    taps = torch.zeros_like(n)
    odd = (n.abs().remainder(2) == 1)
    taps[odd] = -1.0 / (math.pi ** 2 * det_size ** 2 * n[odd] ** 2)
    taps[num_det - 1] = 1.0 / (4.0 * det_size ** 2)
    return taps.to(dtype or torch.get_default_dtype())


def sheeplogan(num_det, det_size, dtype=None):
    n = torch.arange(-(num_det - 1), num_det, dtype=torch.float64)
    taps = -2.0 / (math.pi ** 2 * det_size ** 2 * (4.0 * n * n - 1.0))
    return taps.to(dtype or torch.get_default_dtype())


def filter_generate(num_det, det_size, filter_type="ramp", dtype=None):
    if filter_type not in FILTERS:
        raise AssertionError(f"Filter type {filter_type} not supported")
    return {"ramp": ramp, "sheeplogan": sheeplogan}[filter_type](num_det, det_size, dtype)
