"""Compiled index tensors against the literal known answers of the reference's tests."""
import json
import os

import pytest
import torch

import kat_circuits

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kat_compile.json")
with open(GOLDEN) as f:
    KATS = json.load(f)


def _expected(spec, got):
    if isinstance(spec, dict):
        if "arange" in spec:
            return torch.arange(*spec["arange"])
        if "full" in spec:
            return torch.full_like(got, spec["full"])
        if "bcast" in spec:
            return torch.tensor(spec["bcast"]).expand_as(got)
    return torch.tensor(spec)


@pytest.mark.parametrize("name", sorted(KATS.keys()))
def test_compile_known_answers(name):
    objs = kat_circuits.BUILDERS[name]()
    for expr, spec in KATS[name]:
        got = eval(expr, {"torch": torch}, objs)
        if isinstance(got, torch.Tensor):
            got = got.detach().cpu()
            exp = _expected(spec, got).to(got.dtype)
            assert got.shape == exp.shape, f"{name}: {expr} shape {tuple(got.shape)} != {tuple(exp.shape)}"
            if got.is_floating_point():
                assert torch.allclose(got, exp, atol=1e-6), f"{name}: {expr}"
            else:
                assert torch.equal(got, exp), f"{name}: {expr}\n got {got}\n exp {exp}"
        else:
            assert got == spec, f"{name}: {expr} = {got} != {spec}"


@pytest.mark.parametrize("block_size", [2, 4, 8, 16, 32])
def test_sum_layer_corner_cases(block_size):
    """Index formulas of ``tests/layer/sum_layer_test.py:181-218`` for several block sizes."""
    L = kat_circuits.sum_layer(block_size)["layer"]
    bs = block_size
    assert torch.equal(L.partitioned_nids[0], torch.arange(bs, bs * 7, bs))
    for j, lo in enumerate([1, 1, 3, 3, 5, 5]):
        assert torch.equal(L.partitioned_cids[0][j], torch.arange(bs * lo, bs * (lo + 2)))
        assert torch.equal(L.partitioned_pids[0][j], torch.arange(bs * bs * (2 * j + 1), bs * bs * (2 * j + 3), bs))
        assert torch.equal(L.partitioned_pfids[0][j], torch.arange(bs * bs * 2 * j, bs * bs * (2 * j + 2), bs))
    assert torch.equal(L.partitioned_chids[0], torch.arange(bs, bs * 7, bs))
    par = [[1, 2], [1, 2], [3, 4], [3, 4], [5, 6], [5, 6]]
    ppid = [[1, 3], [2, 4], [5, 7], [6, 8], [9, 11], [10, 12]]
    for j in range(6):
        assert torch.equal(L.partitioned_parids[0][j], torch.tensor(par[j]) * bs)
        assert torch.equal(L.partitioned_parpids[0][j], torch.tensor(ppid[j]) * bs * bs)
