"""Block-size autotuning: candidate generation (CPU) and a measured choice on the device."""
import pytest
import torch

import parity_utils as U
from pyjuice_amd import autotune


def test_block_size_candidates():
    assert autotune.block_size_candidates(256) == [256, 128, 64]
    assert autotune.block_size_candidates(512, smallest=32) == [512, 256, 128, 64, 32]
    assert autotune.block_size_candidates(48) == [48] or autotune.block_size_candidates(48) == []
    assert autotune.block_size_candidates(32) == [32]


@pytest.mark.gpu
def test_tune_block_size_measures_every_candidate():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import numpy as np
    import pyjuice_amd as juice
    import pyjuice_amd.nodes.distributions as dists
    U.say("[autotune] hclt 12 vars, 128 latents")
    ns0, x, cards = U.hclt_categorical(num_vars=12, K=8, L=128)
    xt = torch.from_numpy(x).float()

    def make_nodes(bs):
        torch.manual_seed(0)
        return juice.structures.HCLT(xt, num_latents=128, block_size=bs, input_dist=dists.Categorical(num_cats=8),
                                     chunk_size=32)

    xb = torch.from_numpy(U.random_data(cards, 64, seed=2)).cuda()
    best, ms = autotune.tune_block_size(make_nodes, xb, [128, 64, 32], steps=2, warmup=1, precision="bf16")
    assert set(ms) == {128, 64, 32} and best in ms and all(v > 0 for v in ms.values())
    # the first candidate is kept unless another is faster by more than the 3 % margin
    assert best == 128 or ms[best] < 0.97 * ms[128]


def test_margin_keeps_first_candidate(monkeypatch):
    """The choice rule alone (no device): a later candidate must beat the kept one by the margin."""
    import pyjuice_amd
    times = iter([10.0, 9.8, 9.0])
    monkeypatch.setattr(autotune, "_time_steps", lambda step, steps, warmup: next(times))

    class _PC:
        def to(self, device):
            return self

    monkeypatch.setattr(pyjuice_amd, "compile", lambda ns, verbose=False: _PC())
    import pyjuice_amd.optim as optim
    monkeypatch.setattr(optim, "TrainStepGraph", lambda pc, x, **kw: None)
    monkeypatch.setattr(torch.cuda, "empty_cache", lambda: None)
    best, ms = autotune.tune_block_size(lambda bs: None, torch.zeros(2, 2), [256, 128, 64])
    assert ms == {256: 10.0, 128: 9.8, 64: 9.0} and best == 64      # 9.0 < 0.97 * 10.0; 9.8 is not
