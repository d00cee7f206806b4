"""``pjamd_sum_forward_split_noalloc`` (fixed device-resident workspace, nothing allocated)
against ``pjamd_sum_forward_split`` with a caller-owned workspace, through the C ABI: two calls in
a row, the second on a smaller batch, each bit-identical to the workspace form (same kernels).
Also: a request beyond the fixed workspace is refused before any launch."""
import numpy as np
import pytest
import torch

import parity_utils as U
from pyjuice_amd import _native as N
from pyjuice_amd.layer.sum_layer import PARAMS_ALIGNED
from pyjuice_amd.model import TensorCircuit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _split_layer(pc):
    """(layer, forward table, child-run flags) of the transition sum layer (1024-row blocks)."""
    for g in pc.inner_layer_groups:
        for layer in g:
            if layer.is_sum() and layer.block_size == 1024:
                run = layer._fw_child_run[0] | (PARAMS_ALIGNED if layer._fw_aligned[0] else 0)
                return layer, layer._dev["fw"][0], run
    raise AssertionError("no 1024-row sum layer")


@pytest.mark.parametrize("precision", [1, 2])
def test_noalloc_split_forward_matches_workspace_form(precision, device):
    U.say(f"[no-alloc split forward] precision={precision}")
    ns, cards = U.hmm_small(T=4, L=1024, V=16)
    pc = TensorCircuit(ns, precision="bf16")
    pc.to(device)
    layer, (nids, cids, pids, _pfids, nch, _tot), run = _split_layer(pc)
    assert run & PARAMS_ALIGNED and (run & ~PARAMS_ALIGNED) >= 128
    lib = N.lib()
    bs, n_blocks, K = layer.block_size, nids.numel(), cids.size(1)
    nsplit = 4
    for B in (256, 64):                      # the second call on a smaller batch
        X = U.random_data(cards, B, seed=B)
        pc(torch.from_numpy(X).to(device))   # element rows of the layer below
        slab = pc._slab([i for i, g in enumerate(pc.inner_layer_groups) if layer in list(g)][0])
        rows = n_blocks * bs
        ws = torch.empty([nsplit * (rows + rows // 64) * B], device=device, dtype=torch.float32)
        out_a = torch.full_like(pc.node_mars, float("nan"))
        out_b = torch.full_like(pc.node_mars, float("nan"))
        s = N.stream()
        rc = lib.pjamd_sum_forward_split(N.ptr(out_a), N.ptr(slab), N.ptr(pc.params), N.ptr(nids), N.ptr(cids),
                                         N.ptr(pids), N.ptr(nch), n_blocks, K, bs, B, run, nsplit, N.ptr(ws),
                                         precision, s)
        assert rc == 0, rc
        rc = lib.pjamd_sum_forward_split_noalloc(N.ptr(out_b), N.ptr(slab), N.ptr(pc.params), N.ptr(nids),
                                                 N.ptr(cids), N.ptr(pids), N.ptr(nch), n_blocks, K, bs, B, run,
                                                 nsplit, precision, s)
        assert rc == 0, rc
        torch.cuda.synchronize()
        a0 = int(nids.min())
        ra = out_a[a0:a0 + rows].cpu().numpy()
        rb = out_b[a0:a0 + rows].cpu().numpy()
        assert np.isfinite(ra).all()
        assert np.array_equal(ra, rb)
        # and both equal the layer's own forward on these element rows
        assert np.array_equal(ra, pc.node_mars[a0:a0 + rows].cpu().numpy()) or \
            np.abs(ra - pc.node_mars[a0:a0 + rows].cpu().numpy()).max() <= 4e-3


def test_noalloc_split_forward_refuses_oversized_requests(device):
    lib = N.lib()
    null = None
    # 16 splits x (1024 rows + 16) x 4096 columns needs 68 M floats > the 8 M-float workspace
    rc = lib.pjamd_sum_forward_split_noalloc(null, null, null, null, null, null, null, 1, 1024, 1024, 4096, 1024,
                                             16, 1, null)
    assert rc == -3
    assert b"workspace" in lib.pjamd_error_string(rc)
