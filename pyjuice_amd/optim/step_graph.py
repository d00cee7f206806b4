"""A whole EM training step (clear flows, forward, flow backward, mini-batch EM) recorded once as
a HIP graph and replayed per mini-batch.

The reference captures only the inner-layer loops of ``forward`` / ``backward``
(``record_cudagraph`` / ``apply_cudagraph``, ``tensorcircuit.py:287-309``, ``:481-507``).  Here
the input layers, the batched parameter-flow launch and the EM updates are in the same graph, so
a step is one graph launch.  Every kernel of the step clears buffers with its own fill kernel
(``pjamd_fill_f32``) or stores over them, so the graph records no memset nodes.

The step is recorded with the circuit's state at capture time: parameter flows start from zero
each step (``flows_memory = 0``), the batch size and the input buffer are fixed, and changing the
circuit (``pc.to``, new parameters tensor) needs a new graph.

Input layers that allow it run their EM update inside the flow backward
(``InputLayer.backward(fused_em=...)``, ``pjamd_input_backward_em``): the parameter flows are
still stored, and the step's results are those of backward followed by ``mini_batch_em``.
"""
from __future__ import annotations

from typing import Optional

import torch

__all__ = ["TrainStepGraph"]


class TrainStepGraph:
    """``step = TrainStepGraph(pc, x_example, lr, pseudocount)``; then ``lls = step(x)`` copies
    ``x`` into the static input buffer, replays the recorded step and returns the static
    log-likelihood tensor ``[B, num_roots]`` of that step (overwritten by the next replay)."""

    def __init__(self, pc, x_example: torch.Tensor, lr: float = 0.1, pseudocount: float = 0.1,
                 precision: Optional[str] = None, keep_zero_params: bool = False, fuse_em: bool = True):
        assert x_example.is_cuda and x_example.dim() == 2
        self.pc = pc
        self.fuse_em = fuse_em
        self.lr, self.pseudocount, self.keep_zero = float(lr), float(pseudocount), keep_zero_params
        self.precision = precision
        self.x = x_example.detach().clone().contiguous()
        # one eager step first (on x_example): buffers, launch tables and caches exist before the
        # capture, which records kernels without running them
        self.warmup_lls = self._run()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.lls = self._run()
        torch.cuda.synchronize()

    @staticmethod
    def run_step(pc, x: torch.Tensor, lr: float = 0.1, pseudocount: float = 0.1, precision: Optional[str] = None,
                 keep_zero_params: bool = False, fuse_em: bool = True) -> torch.Tensor:
        """One eager training step, the sequence the graph records: clear the parameter flows,
        forward, flow backward (with the input layers' EM inside it where ``fuse_em`` allows),
        mini-batch EM."""
        with torch.no_grad():
            pc.init_param_flows(flows_memory=0.0)
            lls = pc.forward(x, precision=precision)
            pc.backward(x, flows_memory=0.0, precision=precision,
                        fused_em=(lr, pseudocount, keep_zero_params) if fuse_em else None)
            pc.mini_batch_em(step_size=lr, pseudocount=pseudocount, keep_zero_params=keep_zero_params)
        return lls

    def _run(self) -> torch.Tensor:
        return self.run_step(self.pc, self.x, self.lr, self.pseudocount, self.precision, self.keep_zero,
                             self.fuse_em)

    def __call__(self, x: Optional[torch.Tensor] = None) -> torch.Tensor:
        if x is not None:
            assert x.shape == self.x.shape, "the recorded step has a fixed batch shape"
            self.x.copy_(x)
        self.graph.replay()
        return self.lls
