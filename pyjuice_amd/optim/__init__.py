"""Training-loop glue (``pyjuice/optim/optim.py``): ``zero_grad`` clears the parameter flows,
``step`` runs one mini-batch EM update."""
from __future__ import annotations

from typing import Dict, Optional

__all__ = ["CircuitOptimizer", "TrainStepGraph"]


class CircuitOptimizer:
    SUPPORTED_OPTIM_METHODS = ["EM"]

    def __init__(self, pc, base_optimizer=None, method: str = "EM", lr: float = 0.1, pseudocount: float = 0.1,
                 **kwargs):
        self.pc = pc
        self.base_optimizer = base_optimizer
        assert method in self.SUPPORTED_OPTIM_METHODS, f"CircuitOptimizer has no method {method!r}."
        self.method = method
        self.lr = lr
        self.pseudocount = pseudocount

    def zero_grad(self):
        if self.base_optimizer is not None:
            self.base_optimizer.zero_grad()
        self.pc.init_param_flows(flows_memory=0.0)

    def step(self, closure=None):
        if self.base_optimizer is not None:
            self.base_optimizer.step()
        self.pc.mini_batch_em(step_size=self.lr, pseudocount=self.pseudocount)

    def state_dict(self) -> Dict:
        sd = self.base_optimizer.state_dict() if self.base_optimizer is not None else dict()
        sd["pc_states"] = {"method": self.method, "lr": self.lr, "pseudocount": self.pseudocount}
        return sd

    def load_state_dict(self, state_dict: Dict):
        st = dict(state_dict)
        pcs = st.pop("pc_states")
        self.method, self.lr, self.pseudocount = pcs["method"], pcs["lr"], pcs["pseudocount"]
        if self.base_optimizer is not None:
            self.base_optimizer.load_state_dict(st)


from .step_graph import TrainStepGraph  # noqa: E402
