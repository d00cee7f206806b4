from .tensorcircuit import TensorCircuit, compile
from .autograd import log_likelihood
from .top_down import eval_partition_fn, eval_top_down_probs

__all__ = ["TensorCircuit", "compile", "log_likelihood", "eval_top_down_probs", "eval_partition_fn"]
