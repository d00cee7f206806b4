from .tensorcircuit import TensorCircuit, compile
from .autograd import log_likelihood

__all__ = ["TensorCircuit", "compile", "log_likelihood"]
