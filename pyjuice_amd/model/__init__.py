from .tensorcircuit import TensorCircuit, compile

__all__ = ["TensorCircuit", "compile"]
