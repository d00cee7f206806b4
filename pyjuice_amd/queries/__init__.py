from .marginal import marginal
from .sample import sample

__all__ = ["marginal", "sample"]
