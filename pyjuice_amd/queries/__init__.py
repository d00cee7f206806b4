from .sample import sample

__all__ = ["sample"]
