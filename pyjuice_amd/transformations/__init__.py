from .prune import prune_sum_blocks

__all__ = ["prune_sum_blocks"]
