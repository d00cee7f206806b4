from .hclt import HCLT
from .pd import PD
from .hmm import HMM
from .compilation import BayesianTreeToHiddenRegionGraph

__all__ = ["HCLT", "PD", "HMM", "BayesianTreeToHiddenRegionGraph"]
