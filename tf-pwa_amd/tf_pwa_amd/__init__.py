"""MI355X engine for tf-pwa's unbinned-likelihood hot path.

The drop-in surface mirrors tf_pwa: ``ConfigLoader`` (config.yml -> decay group, variables,
data, ``Model``/``FCN``), ``Model.nll / nll_grad / nll_grad_batch / nll_grad_hessian`` and
``FCN``. Per-event work runs in HIP kernels of ``libtfpwa_amd.so`` through a C ABI.
"""

from .amp import AmplitudeModel
from .config_loader import ConfigLoader
from .data import EventSet
from .model import FCN, Model, get_nll_model, register_nll_model
from .serving import EventPool, FitServer, PoolFull
from .variable import VarsManager

__all__ = ["AmplitudeModel", "ConfigLoader", "EventPool", "EventSet", "FCN", "FitServer",
           "Model", "PoolFull", "VarsManager", "get_nll_model", "register_nll_model"]
