"""sarlabtdc -- time-domain correlation SAR processor for MI355X.

Drop-in for the reference package (py-sarlabtdc/sarlabtdc/__init__.py:31,
which re-exports the compiled ``_sarlabtdc`` module): ``TdcProcessor`` keeps
the reference's constructor, ``start``, ``setRawData``, ``setFocusGrid`` and
``getFocusedImage``. The compiled module is required: importing this package
without the built HIP library raises ImportError.
"""
from ._sarlabtdc import *  # noqa: F401,F403
from ._sarlabtdc import FORM_PLAIN, FORM_PRECISION, Server, TdcProcessor, device_count

__all__ = ["TdcProcessor", "Server", "device_count", "FORM_PLAIN", "FORM_PRECISION"]
