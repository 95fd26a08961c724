"""MI355X-native state-vector backend with qibojit's plugin interface."""

__version__ = "0.1.0"

from qibojit_amd.backends import MetaBackend  # noqa: E402

__all__ = ["MetaBackend", "__version__"]
