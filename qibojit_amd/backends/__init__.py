"""Plugin entry: platform registry and loader (src/qibojit/backends/__init__.py).

``MetaBackend.load("hip")`` constructs the HIP backend; "numba", "cupy" and
"cuquantum" are forwarded to an installed qibojit, as its own MetaBackend
loads them.  A platform that cannot be constructed raises ImportError
(ModuleNotFoundError included), which ``list_available`` reports as False.
"""

PLATFORMS = ("numba", "cupy", "cuquantum", "hip")


def _upstream(platform):
    from qibojit.backends import MetaBackend as UpstreamMeta  # pylint: disable=import-outside-toplevel

    return UpstreamMeta.load(platform)


class MetaBackend:
    """Meta-backend class which takes care of loading the qibojit backends."""

    @staticmethod
    def load(platform=None):
        """Load the backend for ``platform`` (default: hip if a device exists)."""
        if platform == "hip" or platform is None:
            from qibojit_amd.backends.hip import HipBackend  # pylint: disable=import-outside-toplevel

            return HipBackend()
        if platform in ("numba", "cupy", "cuquantum"):
            return _upstream(platform)
        raise ValueError(f"Unknown platform {platform}; available: {PLATFORMS}.")

    def list_available(self):
        """Lists all the available qibojit backends."""
        available = {}
        for platform in PLATFORMS:
            try:
                MetaBackend.load(platform=platform)
                available[platform] = True
            except Exception:  # pylint: disable=broad-except
                available[platform] = False
        return available
