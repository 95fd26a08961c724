"""Configuration constants of the qibo interface (qibo.config)."""

# Shot count from which frequencies are sampled with the Metropolis chains of
# ``measure_frequencies`` (used at src/qibojit/backends/cpu.py:384).  qibo
# 0.3.x sets it to 100000.
SHOT_METROPOLIS_THRESHOLD = 100000

# Default absolute tolerance of qibo's ``assert_allclose``.
PRECISION_TOL = 1e-8


def raise_error(exception, message=None):
    """Raise ``exception`` with ``message`` (qibo.config.raise_error)."""
    raise exception(message)
