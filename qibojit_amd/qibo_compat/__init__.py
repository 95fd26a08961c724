"""Minimal stand-ins for the parts of ``qibo`` the state-vector path consumes.

qibojit's backends subclass ``qibo.backends.Backend`` and read gate objects
created by ``qibo.gates`` (``src/qibojit/backends/cpu.py:9-12``).  When qibo
is importable, :mod:`qibojit_amd._qibo` uses it; otherwise it uses these
modules, which expose the same names, fields and methods for the gates,
circuits and matrices the path needs (qibo 0.3.x conventions).
"""
