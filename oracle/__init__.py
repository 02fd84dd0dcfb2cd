"""CPU oracle for the H/C transport step (test infrastructure only).

This package wraps ``liboracle_mp.so`` (the C restatement in ``mp_oracle.c``)
and, where it was built, ``_ref/libref_phase.so`` (the reference's own
``src/phase`` compiled unmodified).  Only ``tests/``, the smoke check and the
``cpu_baseline`` leg of ``bench.py`` import it; the product package never does.
"""
from .oracle import *  # noqa: F401,F403
