"""CPU oracle of the reference cloth path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py may use
this package, and only as the checker (or the timed CPU baseline).  The
product (simulator_amd) never imports it.
"""
from .bind import *  # noqa: F401,F403
