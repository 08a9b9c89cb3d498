"""The workload and model dictionaries of the tests: those of
viennaemc_amd.workloads (shared with bench.py and the oracle bindings)."""
import os
import sys

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)

from viennaemc_amd.workloads import *  # noqa: E402,F401,F403
from viennaemc_amd.workloads import _splitmix64  # noqa: E402,F401
