"""Re-export of the package workload builder for tests."""

from fdtdx_amd.scenarios import *  # noqa: F401,F403
from fdtdx_amd.scenarios import build  # noqa: F401
