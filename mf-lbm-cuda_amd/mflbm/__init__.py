"""Host-side Python helpers of the MI355X colour-gradient LBM step.

The compute path is the HIP shared library ``libmflbm.so`` behind the C ABI declared in
``include/mflbm.h``; ``mflbm.gpu`` binds it with ctypes.  ``config``, ``geometry`` and
``cases`` hold the control-file, geometry-file and workload logic of the host side.
"""
from . import config, geometry, cases  # noqa: F401
