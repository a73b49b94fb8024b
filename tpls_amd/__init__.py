"""MI355X-native TPLS stencil core: host package.

The compute path is the HIP library ``tpls_amd/libtpls_hip.so`` (sources in
``tpls_amd/csrc``) behind the C ABI declared in ``include/tpls_core.h``.
"""
from .config import CaseParams  # noqa: F401
