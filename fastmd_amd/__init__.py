"""fastmd_amd -- MI355X-native bead-spring (Kremer-Grest) MD engine.

Native library: fastmd_amd/libfastmd.so (C ABI in include/fmd/fmd.h), built by
``python -m fastmd_amd.build``; CLI drop-in: fastmd_amd/bin/fastmd <run.conf>.
"""
from .api import (  # noqa: F401
    ERR_ARG, ERR_BUDGET, ERR_NAN, ERR_NBR_OVERFLOW, ERR_PARSE, ERR_STRETCH, LANGEVIN, NPT_NH,
    NVT_NH, OK, TERM_ALL, TERM_ANGLE, TERM_BOND, TERM_LJ, Desc, Engine, FmdError, Server, Topo,
    admission_replay, lib, size_class, system_bytes,
)
from .build import LIB as LIB_PATH  # noqa: F401
