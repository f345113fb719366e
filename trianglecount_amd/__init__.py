"""trianglecount_amd -- MI355X-native exact triangle counting (LSTC hot path).

C ABI: include/lstc.h, implemented by trianglecount_amd/csrc (gfx950 HIP) and
built into trianglecount_amd/_lib/liblstc.so.  This package is the Python host
side over that ABI; torch supplies device memory and streams only.
"""
from ._lib import LstcError, Tune, GraphDesc, load, check, header_functions  # noqa: F401
from .core import (BatchPlan, Graph, Server, Workspace, clean_orient, count, count_batch,  # noqa: F401
                   generate_chung_lu,
                   generate_er, generate_grid, generate_rmat, halo_extract, halo_extract_parts, partition_random,
                   partition_contig, partition_from_file,
                   stream_handle, autotune, shape_class, tune_resolve, work_stats, KernelTimer, triangle_support, selftest, in_lists, segments, stream_copy)

__version__ = "1.0.0"
