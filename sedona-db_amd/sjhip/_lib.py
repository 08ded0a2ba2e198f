"""ctypes binding of libsjhip.so (include/gpuspatial_hip.h).

The structs mirror the C ABI member for member.  Loading fails loudly when
the library is missing: there is no CPU fallback in this package.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsjhip.so")

c_u32p = ctypes.POINTER(ctypes.c_uint32)


class ArrowSchema(ctypes.Structure):
    pass


ArrowSchema._fields_ = [
    ("format", ctypes.c_char_p),
    ("name", ctypes.c_char_p),
    ("metadata", ctypes.c_char_p),
    ("flags", ctypes.c_int64),
    ("n_children", ctypes.c_int64),
    ("children", ctypes.POINTER(ctypes.POINTER(ArrowSchema))),
    ("dictionary", ctypes.POINTER(ArrowSchema)),
    ("release", ctypes.CFUNCTYPE(None, ctypes.POINTER(ArrowSchema))),
    ("private_data", ctypes.c_void_p),
]


class ArrowArray(ctypes.Structure):
    pass


ArrowArray._fields_ = [
    ("length", ctypes.c_int64),
    ("null_count", ctypes.c_int64),
    ("offset", ctypes.c_int64),
    ("n_buffers", ctypes.c_int64),
    ("n_children", ctypes.c_int64),
    ("buffers", ctypes.POINTER(ctypes.c_void_p)),
    ("children", ctypes.POINTER(ctypes.POINTER(ArrowArray))),
    ("dictionary", ctypes.POINTER(ArrowArray)),
    ("release", ctypes.CFUNCTYPE(None, ctypes.POINTER(ArrowArray))),
    ("private_data", ctypes.c_void_p),
]


class ArrowDeviceArray(ctypes.Structure):
    _fields_ = [
        ("array", ArrowArray),
        ("device_id", ctypes.c_int64),
        ("device_type", ctypes.c_int32),
        ("sync_event", ctypes.c_void_p),
        ("reserved", ctypes.c_int64 * 3),
    ]


ARROW_DEVICE_CPU = 1
ARROW_DEVICE_ROCM = 10


class GpuSpatialRuntimeConfig(ctypes.Structure):
    _fields_ = [
        ("ptx_root", ctypes.c_char_p),
        ("device_id", ctypes.c_int),
        ("use_cuda_memory_pool", ctypes.c_bool),
        ("cuda_memory_pool_init_precent", ctypes.c_int),
    ]


class GpuSpatialRuntime(ctypes.Structure):
    pass


GpuSpatialRuntime._fields_ = [
    ("init", ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(GpuSpatialRuntime), ctypes.POINTER(GpuSpatialRuntimeConfig))),
    ("release", ctypes.CFUNCTYPE(None, ctypes.POINTER(GpuSpatialRuntime))),
    ("get_last_error", ctypes.CFUNCTYPE(ctypes.c_char_p, ctypes.POINTER(GpuSpatialRuntime))),
    ("private_data", ctypes.c_void_p),
]


class GpuSpatialIndexConfig(ctypes.Structure):
    _fields_ = [("runtime", ctypes.POINTER(GpuSpatialRuntime)), ("concurrency", ctypes.c_uint32)]


class SedonaSpatialIndexContext(ctypes.Structure):
    _fields_ = [("private_data", ctypes.c_void_p)]


PROBE_CB = ctypes.CFUNCTYPE(ctypes.c_int, c_u32p, c_u32p, ctypes.c_uint32, ctypes.c_void_p)


class SedonaFloatIndex2D(ctypes.Structure):
    pass


_IX = ctypes.POINTER(SedonaFloatIndex2D)
_CTX = ctypes.POINTER(SedonaSpatialIndexContext)
SedonaFloatIndex2D._fields_ = [
    ("clear", ctypes.CFUNCTYPE(ctypes.c_int, _IX)),
    ("create_context", ctypes.CFUNCTYPE(None, _CTX)),
    ("destroy_context", ctypes.CFUNCTYPE(None, _CTX)),
    ("push_build", ctypes.CFUNCTYPE(ctypes.c_int, _IX, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32)),
    ("finish_building", ctypes.CFUNCTYPE(ctypes.c_int, _IX)),
    ("probe", ctypes.CFUNCTYPE(ctypes.c_int, _IX, _CTX, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32, PROBE_CB,
                               ctypes.c_void_p)),
    ("get_last_error", ctypes.CFUNCTYPE(ctypes.c_char_p, _IX)),
    ("context_get_last_error", ctypes.CFUNCTYPE(ctypes.c_char_p, _CTX)),
    ("release", ctypes.CFUNCTYPE(None, _IX)),
    ("private_data", ctypes.c_void_p),
]


class GpuSpatialRefinerConfig(ctypes.Structure):
    _fields_ = [
        ("runtime", ctypes.POINTER(GpuSpatialRuntime)),
        ("concurrency", ctypes.c_uint32),
        ("compress_bvh", ctypes.c_bool),
        ("pipeline_batches", ctypes.c_uint32),
    ]


class SedonaSpatialRefiner(ctypes.Structure):
    pass


_RF = ctypes.POINTER(SedonaSpatialRefiner)
_SCH = ctypes.POINTER(ArrowSchema)
_ARR = ctypes.POINTER(ArrowArray)
SedonaSpatialRefiner._fields_ = [
    ("clear", ctypes.CFUNCTYPE(ctypes.c_int, _RF)),
    ("init_build_schema", ctypes.CFUNCTYPE(ctypes.c_int, _RF, _SCH)),
    ("push_build", ctypes.CFUNCTYPE(ctypes.c_int, _RF, _ARR)),
    ("finish_building", ctypes.CFUNCTYPE(ctypes.c_int, _RF)),
    ("refine", ctypes.CFUNCTYPE(ctypes.c_int, _RF, _SCH, _ARR, ctypes.c_int, c_u32p, c_u32p, ctypes.c_uint32,
                                c_u32p)),
    ("get_last_error", ctypes.CFUNCTYPE(ctypes.c_char_p, _RF)),
    ("release", ctypes.CFUNCTYPE(None, _RF)),
    ("private_data", ctypes.c_void_p),
]


class SjhipJoinResult(ctypes.Structure):
    pass


SjhipJoinResult._fields_ = [
    ("build_indices", ctypes.c_void_p),
    ("probe_indices", ctypes.c_void_p),
    ("size", ctypes.c_uint64),
    ("candidates", ctypes.c_uint64),
    ("fallback_tests", ctypes.c_uint64),
    ("segment_tests", ctypes.c_uint64),
    ("refine_candidates", ctypes.c_uint64),
    ("stage_ms", ctypes.c_double * 6),
    ("release", ctypes.CFUNCTYPE(None, ctypes.POINTER(SjhipJoinResult))),
    ("private_data", ctypes.c_void_p),
]


class SjhipJoinOptions(ctypes.Structure):
    _fields_ = [
        ("predicate", ctypes.c_int32),
        ("distance", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("sort_output", ctypes.c_int32),
        ("sort_probes", ctypes.c_int32),
        ("profile", ctypes.c_int32),
        ("refine_order", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
    ]


class SjhipTuneResult(ctypes.Structure):
    _fields_ = [("leaf", ctypes.c_int32), ("refine_order", ctypes.c_int32), ("pip_waves", ctypes.c_int32),
                ("evaluations", ctypes.c_int32), ("best_ms", ctypes.c_double), ("seg_leaf", ctypes.c_int32)]


class SjhipServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("memory_budget", ctypes.c_uint64), ("graphs", ctypes.c_int32)]


class SjhipPartitionStats(ctypes.Structure):
    _fields_ = [("partitions", ctypes.c_uint32), ("max_partition_rows", ctypes.c_uint64),
                ("max_partition_bytes", ctypes.c_uint64), ("pairs", ctypes.c_uint64), ("build_ms", ctypes.c_double),
                ("join_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


class SjhipRequestTiming(ctypes.Structure):
    _fields_ = [("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("estimate_bytes", ctypes.c_uint64)]


# Every symbol the public headers declare (checked by the CPU test suite).
EXPORTED = [
    "GpuSpatialRuntimeCreate", "GpuSpatialIndexFloat2DCreate", "GpuSpatialRefinerCreate",
    "SjhipExtVersion", "SjhipRefinerRefineDistance", "SjhipRefinerKnn", "SjhipJoinDevice",
    "SjhipJoinResultCopyToHost", "SjhipIndexBuildFromRefiner", "SjhipRefinerPushBuildDevice", "SjhipLastError",
    "SjhipComputeRects", "SjhipDeviceAlloc", "SjhipDeviceFree", "SjhipCopyToDevice", "SjhipCopyToHost",
    "SjhipSetDevice", "SjhipCopyKernel",
    "SjhipServerCreate", "SjhipServerSubmit", "SjhipServerWait", "SjhipServerDestroy", "SjhipRefinerKnnDevice",
    "SjhipAutotune", "SjhipRefinerSetSegmentLeaf", "SjhipIndexSetTuning", "SjhipRefinerSetPrecision",
    "SjhipHostRaycast", "SjhipHostRaycastF32", "SjhipHostCollinear", "SjhipHostSegmentIntersects",
    "SjhipHostOrient2d", "SjhipHostHypot", "SjhipHostCoordToLine", "SjhipHostRaycastBatch", "SjhipHostHypotBatch",
    "SjhipTestingGraphSortPairs", "SjhipJoinDeviceNoAlloc", "SjhipRefinerKnnTies", "SjhipHostSegmentWithinF32",
    "SjhipRefinerPairDistances", "SjhipIndexSetFilterMode", "SjhipJoinHostStream", "SjhipPartitionedJoin",
    "SjhipHostSegmentsCrossF32",
]

_lib = None


def lib():
    """The loaded libsjhip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsjhip.so not found at {LIB_PATH}: build it with `make -C sedona-db_amd` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        L.GpuSpatialRuntimeCreate.argtypes = [ctypes.POINTER(GpuSpatialRuntime)]
        L.GpuSpatialRuntimeCreate.restype = None
        L.GpuSpatialIndexFloat2DCreate.argtypes = [ctypes.POINTER(SedonaFloatIndex2D),
                                                   ctypes.POINTER(GpuSpatialIndexConfig)]
        L.GpuSpatialRefinerCreate.argtypes = [ctypes.POINTER(SedonaSpatialRefiner),
                                              ctypes.POINTER(GpuSpatialRefinerConfig)]
        L.SjhipRefinerRefineDistance.argtypes = [_RF, _SCH, _ARR, ctypes.c_double, c_u32p, c_u32p, ctypes.c_uint32,
                                                 c_u32p]
        L.SjhipRefinerKnn.argtypes = [_RF, _SCH, _ARR, ctypes.c_uint32, c_u32p, ctypes.POINTER(ctypes.c_double)]
        L.SjhipRefinerPairDistances.argtypes = [_RF, _SCH, _ARR, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                ctypes.c_void_p]
        L.SjhipRefinerKnnTies.argtypes = [_RF, _SCH, _ARR, ctypes.c_uint32, ctypes.c_int32, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_uint64)]
        L.SjhipJoinDevice.argtypes = [_IX, _RF, _SCH, ctypes.POINTER(ArrowDeviceArray),
                                      ctypes.POINTER(SjhipJoinOptions), ctypes.POINTER(SjhipJoinResult)]
        L.SjhipJoinResultCopyToHost.argtypes = [ctypes.POINTER(SjhipJoinResult), c_u32p, c_u32p]
        L.SjhipPartitionedJoin.argtypes = [ctypes.POINTER(GpuSpatialRuntime), _SCH, _ARR, _SCH,
                                           ctypes.POINTER(ArrowDeviceArray), ctypes.POINTER(SjhipJoinOptions),
                                           ctypes.c_uint64, PROBE_CB, ctypes.c_void_p,
                                           ctypes.POINTER(SjhipPartitionStats)]
        L.SjhipJoinHostStream.argtypes = [_IX, _RF, _SCH, _ARR, ctypes.POINTER(SjhipJoinOptions), ctypes.c_int64,
                                          ctypes.c_int32, PROBE_CB, ctypes.c_void_p]
        L.SjhipJoinDeviceNoAlloc.argtypes = [_IX, _RF, _SCH, ctypes.POINTER(ArrowDeviceArray),
                                             ctypes.POINTER(SjhipJoinOptions), ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        L.SjhipServerCreate.argtypes = [_IX, _RF, ctypes.POINTER(SjhipServerConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.SjhipServerSubmit.argtypes = [ctypes.c_void_p, _SCH, ctypes.POINTER(ArrowDeviceArray),
                                        ctypes.POINTER(SjhipJoinOptions), ctypes.POINTER(ctypes.c_uint64)]
        L.SjhipServerWait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(SjhipJoinResult),
                                      ctypes.POINTER(SjhipRequestTiming)]
        L.SjhipServerDestroy.argtypes = [ctypes.c_void_p]
        L.SjhipAutotune.argtypes = [_IX, _RF, _SCH, ctypes.POINTER(ArrowDeviceArray), ctypes.POINTER(SjhipJoinOptions),
                                    ctypes.POINTER(SjhipTuneResult)]
        L.SjhipRefinerSetSegmentLeaf.argtypes = [_RF, ctypes.c_int32]
        L.SjhipIndexSetTuning.argtypes = [_IX, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.SjhipRefinerSetPrecision.argtypes = [_RF, ctypes.c_int32]
        L.SjhipIndexSetFilterMode.argtypes = [_IX, ctypes.c_int32]
        L.SjhipRefinerKnnDevice.argtypes = [_RF, _SCH, ctypes.POINTER(ArrowDeviceArray), ctypes.c_uint32,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.SjhipIndexBuildFromRefiner.argtypes = [_IX, _RF]
        L.SjhipRefinerPushBuildDevice.argtypes = [_RF, _SCH, ctypes.POINTER(ArrowDeviceArray)]
        L.SjhipLastError.restype = ctypes.c_char_p
        L.SjhipComputeRects.argtypes = [_SCH, _ARR, ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        L.SjhipDeviceAlloc.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
        L.SjhipDeviceFree.argtypes = [ctypes.c_void_p]
        L.SjhipCopyToDevice.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.SjhipCopyToHost.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.SjhipSetDevice.argtypes = [ctypes.c_int]
        L.SjhipCopyKernel.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
        d6 = [ctypes.c_double] * 6
        L.SjhipHostRaycast.argtypes = d6
        L.SjhipHostRaycastF32.argtypes = [ctypes.c_float] * 7
        L.SjhipHostCollinear.argtypes = d6
        L.SjhipHostSegmentIntersects.argtypes = [ctypes.c_double] * 8
        L.SjhipHostOrient2d.argtypes = d6
        L.SjhipHostHypot.argtypes = [ctypes.c_double] * 2
        L.SjhipHostHypot.restype = ctypes.c_double
        L.SjhipHostCoordToLine.argtypes = d6
        L.SjhipHostCoordToLine.restype = ctypes.c_double
        L.SjhipHostRaycastBatch.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.SjhipHostHypotBatch.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.SjhipHostSegmentWithinF32.argtypes = [ctypes.c_float] * 8 + [ctypes.c_double]
        L.SjhipHostSegmentsCrossF32.argtypes = [ctypes.c_float] * 9
        L.SjhipTestingGraphSortPairs.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                                                         ctypes.c_int]
        _lib = L
    return _lib
