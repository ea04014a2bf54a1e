"""ctypes binding of libdinara_hip.so, the MI355X library (its C ABI is include/dinara.h).

Loading fails loudly when the library is missing or cannot load: this package has no CPU path.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libdinara_hip.so")

UNSUPPORTED_SYMBOLS = -1
OUT_OF_MEMORY = -2
INVALID_COSTS = -3
ABOVE_MAX = -4
OUTSIDE_BAND = -5
INVALID_MODE = -6
INVALID_LENGTH = -7
NO_DEVICE = -8
SCRATCH_FULL = -9

MESSAGES = {
    UNSUPPORTED_SYMBOLS: "a sequence holds a byte the library refuses",
    OUT_OF_MEMORY: "out of memory",
    INVALID_COSTS: "costs no alignment can be searched by",
    ABOVE_MAX: "every alignment costs more than the cap",
    OUTSIDE_BAND: "no alignment stays inside the band",
    INVALID_MODE: "a mode this library does not serve",
    INVALID_LENGTH: "a sequence's length below zero",
    NO_DEVICE: "no such GPU",
    SCRATCH_FULL: "the batch needs more than the fixed device scratch",
}

BATCH_DISTANCES = 0
BATCH_ALIGNMENTS = 1


class CCosts(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "mismatch", "opening", "extension", "opening2", "extension2", "deletion_opening", "deletion_extension",
        "deletion_opening2", "deletion_extension2")]


class CMode(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "kind", "reference_start", "reference_end", "query_start", "query_end", "match_score", "anchor", "zdrop",
        "end_bonus")]


class COptions(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "band_low", "band_high", "max_cost", "eqx", "right_ties", "max_memory")]


class CAlignment(ctypes.Structure):
    _fields_ = [("cost", ctypes.c_int64), ("score", ctypes.c_int64), ("reference_start", ctypes.c_int64),
                ("reference_end", ctypes.c_int64), ("query_start", ctypes.c_int64), ("query_end", ctypes.c_int64),
                ("cigar", ctypes.c_void_p), ("cigar_length", ctypes.c_int64)]


class CScoring(ctypes.Structure):
    _fields_ = [("alphabet", ctypes.c_char_p), ("letters", ctypes.c_int64), ("table", ctypes.c_void_p),
                ("opening", ctypes.c_int64), ("extension", ctypes.c_int64)]


_LIB = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises RuntimeError when it is not built or cannot load."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBRARY):
        raise RuntimeError(f"{LIBRARY} is not built: run dinara-align_amd/build.py (hipcc, gfx950)")
    handle = ctypes.CDLL(LIBRARY)
    i64, vp = ctypes.c_int64, ctypes.c_void_p
    pi64 = ctypes.POINTER(ctypes.c_int64)
    batch_args = [i64, vp, vp, vp, vp]
    handle.dinara_device_count.argtypes = []
    handle.dinara_device_count.restype = i64
    handle.dinara_free.argtypes = [vp]
    handle.dinara_free.restype = None
    handle.dinara_gpu_distances.argtypes = [i64, *batch_args, vp, vp, vp, i64, pi64]
    handle.dinara_gpu_distances.restype = i64
    handle.dinara_gpu_distances_static.argtypes = [i64, *batch_args, vp, vp, vp, i64, pi64]
    handle.dinara_gpu_distances_static.restype = i64
    handle.dinara_gpu_alignments.argtypes = [i64, *batch_args, vp, vp, vp, i64, vp, pi64]
    handle.dinara_gpu_alignments.restype = i64
    handle.dinara_distances.argtypes = [*batch_args, vp, vp, vp, i64, pi64]
    handle.dinara_distances.restype = i64
    handle.dinara_alignments.argtypes = [*batch_args, vp, vp, vp, i64, vp, pi64]
    handle.dinara_alignments.restype = i64
    handle.dinara_gpu_scores.argtypes = [i64, *batch_args, vp, vp, i64, pi64]
    handle.dinara_gpu_scores.restype = i64
    handle.dinara_gpu_scoring_alignments.argtypes = [i64, *batch_args, vp, vp, i64, i64, vp, pi64]
    handle.dinara_gpu_scoring_alignments.restype = i64
    handle.dinara_batch_new.argtypes = [i64, i64, *batch_args, vp, vp, vp, i64, pi64]
    handle.dinara_batch_new.restype = vp
    handle.dinara_batch_run.argtypes = [vp]
    handle.dinara_batch_run.restype = i64
    handle.dinara_batch_distances.argtypes = [vp, pi64]
    handle.dinara_batch_distances.restype = i64
    handle.dinara_batch_alignments.argtypes = [vp, vp, pi64]
    handle.dinara_batch_alignments.restype = i64
    handle.dinara_batch_free.argtypes = [vp]
    handle.dinara_batch_free.restype = None
    handle.dinara_batch_stat.argtypes = [vp, i64]
    handle.dinara_batch_stat.restype = ctypes.c_double
    handle.dinara_scoring_batch_new.argtypes = [i64, i64, *batch_args, vp, vp, i64, i64, pi64]
    handle.dinara_scoring_batch_new.restype = vp
    handle.dinara_scoring_batch_run.argtypes = [vp]
    handle.dinara_scoring_batch_run.restype = i64
    handle.dinara_scoring_batch_scores.argtypes = [vp, pi64]
    handle.dinara_scoring_batch_scores.restype = i64
    handle.dinara_scoring_batch_alignments.argtypes = [vp, vp, vp, vp, pi64]
    handle.dinara_scoring_batch_alignments.restype = i64
    handle.dinara_scoring_batch_free.argtypes = [vp]
    handle.dinara_scoring_batch_free.restype = None
    handle.dinara_scoring_batch_stat.argtypes = [vp, i64]
    handle.dinara_scoring_batch_stat.restype = ctypes.c_double
    handle.dinara_batch_features.argtypes = [vp, vp]
    handle.dinara_batch_features.restype = i64
    handle.dinara_batch_seed.argtypes = [vp, vp]
    handle.dinara_batch_seed.restype = i64
    handle.dinara_batch_seed_learned.argtypes = [vp, vp, ctypes.c_double]
    handle.dinara_batch_seed_learned_by_length.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double]
    handle.dinara_batch_seed_learned_by_length.restype = i64
    handle.dinara_batch_seed_learned.restype = i64
    handle.dinara_batch_set_retry.argtypes = [vp, i64]
    handle.dinara_batch_set_retry.restype = i64
    handle.dinara_batch_set_graph.argtypes = [vp, i64]
    handle.dinara_batch_set_graph.restype = i64
    handle.dinara_batch_set_keep_first.argtypes = [vp, i64]
    handle.dinara_batch_set_keep_first.restype = i64
    handle.dinara_batch_set_side_by_side.argtypes = [vp, i64]
    handle.dinara_batch_set_side_by_side.restype = i64
    handle.dinara_batch_set_persistent.argtypes = [vp, i64]
    handle.dinara_batch_set_persistent.restype = i64
    handle.dinara_batch_set_whole.argtypes = [vp, i64]
    handle.dinara_batch_set_whole.restype = i64
    handle.dinara_batch_set_tuning.argtypes = [vp, i64, vp]
    handle.dinara_batch_set_tuning.restype = i64
    handle.dinara_batch_set_bins.argtypes = [vp, i64]
    handle.dinara_batch_set_bins.restype = i64
    handle.dinara_batch_autotune.argtypes = [vp, vp]
    handle.dinara_batch_autotune.restype = i64
    handle.dinara_mlp_new.argtypes = [i64, ctypes.c_uint64]
    handle.dinara_mlp_new.restype = vp
    handle.dinara_mlp_free.argtypes = [vp]
    handle.dinara_mlp_free.restype = None
    handle.dinara_mlp_parameter_count.argtypes = [vp]
    handle.dinara_mlp_parameter_count.restype = i64
    handle.dinara_mlp_get_weights.argtypes = [vp, vp]
    handle.dinara_mlp_get_weights.restype = i64
    handle.dinara_mlp_set_weights.argtypes = [vp, vp]
    handle.dinara_mlp_set_weights.restype = i64
    handle.dinara_mlp_gradients.argtypes = [vp, i64, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
    handle.dinara_mlp_gradients.restype = i64
    handle.dinara_mlp_train.argtypes = [vp, i64, vp, vp, i64, i64, ctypes.c_double, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_double)]
    handle.dinara_mlp_train.restype = i64
    handle.dinara_mlp_predict.argtypes = [vp, i64, vp, vp]
    handle.dinara_mlp_predict.restype = i64
    _LIB = handle
    return handle


class Texts:
    """A batch's sequences as the C ABI takes them: arrays of first bytes and of lengths, kept alive."""

    def __init__(self, sequences):
        self.items = [s if isinstance(s, bytes) else s.encode("ascii") for s in sequences]
        count = len(self.items)
        self.pointers = (ctypes.c_char_p * max(count, 1))(*self.items)
        self.lengths = (ctypes.c_int64 * max(count, 1))(*[len(s) for s in self.items])

    def args(self):
        return ctypes.cast(self.pointers, ctypes.c_void_p), ctypes.cast(self.lengths, ctypes.c_void_p)
