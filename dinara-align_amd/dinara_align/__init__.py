"""dinara_align on an AMD Instinct MI355X: the reference package's batch calls, run by hand-written HIP kernels.

A drop-in for the reference's ``dinara_align`` on its batched device path (reference
python/dinara_align/__init__.py:435-621): the same value types (``Costs``, ``Mode``, ``Band``, ``Scoring``,
``Alignment``, in ``types.py``), the same ``distance``/``align``/``score``/``distances``/``alignments`` calls
with their argument meanings and ``ValueError`` refusals, plus the Mojo API's batch ``scores`` and device
``placement`` (reference api.mojo:843-1001,1216-1252). Every call reaches ``libdinara_hip.so`` through ctypes
and runs on the GPU; there is no CPU path, and a missing library raises.

>>> import dinara_align as da
>>> da.distance("ACGTACGTTTGCA", "ACGTCGTTTTGCA")
2
>>> found = da.align("ACGTACGTTTGCA", "ACGTCGTTTTGCA", da.Costs.affine(4, 6, 2))
>>> found.cost, found.cigar
(12, '4=3X6=')
"""

from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Union

from . import _lib
from .types import (UNBOUNDED, Alignment, Anchor, Band, Costs, Mode, Placement, Scoring, Text, as_bytes,
                    _integers)

__all__ = [
    "Alignment",
    "Anchor",
    "Band",
    "Costs",
    "Mode",
    "Placement",
    "Scoring",
    "UNBOUNDED",
    "align",
    "alignments",
    "distance",
    "distances",
    "score",
    "scores",
    "device_count",
]

_INT64_MIN, _INT64_MAX = -(1 << 63), (1 << 63) - 1


def _fail(code: int):
    """The library's batch-wide or per-pair code as the ValueError the reference package raises."""
    raise ValueError(_lib.MESSAGES.get(code, f"the library returned code {code}"))


def _options(band: Optional[Band], max_cost: Optional[int], eqx: bool, ties: str,
             max_memory: Optional[int] = None) -> _lib.COptions:
    """A dinara_options struct: UNBOUNDED band edges become the struct's no-band values, no cap is -1."""
    if ties not in ("left", "right"):
        raise ValueError(f"ties is 'left' or 'right', not {ties!r}")
    band = band if band is not None else Band()
    _integers(band_low=band.low, band_high=band.high)
    if max_cost is not None:
        _integers(max_cost=max_cost)
    if max_memory is not None:
        _integers(max_memory=max_memory)
    return _lib.COptions(
        band_low=_INT64_MIN if band.low <= -UNBOUNDED else band.low,
        band_high=_INT64_MAX if band.high >= UNBOUNDED else band.high,
        max_cost=-1 if max_cost is None else max_cost,
        eqx=int(bool(eqx)),
        right_ties=int(ties == "right"),
        max_memory=0 if max_memory is None else max_memory)


def _where(placement: Optional[Placement], threads: int = 0) -> Placement:
    if placement is not None:
        return placement
    _integers(threads=threads)
    return Placement(0, threads)


def device_count() -> int:
    """How many GPUs the library sees."""
    return int(_lib.lib().dinara_device_count())


def _texts(references: Sequence[Text], queries: Sequence[Text]):
    refs = _lib.Texts([as_bytes(item) for item in references])
    qrys = _lib.Texts([as_bytes(item) for item in queries])
    if len(refs.items) != len(qrys.items):
        raise ValueError(f"{len(refs.items)} references but {len(qrys.items)} queries")
    return refs, qrys, len(refs.items)


def _settle(codes, max_cost: Optional[int]) -> list:
    """Per-pair results: a value stays, a pair over the cap becomes None, any other code raises."""
    if max_cost is not None and max_cost < 0:
        # A negative cap admits nothing; the call above still checked the whole batch (reference api.mojo:710-711).
        return [None] * len(codes)
    out = []
    for item in codes:
        if not isinstance(item, int) or item >= 0:
            out.append(item)
        elif item == _lib.ABOVE_MAX and max_cost is not None:
            out.append(None)
        else:
            _fail(item)
    return out


def distances(references: Sequence[Text], queries: Sequence[Text], costs: Costs = Costs(), mode: Mode = Mode.GLOBAL,
              *, band: Optional[Band] = None, max_cost: Optional[int] = None, threads: int = 0,
              placement: Optional[Placement] = None) -> list:
    """Each pair's least cost, or None above `max_cost` (reference python/dinara_align/__init__.py:545-563)."""
    options = _options(band, max_cost, True, "left")
    where = _where(placement, threads)
    refs, qrys, count = _texts(references, queries)
    results = (ctypes.c_int64 * max(count, 1))()
    c_costs, c_mode = costs.as_c(), mode.as_c()
    code = _lib.lib().dinara_gpu_distances(where.gpu_id, count, *refs.args(), *qrys.args(), ctypes.byref(c_costs),
                                           ctypes.byref(c_mode), ctypes.byref(options), where.threads, results)
    if code != 0:
        _fail(code)
    return _settle(list(results[:count]), max_cost)


def _run_alignments(references, queries, costs, mode: Mode, options, where: Placement, eqx: bool) -> list:
    """Alignment objects, or the per-pair code where a pair failed."""
    refs, qrys, count = _texts(references, queries)
    found = (_lib.CAlignment * max(count, 1))()
    statuses = (ctypes.c_int64 * max(count, 1))()
    handle = _lib.lib()
    c_mode = mode.as_c()
    if isinstance(costs, Scoring):
        c_scoring, keep_alive = costs.as_c()
        code = handle.dinara_gpu_scoring_alignments(where.gpu_id, count, *refs.args(), *qrys.args(),
                                                    ctypes.byref(c_scoring), ctypes.byref(c_mode), int(bool(eqx)),
                                                    where.threads, found, statuses)
    else:
        c_costs = costs.as_c()
        code = handle.dinara_gpu_alignments(where.gpu_id, count, *refs.args(), *qrys.args(), ctypes.byref(c_costs),
                                            ctypes.byref(c_mode), ctypes.byref(options), where.threads, found,
                                            statuses)
    if code != 0:
        _fail(code)
    out = []
    for i in range(count):
        if statuses[i] != 0:
            out.append(int(statuses[i]))
            continue
        item = found[i]
        text = ctypes.string_at(item.cigar, item.cigar_length).decode("ascii") if item.cigar else ""
        handle.dinara_free(item.cigar)
        out.append(Alignment(cost=item.cost, score=item.score, cigar=text, reference_start=item.reference_start,
                             reference_end=item.reference_end, query_start=item.query_start,
                             query_end=item.query_end))
    return out


def alignments(references: Sequence[Text], queries: Sequence[Text], costs: Union[Costs, Scoring] = Costs(),
               mode: Mode = Mode.GLOBAL, *, band: Optional[Band] = None, max_cost: Optional[int] = None,
               ties: str = "left", eqx: bool = True, max_memory: Optional[int] = None, threads: int = 0,
               placement: Optional[Placement] = None) -> list:
    """Each pair's optimal alignment, the one `ties` names among equals, or None above `max_cost` (reference
    python/dinara_align/__init__.py:566-588; under a Scoring, reference api.mojo:895-1001). `max_memory` is
    checked and passed on; the device traces every pair whole, as the reference does for a pair whose kept
    fronts stay within it."""
    options = _options(band, max_cost, eqx, ties, max_memory)
    found = _run_alignments(references, queries, costs, mode, options, _where(placement, threads), eqx)
    return _settle(found, max_cost)


def scores(references: Sequence[Text], queries: Sequence[Text], scoring: Scoring, mode: Mode = Mode.GLOBAL, *,
           placement: Optional[Placement] = None) -> List[int]:
    """Each pair's best score under `scoring`, global or local (reference api.mojo:843-892)."""
    where = _where(placement)
    refs, qrys, count = _texts(references, queries)
    results = (ctypes.c_int64 * max(count, 1))()
    c_scoring, keep_alive = scoring.as_c()
    c_mode = mode.as_c()
    code = _lib.lib().dinara_gpu_scores(where.gpu_id, count, *refs.args(), *qrys.args(), ctypes.byref(c_scoring),
                                        ctypes.byref(c_mode), where.threads, results)
    if code != 0:
        _fail(code)
    return list(results[:count])


def distance(reference: Text, query: Text, costs: Costs = Costs(), mode: Mode = Mode.GLOBAL, *,
             band: Optional[Band] = None, max_cost: Optional[int] = None,
             placement: Optional[Placement] = None) -> Optional[int]:
    """One pair's least cost, or None above `max_cost` (reference python/dinara_align/__init__.py:435-447)."""
    return distances([reference], [query], costs, mode, band=band, max_cost=max_cost, placement=placement)[0]


def align(reference: Text, query: Text, costs: Union[Costs, Scoring] = Costs(), mode: Mode = Mode.GLOBAL, *,
          band: Optional[Band] = None, max_cost: Optional[int] = None, ties: str = "left", eqx: bool = True,
          max_memory: Optional[int] = None, placement: Optional[Placement] = None) -> Optional[Alignment]:
    """One pair's optimal alignment, or None above `max_cost` (reference python/dinara_align/__init__.py:450-472)."""
    return alignments([reference], [query], costs, mode, band=band, max_cost=max_cost, ties=ties, eqx=eqx,
                      max_memory=max_memory, placement=placement)[0]


def score(reference: Text, query: Text, costs: Union[Costs, Scoring] = Costs(), mode: Mode = Mode.GLOBAL, *,
          band: Optional[Band] = None, placement: Optional[Placement] = None) -> int:
    """One pair's best score: a Scoring's own, or minus the least cost under Costs (reference
    python/dinara_align/__init__.py:518-531)."""
    if isinstance(costs, Scoring):
        return scores([reference], [query], costs, mode, placement=placement)[0]
    return -distance(reference, query, costs, mode, band=band, placement=placement)
