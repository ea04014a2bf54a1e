"""dinara_gpu_distances_static, the unit-cost distances call that allocates no device memory, against
dinara_gpu_distances (the allocating call on the same kernels), through the C ABI.

Two calls in a row share the device's fixed scratch: the second, smaller batch must not see the first's contents.
A batch larger than the scratch is refused on the host before anything runs, and costs the bit-parallel path does
not serve are refused as DINARA_INVALID_MODE.
"""

import ctypes

import pytest

from dinara_align import _lib, synthetic

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def call(name, refs, qrys, costs=None, options=None):
    h = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    out = (ctypes.c_int64 * max(len(refs), 1))()
    code = getattr(h, name)(0, len(refs), *R.args(), *Q.args(), ctypes.byref(costs) if costs else None, None,
                            ctypes.byref(options) if options else None, 4, out)
    return code, list(out)[: len(refs)]


def both(refs, qrys, costs=None, options=None):
    print(f"[static] {len(refs)} pairs, longest {max(max(len(r), len(q)) for r, q in zip(refs, qrys))}", flush=True)
    code, want = call("dinara_gpu_distances", refs, qrys, costs, options)
    assert code == 0
    code, found = call("dinara_gpu_distances_static", refs, qrys, costs, options)
    assert code == 0
    return want, found


def test_static_distances_equal_the_allocating_call_twice_in_a_row():
    refs, qrys = synthetic.uneven(4000, seed=91, clip=(50, 3000))
    want, found = both(refs, qrys)
    assert found == want
    # A smaller batch after it: the scratch keeps the first call's queues and results, which must not leak in.
    short_refs, short_qrys = synthetic.edit_short(700, seed=92)
    want, found = both(short_refs, short_qrys)
    assert found == want


def test_static_distances_scaled_costs_caps_and_odd_pairs():
    refs, qrys = synthetic.uneven(1500, seed=93, clip=(1, 1200))
    # Empty and one-sided pairs, a byte outside ACGT, a refused byte.
    refs += [b"", b"", b"ACGT", b"NNNNACGTNN", b"AC\xffGT"]
    qrys += [b"", b"ACG", b"", b"ACGTTTT", b"ACGT"]
    scaled = _lib.CCosts(3, 0, 3, -1, 0, 0, 0, -1, 0)
    want, found = both(refs, qrys, scaled)
    assert found == want
    assert found[-1] == _lib.UNSUPPORTED_SYMBOLS
    capped = _lib.COptions(-(1 << 62), 1 << 62, 40, 1, 0, 0)
    want, found = both(refs, qrys, None, capped)
    assert found == want
    assert _lib.ABOVE_MAX in found


def test_static_distances_refusals():
    refs, qrys = synthetic.edit_short(64, seed=94)
    affine = _lib.CCosts(4, 6, 2, -1, 0, 0, 0, -1, 0)
    code, _ = call("dinara_gpu_distances_static", refs, qrys, affine)
    assert code == _lib.INVALID_MODE
    banded = _lib.COptions(-10, 10, -1, 1, 0, 0)
    code, _ = call("dinara_gpu_distances_static", refs, qrys, None, banded)
    assert code == _lib.INVALID_MODE
    # More letters than the scratch holds at 2 bits a letter: refused before any device work.
    big = [b"A" * (1 << 20)] * (SCRATCH_BYTES // (1 << 19) + 1)
    code, _ = call("dinara_gpu_distances_static", big, big)
    assert code == _lib.SCRATCH_FULL
    # The scratch still serves after a refusal.
    want, found = both(refs, qrys)
    assert found == want
