"""The C ABI's threading contract: the library keeps no state between calls that one caller's batch could see from
another's, and may be called from many host threads at once (reference c/dinara.h:14-16; its pthreads batches,
c/dinara.mojo:822-990). Several threads call dinara_alignments / dinara_distances at the same time, on different
batches and costs, and each must get exactly what the same call made alone returns.
"""

import ctypes
import threading

import pytest

from dinara_align import _lib, synthetic

pytestmark = pytest.mark.gpu


def alignments(refs, qrys, costs, threads):
    h = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    found = (_lib.CAlignment * len(refs))()
    statuses = (ctypes.c_int64 * len(refs))()
    code = h.dinara_alignments(len(refs), *R.args(), *Q.args(), ctypes.byref(costs) if costs else None, None, None,
                               threads, found, statuses)
    out = []
    for i in range(len(refs)):
        if statuses[i] == 0:
            out.append((found[i].cost, ctypes.string_at(found[i].cigar, found[i].cigar_length)))
            h.dinara_free(found[i].cigar)
        else:
            out.append(statuses[i])
    return code, out


def distances(refs, qrys, costs, threads):
    h = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    out = (ctypes.c_int64 * len(refs))()
    code = h.dinara_distances(len(refs), *R.args(), *Q.args(), ctypes.byref(costs) if costs else None, None, None,
                              threads, out)
    return code, list(out)


def test_concurrent_calls_from_host_threads_match_calls_alone():
    affine = _lib.CCosts(4, 6, 2, -1, 0, 0, 0, -1, 0)
    jobs = []
    for t in range(6):
        if t % 3 == 0:
            refs, qrys = synthetic.affine_reads(1500, seed=300 + t)
            jobs.append((alignments, refs, qrys, affine))
        elif t % 3 == 1:
            refs, qrys = synthetic.uneven(800, seed=300 + t, clip=(20, 2500))
            jobs.append((alignments, refs, qrys, None))
        else:
            refs, qrys = synthetic.uneven(1200, seed=300 + t, clip=(20, 2500))
            jobs.append((distances, refs, qrys, affine))
    alone = [fn(refs, qrys, costs, 2) for fn, refs, qrys, costs in jobs]
    assert all(code == 0 for code, _ in alone)
    print(f"[threads] {len(jobs)} batches alone done", flush=True)
    together = [None] * len(jobs)
    start = threading.Barrier(len(jobs))

    def worker(i):
        fn, refs, qrys, costs = jobs[i]
        start.wait()
        together[i] = fn(refs, qrys, costs, 2)

    pool = [threading.Thread(target=worker, args=(i,)) for i in range(len(jobs))]
    for t in pool:
        t.start()
    for t in pool:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in pool)
    print(f"[threads] {len(jobs)} batches at once done", flush=True)
    for i, (want, got) in enumerate(zip(alone, together)):
        assert got == want, i
