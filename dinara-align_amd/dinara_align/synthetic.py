"""Seeded synthetic pairs for the benchmark configurations, generated in bulk with numpy.

Each generator returns (references, queries) as lists of ASCII ``bytes``. A mutation walks the reference and,
at each letter with probability ``rate``, makes one event: a substitution (to a different letter), an
insertion before the letter, or a deletion starting at it; indel lengths are 1 or drawn geometric.
"""

from __future__ import annotations

from typing import List, Tuple

import numpy as np

DNA = np.frombuffer(b"ACGT", dtype=np.uint8)
PROTEIN = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)


def _split(codes: np.ndarray, owner: np.ndarray, pairs: int, letters: np.ndarray) -> List[bytes]:
    counts = np.bincount(owner, minlength=pairs)
    text = letters[codes].tobytes()
    out, at = [], 0
    for count in counts.tolist():
        out.append(text[at:at + count])
        at += count
    return out


def mutate(codes: np.ndarray, owner: np.ndarray, rate: np.ndarray, weights: Tuple[float, float, float],
           alphabet: int, rng: np.random.Generator, indel_p: float = 0.0) -> Tuple[np.ndarray, np.ndarray]:
    """Mutates the concatenated sequences `codes` (pair of each letter `owner`, rate per pair `rate`).
    `weights` are substitution, insertion, deletion; `indel_p` > 0 draws indel lengths Geometric(indel_p)."""
    total = len(codes)
    events = np.flatnonzero(rng.random(total) < rate[owner])
    kind = rng.choice(3, size=len(events), p=np.asarray(weights) / np.sum(weights))
    sub = np.zeros(total, dtype=bool)
    ins = np.zeros(total, dtype=bool)
    dele = np.zeros(total, dtype=bool)
    sub[events[kind == 0]] = True
    ins[events[kind == 1]] = True
    dele[events[kind == 2]] = True
    out = codes.copy()
    out[sub] = (out[sub] + rng.integers(1, alphabet, size=int(sub.sum()))) % alphabet
    keep = np.ones(total, dtype=bool)
    starts = np.flatnonzero(dele)
    if len(starts):
        lengths = rng.geometric(indel_p, size=len(starts)) if indel_p > 0 else np.ones(len(starts), dtype=np.int64)
        at = np.repeat(starts, lengths) + (np.arange(lengths.sum()) - np.repeat(np.cumsum(lengths) - lengths, lengths))
        inside = at < total
        at = at[inside]
        same = owner[at] == np.repeat(owner[starts], lengths)[inside]
        keep[at[same]] = False
    places = np.flatnonzero(ins)
    if len(places):
        lengths = rng.geometric(indel_p, size=len(places)) if indel_p > 0 else np.ones(len(places), dtype=np.int64)
        where = np.repeat(places, lengths)
        values = rng.integers(0, alphabet, size=len(where)).astype(out.dtype)
        out = np.insert(out, where, values)
        owner = np.insert(owner, where, owner[where])
        keep = np.insert(keep, where, True)
    return out[keep], owner[keep]


def _uniform(rng, lengths: np.ndarray, alphabet: int):
    owner = np.repeat(np.arange(len(lengths), dtype=np.int32), lengths)
    codes = rng.integers(0, alphabet, size=int(lengths.sum())).astype(np.uint8)
    return codes, owner


def edit_short(pairs: int = 10_000, seed: int = 1, length: int = 150, rate: float = 0.02):
    """Unit-cost distances: `length` bp uniform ACGT references, targets mutated at `rate` a base, substitutions,
    insertions and deletions alike, indels of one letter."""
    rng = np.random.Generator(np.random.PCG64(seed))
    codes, owner = _uniform(rng, np.full(pairs, length), 4)
    mutated, mowner = mutate(codes, owner, np.full(pairs, rate), (1, 1, 1), 4, rng)
    return _split(codes, owner, pairs, DNA), _split(mutated, mowner, pairs, DNA)


def affine_reads(pairs: int = 100_000, seed: int = 2, length: int = 250, rate: float = 0.05):
    """Gap-affine alignments: `length` bp uniform references, targets mutated at `rate` a base, 60% substitutions,
    20% insertions, 20% deletions, indel lengths Geometric(0.7)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    codes, owner = _uniform(rng, np.full(pairs, length), 4)
    mutated, mowner = mutate(codes, owner, np.full(pairs, rate), (0.6, 0.2, 0.2), 4, rng, indel_p=0.7)
    return _split(codes, owner, pairs, DNA), _split(mutated, mowner, pairs, DNA)


def uneven(pairs: int = 200_000, seed: int = 3, sigma: float = 1.0, clip=(100, 20_000)):
    """Uneven lengths: log-normal around 1.5 kbp (sigma `sigma`) clipped to `clip`, a rate per pair uniform in
    1%..15%, half substitutions, a quarter insertions and deletions, indels of one letter."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = np.clip(np.round(rng.lognormal(np.log(1500.0), sigma, size=pairs)), clip[0], clip[1]).astype(np.int64)
    rates = rng.uniform(0.01, 0.15, size=pairs)
    codes, owner = _uniform(rng, lengths, 4)
    mutated, mowner = mutate(codes, owner, rates, (0.5, 0.25, 0.25), 4, rng)
    return _split(codes, owner, pairs, DNA), _split(mutated, mowner, pairs, DNA)


def _uneven_chunk(job):
    pairs, seed, clip = job
    return uneven(pairs, seed=seed, clip=clip)


def uneven_parallel(pairs: int, seed: int = 3, clip=(100, 20_000), chunk: int = 50_000, workers: int = 16):
    """`uneven` in chunks of `chunk` pairs on `workers` processes, chunk i seeded `seed * 1000 + i` (the same pairs
    for the same arguments whatever the machine)."""
    jobs = [(min(chunk, pairs - at), seed * 1000 + i, clip) for i, at in enumerate(range(0, pairs, chunk))]
    if len(jobs) <= 1 or workers <= 1:
        parts = [_uneven_chunk(j) for j in jobs]
    else:
        import multiprocessing
        # Spawned, not forked: safe after the calling process has touched the GPU. The pool is closed and joined,
        # not terminated, so each worker exits on its own.
        pool = multiprocessing.get_context("spawn").Pool(min(workers, len(jobs)))
        try:
            parts = pool.map(_uneven_chunk, jobs)
        finally:
            pool.close()
            pool.join()
    refs, qrys = [], []
    for r, q in parts:
        refs.extend(r)
        qrys.extend(q)
    return refs, qrys


def protein_table(seed: int = 5) -> Tuple[str, tuple]:
    """A symmetric 20-letter table: its diagonal uniform in +4..+11, the rest uniform in -4..+3."""
    rng = np.random.Generator(np.random.PCG64(seed))
    size = len(PROTEIN)
    table = np.zeros((size, size), dtype=np.int64)
    for a in range(size):
        table[a, a] = rng.integers(4, 12)
        for b in range(a + 1, size):
            table[a, b] = table[b, a] = rng.integers(-4, 4)
    return PROTEIN.tobytes().decode("ascii"), tuple(int(v) for v in table.reshape(-1))


def local_protein(pairs: int = 50_000, seed: int = 5, query: int = 300, window: int = 3_000, rate: float = 0.30):
    """Local alignments: a `query`-residue uniform query against a `window`-residue uniform window holding a copy
    of the query mutated at `rate` (substitutions, insertions, deletions alike, length 1) at a random offset."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    size = len(PROTEIN)
    q_codes, q_owner = _uniform(rng, np.full(pairs, query), size)
    mutated, m_owner = mutate(q_codes, q_owner, np.full(pairs, rate), (1, 1, 1), size, rng)
    w_codes = rng.integers(0, size, size=pairs * window).astype(np.uint8).reshape(pairs, window)
    counts = np.bincount(m_owner, minlength=pairs)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    offsets = rng.integers(0, np.maximum(window - counts, 0) + 1)
    for p in range(pairs):
        piece = mutated[starts[p]:starts[p] + counts[p]][:window]
        w_codes[p, offsets[p]:offsets[p] + len(piece)] = piece
    windows = [PROTEIN[row].tobytes() for row in w_codes]
    return windows, _split(q_codes, q_owner, pairs, PROTEIN)
