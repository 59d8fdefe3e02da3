"""Synthetic search workloads (numpy, seeded).

Proteins: residues i.i.d. from the BLOSUM62 background frequencies
(A .074 R .052 N .045 D .054 C .025 Q .034 E .054 G .074 H .026 I .068
L .099 K .058 M .025 F .047 P .039 S .057 T .051 W .013 Y .032 V .073,
normalised), lengths log-normal (median 300 aa, sigma 0.44, so that one
million proteins hold about 330M aa; clipped to 50..2000) for database
entries.  Homologs of a query are planted into random database entries: the
query is copied with BLOSUM62-weighted substitutions at a target identity
(P(b | a) proportional to p_b * 2^(s(a, b) / 2), b != a: BLOSUM62 is in
half-bits) and indels at rate 0.02 per residue with geometric lengths of
mean 3, then written over a random window of the entry.

Translated reads: 150-nt reads, 70 % reverse-translated from random 50-aa
windows of database proteins (uniformly random synonymous codons, 1 % i.i.d.
nucleotide substitutions, either strand), 30 % uniform random ACGT.

Every generator returns (names, residues, offsets): names a list of str,
residues one ASCII bytes object of all sequences back to back and offsets a
uint64 array of n + 1 starts -- the form gz_queries_create and
gz_database_create take.
"""
import numpy as np

AMINO = b"ARNDCQEGHILKMFPSTWYV"
BACKGROUND = np.array([.074, .052, .045, .054, .025, .034, .054, .074, .026, .068,
                       .099, .058, .025, .047, .039, .057, .051, .013, .032, .073])
BACKGROUND = BACKGROUND / BACKGROUND.sum()
_BACKGROUND_CDF = np.cumsum(BACKGROUND)
_BACKGROUND_CDF[-1] = 1.0

_B62_ROWS = """
 4 -1 -2 -2  0 -1 -1  0 -2 -1 -1 -1 -1 -2 -1  1  0 -3 -2  0
-1  5  0 -2 -3  1  0 -2  0 -3 -2  2 -1 -3 -2 -1 -1 -3 -2 -3
-2  0  6  1 -3  0  0  0  1 -3 -3  0 -2 -3 -2  1  0 -4 -2 -3
-2 -2  1  6 -3  0  2 -1 -1 -3 -4 -1 -3 -3 -1  0 -1 -4 -3 -3
 0 -3 -3 -3  9 -3 -4 -3 -3 -1 -1 -3 -1 -2 -3 -1 -1 -2 -2 -1
-1  1  0  0 -3  5  2 -2  0 -3 -2  1  0 -3 -1  0 -1 -2 -1 -2
-1  0  0  2 -4  2  5 -2  0 -3 -3  1 -2 -3 -1  0 -1 -3 -2 -2
 0 -2  0 -1 -3 -2 -2  6 -2 -4 -4 -2 -3 -3 -2  0 -2 -2 -3 -3
-2  0  1 -1 -3  0  0 -2  8 -3 -3 -1 -2 -1 -2 -1 -2 -2  2 -3
-1 -3 -3 -3 -1 -3 -3 -4 -3  4  2 -3  1  0 -3 -2 -1 -3 -1  3
-1 -2 -3 -4 -1 -2 -3 -4 -3  2  4 -2  2  0 -3 -2 -1 -2 -1  1
-1  2  0 -1 -3  1  1 -2 -1 -3 -2  5 -1 -3 -1  0 -1 -3 -2 -2
-1 -1 -2 -3 -1  0 -2 -3 -2  1  2 -1  5  0 -2 -1 -1 -1 -1  1
-2 -3 -3 -3 -2 -3 -3 -3 -1  0  0 -3  0  6 -4 -2 -2  1  3 -1
-1 -2 -2 -1 -3 -1 -1 -2 -2 -3 -3 -1 -2 -4  7 -1 -1 -4 -3 -2
 1 -1  1  0 -1  0  0  0 -1 -2 -2  0 -1 -2 -1  4  1 -3 -2 -2
 0 -1  0 -1 -1 -1 -1 -2 -2 -1 -1 -1 -1 -2 -1  1  5 -2 -2  0
-3 -3 -4 -4 -2 -2 -3 -2 -2 -3 -2 -3 -1  1 -4 -3 -2 11  2 -3
-2 -2 -2 -3 -2 -1 -2 -3  2 -1 -1 -2 -1  3 -3 -2 -2  2  7 -1
 0 -3 -3 -3 -1 -2 -2 -3 -3  3  1 -2  1 -1 -2 -2  0 -3 -1  4
"""
BLOSUM62_20 = np.array(_B62_ROWS.split(), dtype=np.int64).reshape(20, 20)

# substitution kernel: P(b | a) ~ p_b 2^(s(a, b) / 2), b != a
_SUB = BACKGROUND[None, :] * np.exp2(BLOSUM62_20 / 2.0)
np.fill_diagonal(_SUB, 0.0)
_SUB /= _SUB.sum(axis=1, keepdims=True)
_SUB_CDF = np.cumsum(_SUB, axis=1)

_CODONS = {}
for _i, _c in enumerate("KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"):
    _n = "ACGT"
    codon = _n[_i // 16] + _n[(_i // 4) % 4] + _n[_i % 4]
    _CODONS.setdefault(_c, []).append(codon.encode())
# [20][6] codon table (padded) and counts, in AMINO order
_CODON_TABLE = np.zeros((20, 6, 3), np.uint8)
_CODON_COUNT = np.zeros(20, np.int64)
for _a, _ch in enumerate(AMINO.decode()):
    cods = _CODONS[_ch]
    _CODON_COUNT[_a] = len(cods)
    for _j, cod in enumerate(cods):
        _CODON_TABLE[_a, _j] = np.frombuffer(cod, np.uint8)
_COMPLEMENT = np.zeros(256, np.uint8)
for _x, _y in zip(b"ACGT", b"TGCA"):
    _COMPLEMENT[_x] = _y


# residue by a uniform 16-bit draw: each background probability is rounded to
# a multiple of 2^-16 (an error below 1.6e-5 per letter)
_CODE_TABLE = np.searchsorted(_BACKGROUND_CDF, (np.arange(1 << 16) + 0.5) / (1 << 16),
                              side="right").astype(np.uint8)


def random_codes(rng, n):
    """n residue indexes in AMINO order from the background frequencies."""
    return _CODE_TABLE[rng.integers(0, 1 << 16, n, dtype=np.uint16)]


def lognormal_lengths(rng, n, median=300, sigma=0.44, lo=50, hi=2000):
    x = rng.lognormal(np.log(median), sigma, n)
    return np.clip(np.rint(x), lo, hi).astype(np.int64)


def substitute(rng, codes, identity):
    """Replace a (1 - identity) fraction of positions by BLOSUM62-weighted
    different residues."""
    out = codes.copy()
    m = rng.random(len(out)) >= identity
    idx = np.nonzero(m)[0]
    if len(idx):
        u = rng.random(len(idx))
        rows = _SUB_CDF[out[idx]]
        out[idx] = np.minimum((rows < u[:, None]).sum(axis=1), 19)
    return out


def homolog(rng, codes, identity, indel_rate=0.02, indel_mean=3.0):
    s = substitute(rng, codes, identity)
    n = len(s)
    events = np.nonzero(rng.random(n) < indel_rate)[0]
    if len(events) == 0:
        return s
    parts, last = [], 0
    for e in events:
        if e < last:
            continue
        parts.append(s[last:e])
        k = int(rng.geometric(1.0 / indel_mean))
        if rng.random() < 0.5:
            parts.append(random_codes(rng, k))  # insertion
            last = e
        else:
            last = min(n, e + k)  # deletion
    parts.append(s[last:])
    out = np.concatenate(parts)
    return out if len(out) else s[:1]


def _pack(seqs_codes, prefix):
    names = [f"{prefix}{i}" for i in range(len(seqs_codes))]
    lens = np.array([len(s) for s in seqs_codes], np.uint64)
    offsets = np.zeros(len(seqs_codes) + 1, np.uint64)
    np.cumsum(lens, out=offsets[1:])
    alpha = np.frombuffer(AMINO, np.uint8)
    residues = alpha[np.concatenate(seqs_codes)].tobytes() if seqs_codes else b""
    return names, residues, offsets


def protein_database(rng, n, median=300, lo=50, hi=2000):
    """Database codes as one array + offsets (mutable for planting)."""
    lens = lognormal_lengths(rng, n, median, lo=lo, hi=hi)
    codes = random_codes(rng, int(lens.sum()))
    offs = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=offs[1:])
    return [codes[offs[i]:offs[i + 1]] for i in range(n)]


def plant(rng, db, query_codes, copies, identity=(0.3, 0.9)):
    """Write `copies` homologs of the query over random windows of random
    database entries (entries grow when the homolog is longer)."""
    for _ in range(copies):
        h = homolog(rng, query_codes, rng.uniform(*identity))
        e = int(rng.integers(len(db)))
        s = db[e]
        if len(h) >= len(s):
            db[e] = h.copy()
        else:
            at = int(rng.integers(0, len(s) - len(h) + 1))
            s = s.copy()
            s[at:at + len(h)] = h
            db[e] = s


def protein_workload(seed, n_db, n_queries, qlen=(100, 500), homologs=3, db_median=300):
    """Protein-vs-protein search: uniform query lengths, each query planted
    `homologs` times.  Returns ((qnames, qres, qoffs), (dnames, dres, doffs))."""
    rng = np.random.default_rng(seed)
    db = protein_database(rng, n_db, db_median)
    queries = [random_codes(rng, int(rng.integers(qlen[0], qlen[1] + 1)))
               for _ in range(n_queries)]
    for q in queries:
        plant(rng, db, q, homologs)
    return _pack(queries, "query_"), _pack(db, "subject_")


def pareto_lengths(rng, n, lo=30, median=250, hi=35000):
    # Pareto with x_m = lo and alpha chosen so the median is `median`
    alpha = np.log(2.0) / np.log(median / lo)
    x = lo * (1.0 - rng.random(n)) ** (-1.0 / alpha)
    return np.clip(np.rint(x), lo, hi).astype(np.int64)


def uneven_protein_workload(seed, n_db, n_queries, planted_fraction=0.2, homologs=3):
    """Pareto-tailed query lengths (min 30, median 250, max 35,000)."""
    rng = np.random.default_rng(seed)
    db = protein_database(rng, n_db)
    lens = pareto_lengths(rng, n_queries)
    queries = [random_codes(rng, int(L)) for L in lens]
    for q in queries:
        if rng.random() < planted_fraction:
            plant(rng, db, q, homologs)
    return _pack(queries, "query_"), _pack(db, "subject_")


def translated_reads(rng, db, n_reads, read_nt=150, from_db=0.7, substitution=0.01):
    """Reads as ASCII nucleotides (n_reads x read_nt)."""
    aa = read_nt // 3
    reads = np.empty((n_reads, read_nt), np.uint8)
    is_db = rng.random(n_reads) < from_db
    k = int(is_db.sum())
    if k:
        lens = np.array([len(s) for s in db])
        ok = np.nonzero(lens >= aa)[0]
        pick = ok[rng.integers(0, len(ok), k)]
        windows = np.empty((k, aa), np.uint8)
        for i, e in enumerate(pick):
            st = int(rng.integers(0, lens[e] - aa + 1))
            windows[i] = db[e][st:st + aa]
        choice = np.floor(rng.random((k, aa)) * _CODON_COUNT[windows]).astype(np.int64)
        nt = _CODON_TABLE[windows, choice].reshape(k, read_nt)
        sub = rng.random(nt.shape) < substitution
        if sub.any():
            alpha = np.frombuffer(b"ACGT", np.uint8)
            cur = nt[sub]
            idx = np.searchsorted(alpha, cur)
            nt[sub] = alpha[(idx + rng.integers(1, 4, len(cur))) % 4]
        rev = rng.random(k) < 0.5
        nt[rev] = _COMPLEMENT[nt[rev][:, ::-1]]
        reads[is_db] = nt
    rnd = ~is_db
    if rnd.any():
        reads[rnd] = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, (int(rnd.sum()), read_nt))]
    names = [f"read_{i}" for i in range(n_reads)]
    offsets = np.arange(n_reads + 1, dtype=np.uint64) * read_nt
    return names, reads.tobytes(), offsets


def translated_workload(seed, n_db, n_reads, db_median=300):
    """Translated DNA reads against a protein database."""
    rng = np.random.default_rng(seed)
    db = protein_database(rng, n_db, db_median)
    q = translated_reads(rng, db, n_reads)
    return q, _pack(db, "subject_")


def write_fasta(path, names, residues, offsets, width=60):
    with open(path, "w") as f:
        for i, nm in enumerate(names):
            s = residues[int(offsets[i]):int(offsets[i + 1])].decode()
            f.write(f">{nm}\n")
            for j in range(0, len(s), width):
                f.write(s[j:j + width] + "\n")
