"""Shared test inputs: delimiter-framed random sequences and seeds."""
import numpy as np

DELIM = 25
# BLOSUM62 background frequencies (A R N D C Q E G H I L K M F P S T W Y V)
BACKGROUND = np.array([.074, .052, .045, .054, .025, .034, .054, .074, .026, .068,
                       .099, .058, .025, .047, .039, .057, .051, .013, .032, .073])
BACKGROUND = BACKGROUND / BACKGROUND.sum()


def random_protein(rng, n, ambiguous=0.0):
    s = rng.choice(20, size=n, p=BACKGROUND).astype(np.uint8)
    if ambiguous > 0:
        m = rng.random(n) < ambiguous
        s[m] = rng.integers(20, 25, m.sum())
    return s


def mutate(rng, s, identity, indel=0.0):
    out = []
    i = 0
    while i < len(s):
        if indel and rng.random() < indel:
            if rng.random() < 0.5:
                i += int(rng.geometric(1 / 3))
                continue
            out.extend(random_protein(rng, int(rng.geometric(1 / 3))).tolist())
        c = s[i]
        if rng.random() > identity:
            c = rng.choice(20, p=BACKGROUND)
        out.append(int(c))
        i += 1
    return np.array(out, np.uint8)


def concat(seqs):
    """Leading delimiter, each sequence followed by a delimiter."""
    parts = [np.array([DELIM], np.uint8)]
    offsets = []
    pos = 1
    for s in seqs:
        offsets.append(pos)
        parts.append(np.asarray(s, np.uint8))
        parts.append(np.array([DELIM], np.uint8))
        pos += len(s) + 1
    offsets.append(pos)
    return np.concatenate(parts), np.array(offsets, np.uint32)


def related_pair_set(rng, n, min_len=20, max_len=300, identity=(0.3, 0.95), indel=0.02):
    """Query/database concatenations whose i-th sequences are homologous, plus
    one seed per pair on a shared diagonal (query pos, db pos)."""
    qs, ds = [], []
    for _ in range(n):
        L = int(rng.integers(min_len, max_len))
        a = random_protein(rng, L, ambiguous=0.01)
        b = mutate(rng, a, rng.uniform(*identity), indel)
        if len(b) < 2:
            b = random_protein(rng, 5)
        qs.append(a)
        ds.append(b)
    q, qo = concat(qs)
    d, do = concat(ds)
    qpos, dpos = [], []
    for i in range(n):
        la, lb = qo[i + 1] - qo[i] - 1, do[i + 1] - do[i] - 1
        k = int(rng.integers(0, min(la, lb)))
        qpos.append(qo[i] + k)
        dpos.append(do[i] + k)
    return q, qo, d, do, np.array(qpos, np.uint32), np.array(dpos, np.uint32)
