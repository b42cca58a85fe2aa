"""Independent pure-Python restatement of the path for small inputs.

Written from the reference's definitions with Python big integers and
strings (no limb shifting, no rolling updates), so that a mistake in the C
oracle's word-level code shows up as a disagreement.  Test code only.
"""
from __future__ import annotations

import math

M64 = (1 << 64) - 1
CMS_SEEDS = [0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0, 0x082EFA98EC4E6C89]
NT = {"A": 0x3C8BFBB395C60474, "C": 0x3193C18562A02B4C, "G": 0x20323EE046061444, "T": 0x295C504FAC61002E}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def splitmix64(z: int) -> int:
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def norm(seq: bytes) -> str:
    """ASCII -> ACGT with every non-ACGT byte read as A (bitseq.rs:67-81)."""
    out = []
    for c in seq:
        ch = chr(c).upper()
        out.append(ch if ch in "ACGT" else "A")
    return "".join(out)


def revcomp(s: str) -> str:
    return "".join(COMP[c] for c in reversed(s))


def kmer_int(s: str) -> int:
    v = 0
    for c in s:
        v = v * 4 + "ACGT".index(c)
    return v


def canonical(s: str) -> tuple[int, bool]:
    f, r = kmer_int(s), kmer_int(revcomp(s))
    return (r, True) if r < f else (f, False)


def fold(x: int) -> int:
    """to_u64: XOR of the 64-bit limbs of the k-mer integer."""
    f = 0
    while x:
        f ^= x & M64
        x >>= 64
    return f


def ref_words(x: int, k: int) -> list:
    nw = 1 if k <= 31 else (2 if k <= 63 else 4)
    return [(x >> (64 * (nw - 1 - i))) & M64 for i in range(nw)]


def rol(x: int, r: int) -> int:
    r %= 64
    return ((x << r) | (x >> (64 - r))) & M64 if r else x


def nthash_fwd(s: str) -> int:
    k, h = len(s), 0
    for i, c in enumerate(s):
        h ^= rol(NT[c], k - 1 - i)
    return h


def nthash_rev(s: str) -> int:
    h = 0
    for i, c in enumerate(s):
        h ^= rol(NT[COMP[c]], i)
    return h


def nthash_canonical(s: str) -> int:
    return min(nthash_fwd(s), nthash_rev(s))


def cms_cells(key: int, width: int) -> list:
    return [r * width + (splitmix64(key ^ CMS_SEEDS[r]) & (width - 1)) for r in range(4)]


def hll_index_rho(h: int, p: int) -> tuple[int, int]:
    idx = h >> (64 - p)
    w = h & ((1 << (64 - p)) - 1)
    rho = (64 - p + 1) if w == 0 else ((64 - p) - w.bit_length()) + 1
    return idx, rho


def hll_estimate(regs, p: int) -> float:
    m = float(1 << p)
    alpha = 0.7213 / (1.0 + 1.079 / m)
    s = sum(2.0 ** -int(v) for v in regs)
    zeros = sum(1 for v in regs if v == 0)
    raw = alpha * m * m / s
    if raw <= 2.5 * m:
        return m * math.log(m / zeros) if zeros else raw
    return raw


# ---- CMS table from exact per-cell counts (order-free) ---------------------

def cms_table_from_counts(cells: dict, width: int, bits: int) -> bytes | list:
    cap = 15 if bits == 4 else 65535
    if bits == 16:
        t = [0] * (4 * width)
        for c, n in cells.items():
            t[c] = min(n, cap)
        return t
    t = bytearray(2 * width)
    for c, n in cells.items():
        v = min(n, cap)
        if c & 1:
            t[c >> 1] |= v << 4
        else:
            t[c >> 1] |= v
    return bytes(t)


def cms_estimate(cells: dict, key: int, width: int, bits: int) -> int:
    cap = 15 if bits == 4 else 65535
    return min(min(cells.get(c, 0), cap) for c in cms_cells(key, width))


# ---- passes 1-3 (build.rs:93-328) by brute force ---------------------------

def build_solid(reads: list, k: int, width: int, target=None, bits: int = 4, contigs=()) -> tuple:
    """Returns (sorted [(kmer_int, cov, mask)], cell counts)."""
    t_solid = 2 if target is None else target
    gate = 3 if target is None else target
    cells: dict = {}
    seqs = [norm(r) for r in reads]
    cseqs = [norm(c) for c in contigs]
    for s in seqs:
        for i in range(len(s) - k + 1):
            for c in cms_cells(fold(canonical(s[i:i + k])[0]), width):
                cells[c] = cells.get(c, 0) + 1
    for s in cseqs:
        for i in range(len(s) - k + 1):
            for c in cms_cells(fold(canonical(s[i:i + k])[0]), width):
                cells[c] = cells.get(c, 0) + 10
    acc: dict = {}

    def push(key, edge, cw, ew):
        a = acc.setdefault(key, [0, [0] * 8])
        a[0] += cw
        a[1][edge] += ew

    def pairs(s, gated):
        for i in range(len(s) - k):
            u, v = s[i:i + k], s[i + 1:i + k + 1]
            (uc, urc), (vc, vrc) = canonical(u), canonical(v)
            if gated and not (cms_estimate(cells, fold(uc), width, bits) >= t_solid
                              and cms_estimate(cells, fold(vc), width, bits) >= t_solid):
                continue
            b = "ACGT".index(s[i + k])
            p = "ACGT".index(s[i])
            yield uc, (b + 4 if urc else b), vc, ((3 - p) if vrc else (3 - p) + 4)

    for s in cseqs:
        for uc, eu, vc, ev in pairs(s, False):
            push(uc, eu, 100, 1000)
            push(vc, ev, 100, 1000)
    for s in seqs:
        for uc, eu, vc, ev in pairs(s, True):
            push(uc, eu, 1, 1)
            push(vc, ev, 1, 1)
    out = []
    for key, (cov, edges) in acc.items():
        mask = 0
        for e in range(8):
            if min(edges[e], 65535) >= gate:
                mask |= 1 << e
        if mask:
            out.append((key, max(min(cov, 65535) // 2, 1), mask))
    out.sort()
    return out, cells


# ---- estimator sketch loop (estimator.rs:309-379), sequential -------------

def estimator_registers(pairs: list, n_pairs: int, width: int = 1 << 23):
    """Final hll_total / hll_solid registers (p=14) after n_pairs pairs, with
    the order-dependent `c >= 2` test evaluated at insertion time."""
    k = 21
    cells: dict = {}
    ht, hs = [0] * (1 << 14), [0] * (1 << 14)
    m = 0
    for r1, r2 in pairs[:n_pairs]:
        for s in (norm(r1), norm(r2)):
            if len(s) < k:
                continue
            m += len(s) - k + 1
            for i in range(len(s) - k + 1):
                raw = nthash_canonical(s[i:i + k])
                cs = cms_cells(raw, width)
                for c in cs:
                    cells[c] = min(cells.get(c, 0) + 1, 15)
                cnt = min(cells[c] for c in cs)
                mixed = splitmix64(raw)
                idx, rho = hll_index_rho(mixed, 14)
                ht[idx] = max(ht[idx], rho)
                if cnt >= 2:
                    hs[idx] = max(hs[idx], rho)
    return ht, hs, m


# ---- unitig count of a solid set (walk of build.rs:332-549) ----------------

def count_unitigs(solid: list, k: int) -> int:
    """Number of unitigs the reference's compaction walk emits for a solid
    set [(kmer_int, cov, mask)]: walks from boundary directed k-mers plus
    isolated cycles (build.rs:332-549)."""
    nodes = {key: [mask, 0] for key, _, mask in solid}  # mask, visited
    full = (1 << (2 * k)) - 1

    def rc_int(x):
        return kmer_int(revcomp(to_str(x)))

    def to_str(x):
        return "".join("ACGT"[(x >> (2 * (k - 1 - i))) & 3] for i in range(k))

    def step(x_fwd, b):
        nf = ((x_fwd << 2) | b) & full
        nr = rc_int(nf)
        return (nr, True) if nr < nf else (nf, False)

    def first_base(x_fwd):
        return (x_fwd >> (2 * (k - 1))) & 3

    count = 0
    boundary = []
    for key, (mask, _) in nodes.items():
        fo, fi = bin(mask & 15).count("1"), bin(mask >> 4).count("1")
        if fi != 1 or fo != 1:
            boundary.append((key, False))
        if fo != 1 or fi != 1:
            boundary.append((key, True))

    def walk_from(ukey, urc, b):
        fwd = rc_int(ukey) if urc else ukey
        rc_first = 3 - first_base(fwd)
        ckey, crc = step(fwd, b)
        if ckey in nodes:
            nodes[ckey][1] |= 1 << (rc_first + 4 if not crc else rc_first)
        steps = 0
        while True:
            if ckey not in nodes:
                break
            m = nodes[ckey][0]
            out_m = (m & 15) if not crc else (m >> 4)
            in_m = (m >> 4) if not crc else (m & 15)
            if bin(in_m).count("1") == 1 and bin(out_m).count("1") == 1:
                nb = (out_m & -out_m).bit_length() - 1
                eb = nb if not crc else nb + 4
                if nodes[ckey][1] & (1 << eb):
                    break
                nodes[ckey][1] |= 1 << eb
                cf = rc_int(ckey) if crc else ckey
                rp = 3 - first_base(cf)
                nkey, nrc = step(cf, nb)
                if nkey in nodes:
                    nodes[nkey][1] |= 1 << (rp + 4 if not nrc else rp)
                ckey, crc = nkey, nrc
                steps += 1
            else:
                break
        return ckey, crc

    for ukey, urc in boundary:
        for b in range(4):
            if ukey not in nodes:
                break
            m, vis = nodes[ukey]
            out_m = (m & 15) if not urc else (m >> 4)
            v = (vis & 15) if not urc else (vis >> 4)
            if out_m & (1 << b) and not v & (1 << b):
                nodes[ukey][1] |= 1 << (b if not urc else b + 4)
                walk_from(ukey, urc, b)
                count += 1
    for key in list(nodes):
        m, vis = nodes[key]
        if bin(m & 15).count("1") == 1 and bin(m >> 4).count("1") == 1:
            b = (m & 15).bit_length() - 1
            if not vis & (1 << b):
                nodes[key][1] |= 1 << b
                end, _ = walk_from(key, False, b)
                if end == key:
                    count += 1
    return count


def compact(solid: list, k: int) -> dict:
    """The reference's unitig compaction (build.rs:330-597) on a solid set
    [(kmer_int, cov, mask)] taken in list order (the boundary walks, then the
    cycles), written with strings: every k-mer is a string, a step appends a
    letter and drops the first, orientation is by comparing the integer of the
    string with that of its reverse complement.  Returns seqs, cov (as float32
    values), circular flags and edges (src, dst, ov, support) per node in
    insertion order, forward side then reverse side."""
    import numpy as np

    def to_str(x):
        return "".join("ACGT"[(x >> (2 * (k - 1 - i))) & 3] for i in range(k))

    nodes = {key: [mask, 0, c] for key, c, mask in solid}  # mask, visited, cov
    order = [key for key, _, _ in solid]

    def oriented(key, rc):
        s = to_str(key)
        return revcomp(s) if rc else s

    def step(s, b):
        """next directed k-mer after string s by base b, marking its twin edge"""
        t = s[1:] + "ACGT"[b]
        key, rc = canonical(t)
        rc_first = 3 - "ACGT".index(s[0])
        if key in nodes:
            nodes[key][1] |= 1 << (rc_first + 4 if not rc else rc_first)
        return key, rc

    def walk(key, rc, stop):
        path, total, points = [], 0, 0
        while key in nodes and key != stop:
            m, vis, c = nodes[key]
            out_m = (m & 15) if not rc else (m >> 4)
            in_m = (m >> 4) if not rc else (m & 15)
            if bin(in_m).count("1") != 1 or bin(out_m).count("1") != 1:
                break
            nb = (out_m & -out_m).bit_length() - 1
            eb = nb if not rc else nb + 4
            if vis & (1 << eb):
                break
            nodes[key][1] |= 1 << eb
            key, rc = step(oriented(key, rc), nb)
            path.append("ACGT"[nb])
            total += c
            points += 1
        return key, path, total, points

    unitigs = []  # [seq, cov, circular, edges_fwd, edges_rev]

    def f32_mean(total, points):
        v = np.float32(np.float32(total) / np.float32(points))
        return max(v, np.float32(1.0))

    for key in order:
        for urc in (False, True):
            m = nodes[key][0]
            fo, fi = bin(m & 15).count("1"), bin(m >> 4).count("1")
            if not (fi != 1 or fo != 1):
                continue
            for b in range(4):
                m, vis, c = nodes[key]
                out_m = (m & 15) if not urc else (m >> 4)
                v = (vis & 15) if not urc else (vis >> 4)
                if not (out_m & (1 << b)) or v & (1 << b):
                    continue
                nodes[key][1] |= 1 << (b if not urc else b + 4)
                s = oriented(key, urc)
                nkey, nrc = step(s, b)
                _, tail, total, points = walk(nkey, nrc, None)
                seq = s + "ACGT"[b] + "".join(tail)
                seq = min(seq, revcomp(seq))
                unitigs.append([seq, f32_mean(c + total, 1 + points), False, [], []])
    for key in order:
        m, vis, c = nodes[key]
        if bin(m & 15).count("1") != 1 or bin(m >> 4).count("1") != 1:
            continue
        b = (m & 15).bit_length() - 1
        if vis & (1 << b):
            continue
        nodes[key][1] |= 1 << b
        s = oriented(key, False)
        nkey, nrc = step(s, b)
        end, tail, total, points = walk(nkey, nrc, key)
        seq = s + "ACGT"[b] + "".join(tail)
        if end == key and len(seq) > k:
            if seq[-k:] == seq[:k]:
                seq = seq[:-k]
            elif seq[-(k - 1):] == seq[:k - 1]:
                seq = seq[:-(k - 1)]
            u = [seq, f32_mean(c + total, 1 + points), True, [], []]
            u[3].append([len(unitigs) << 1, k - 1, 10])
            unitigs.append(u)

    def add(lst, target, ov):
        for e in lst:
            if e[0] == target and e[1] == ov:
                e[2] = min(e[2] + 10, 65535)
                return
        lst.append([target, ov, 10])

    for ov in (k, k - 1):
        if ov == 0:
            continue
        pm = {}
        for i, u in enumerate(unitigs):
            if len(u[0]) >= ov:
                pm.setdefault(u[0][:ov], []).append(i << 1)
                pm.setdefault(revcomp(u[0])[:ov], []).append(i << 1 | 1)
        for i, u in enumerate(unitigs):
            if len(u[0]) < ov:
                continue
            for side, strand in ((0, u[0]), (1, revcomp(u[0]))):
                for t in pm.get(strand[-ov:], []):
                    if (t >> 1) != i or u[2]:
                        add(u[3 + side], t, ov)
    src, dst, ovs, sup = [], [], [], []
    for i, u in enumerate(unitigs):
        for side in (0, 1):
            for t, ov, s in u[3 + side]:
                src.append(i << 1 | side)
                dst.append(t)
                ovs.append(ov)
                sup.append(s)
    return {"seqs": [u[0].encode() for u in unitigs], "cov": np.array([u[1] for u in unitigs], np.float32),
            "circular": np.array([1 if u[2] else 0 for u in unitigs], np.uint8), "src": np.array(src, np.uint32),
            "dst": np.array(dst, np.uint32), "ov": np.array(ovs, np.uint8), "support": np.array(sup, np.uint16)}


def solid_from_sequences(seqs: list, k: int, cov_of=None, circular: bool = False) -> list:
    """An exact solid set of the k-mers of error-free sequences: every
    adjacent pair of k-mers sets the out-edge of the first and the in-edge of
    the second (build.rs:268-275 edge encoding); cov from cov_of(key) or 10."""
    acc = {}
    for s in seqs:
        t = s + s[:k] if circular else s
        kms = [t[i:i + k] for i in range(len(t) - k + 1)]
        for i in range(len(kms) - 1):
            u, v = kms[i], kms[i + 1]
            b, p = "ACGT".index(v[-1]), "ACGT".index(u[0])
            uk, urc = canonical(u)
            vk, vrc = canonical(v)
            acc.setdefault(uk, 0)
            acc.setdefault(vk, 0)
            acc[uk] |= 1 << (b + 4 if urc else b)
            acc[vk] |= 1 << ((3 - p) if vrc else (3 - p) + 4)
    return sorted((key, (cov_of(key) if cov_of else 10), m) for key, m in acc.items())


def filter_active_reads(nodes: list, k: int, merged: list, pairs: list) -> tuple:
    """local_multik.rs:259-333 on strings.  nodes: (seq, circular, deleted,
    n_edges_fwd, n_edges_rev); merged: read strings; pairs: (r1, r2).
    Returns (keep_merged, keep_pairs) as lists of bools."""
    kk = min(k, 31)
    boundary = set()
    for seq, circ, deleted, nf, nr in nodes:
        if deleted or len(seq) < kk:
            continue
        starts = range(len(seq) - kk + 1)
        if not (len(seq) < 2000 or circ or nf > 1 or nr > 1):
            # the 350 windows at each end (both ranges clipped to the unitig)
            left = set(range(min(350, len(seq) - kk + 1)))
            right = set(range(min(max(len(seq) - 350, 0), len(seq) - kk), len(seq) - kk + 1))
            starts = sorted(left | right)
        for p in starts:
            boundary.add(nthash_canonical(seq[p:p + kk]))

    def active(r: str) -> bool:
        if len(r) < kk:
            return False
        return any(nthash_canonical(r[o:o + kk]) in boundary for o in (0, (len(r) - kk) // 2, len(r) - kk))

    return [active(r) for r in merged], [active(a) or active(b) for a, b in pairs]


# ---- ingest (trim_merge.rs:17-417), written on Python strings ---------------

ADAPTERS_AGAT = ("AGATCGGAAGAGC", "AGATGTGTATAAGAGACAG")
ADAPTER_CTGT = "CTGTCTCTTATACACATCT"


def _close(window: str, adapter: str) -> bool:
    return sum(1 for x, y in zip(window.upper(), adapter[:10]) if x != y) <= 1


def trim_adapter(seq: str) -> int:
    """Kept length: the first position (within the last 100, at least 10
    bases from the end) where an adapter's first 10 bases match with at most
    one mismatch, behind the AG / CT prefix test."""
    n = len(seq)
    if n < 10:
        return n
    for i in range(max(n - 100, 0), n - 9):
        b0, b1 = seq[i].upper(), seq[i + 1].upper()
        agat = b0 == "A" or b1 == "G"
        ctgt = b0 == "C" or b1 == "T"
        w = seq[i:i + 10]
        if agat and any(_close(w, ad) for ad in ADAPTERS_AGAT):
            return i
        if ctgt and _close(w, ADAPTER_CTGT):
            return i
    return n


def trim_quality(qual: str, min_qual: int, min_len: int):
    """Kept length or None: drop the tail until a 4-base window averages at
    least min_qual (integer mean of Phred = ord - 33, floored at 0)."""
    q = [max(ord(c) - 33, 0) for c in qual]
    if len(q) < min_len:
        return None
    cut = len(q)
    if cut >= 4 and sum(q[-4:]) < 4 * min_qual:
        while cut >= 4 and sum(q[cut - 4:cut]) // 4 < min_qual:
            cut -= 1
    return None if cut < min_len else cut


def merge_pair(r1: str, q1: str, r2: str, q2: str, min_overlap: int = 15, ratio: float = 0.10):
    """-> (merged, overlap, mismatches) or None.  Overlaps are tried from the
    longest down; score = d - 2.5 m over those with m <= floor(d * ratio);
    the search stops once d cannot beat the best score."""
    if len(r1) < min_overlap or len(r2) < min_overlap:
        return None
    rc2 = "".join({"A": "T", "C": "G", "G": "C", "T": "A"}.get(c.upper(), "A") for c in reversed(r2))
    best = None
    for d in range(min(len(r1), len(r2)), min_overlap - 1, -1):
        m = sum(1 for x, y in zip(r1[len(r1) - d:].upper(), rc2[:d]) if x != y)
        if m <= int(d * ratio):
            score = d - 2.5 * m
            if best is None or score > best[0]:
                best = (score, d, m)
        if best is not None and d <= best[0]:
            break
    if best is None:
        return None
    _, d, m = best
    out = [r1[:len(r1) - d]]
    for i in range(d):
        j = len(r1) - d + i
        b1, b2 = r1[j], rc2[i]
        qa = q1[j] if j < len(q1) else "I"
        k2 = len(r2) - 1 - i
        qb = q2[k2] if k2 < len(q2) else "I"
        out.append(b1.upper() if (b1.upper() == b2 or qa >= qb) else b2)
    out.append(rc2[d:])
    return "".join(out), d, m


# ---- PE mapper (mapper.rs:89-279), with dictionaries of Python ints ---------

def _fold_canonical(s: str, k: int) -> tuple:
    """(to_u64 of the canonical k-mer, is_rc): the fold XORs the 64-bit words
    of the reference's integer type for k."""
    x, is_rc = canonical(s)
    nw = 1 if k <= 31 else (2 if k <= 63 else 4)
    f = 0
    for w in ref_words(x, k)[:nw]:
        f ^= w
    return f, is_rc


def map_paired_reads(unitigs: list, pairs: list, k: int):
    """unitigs: sequences (active, linear); pairs: (r1, r2) strings.
    -> (links as tuples in pair order, insert sizes in pair order)."""
    counts, where = {}, {}
    for uid, seq in enumerate(unitigs):
        t = len(seq) - k + 1
        if t <= 0:
            continue
        starts = list(range(t)) if t <= 2000 else list(range(1000)) + list(range(t - 1000, t))
        for p in starts:
            h, rc = _fold_canonical(seq[p:p + k], k)
            counts[h] = counts.get(h, 0) + 1
            where[h] = (uid * 2 + int(rc), p)
    index = {h: v for h, v in where.items() if counts[h] == 1}

    def hit(r):
        for st in (0, 20, 40, 60, 80, 100):
            if st + k <= len(r):
                h, rc = _fold_canonical(r[st:st + k], k)
                if h in index:
                    d, p = index[h]
                    return (d ^ 1 if rc else d), p, st
        return None

    links, sizes = [], []
    for i, (r1, r2) in enumerate(pairs):
        if len(r1) < k or len(r2) < k:
            continue
        a, b = hit(r1), hit(r2)
        if a is None or b is None:
            continue
        (u1, p1, s1), (u2, p2, s2) = a, b
        if u1 // 2 == u2 // 2:
            if len(unitigs[u1 // 2]) < 1500:
                continue
            if u1 % 2 == 0 and u2 % 2 == 1:
                span = (p2 + k + s2) - (p1 - s1)
            elif u1 % 2 == 1 and u2 % 2 == 0:
                span = (p1 + k + s1) - (p2 - s2)
            else:
                span = -1
            if max(len(r1), len(r2), 100) <= span <= 1500:
                sizes.append(span)
        else:
            u2f = u2 ^ 1
            d1 = p1 + k + s1 if u1 % 2 else len(unitigs[u1 // 2]) - (p1 - s1)
            d2 = len(unitigs[u2 // 2]) - (p2 - s2) if u2f % 2 else p2 + k + s2
            d1, d2 = max(d1, 0), max(d2, 0)
            links.append((u1, u2f, i, d1, d2, len(r1), len(r2)))
            links.append((u2f ^ 1, u1 ^ 1, i, d2, d1, len(r2), len(r1)))
    return links, sizes


# ---- consensus polishing (consensus.rs:64-271), on strings -----------------

def polish(seqs: list, eligible: list, reads: list):
    """seqs: node sequences; eligible: bools; reads: strings in order.
    -> (polished sequences, stats dict)."""
    k = 31
    sampled, last = set(), {}
    for i, s in enumerate(seqs):
        if eligible[i]:
            for p in range(0, len(s) - k + 1, 25):
                h = nthash_canonical(s[p:p + k])
                sampled.add(h)
                last[h] = (i, p)
    occ = {h: 0 for h in sampled}
    for i, s in enumerate(seqs):
        if eligible[i]:
            for p in range(len(s) - k + 1):
                h = nthash_canonical(s[p:p + k])
                if h in occ:
                    occ[h] += 1
    seeds = {h: v for h, v in last.items() if occ[h] == 1}
    piles = {i: [[0, 0, 0, 0] for _ in s] for i, s in enumerate(seqs) if eligible[i]}
    st = dict(bases_modified=0, fwd_mapped_reads=0, rev_mapped_reads=0, fwd_pileup_bases=0, rev_pileup_bases=0,
              rejected_anchors=0)
    for r in reads:
        if len(r) < k:
            continue
        hit, anchor = False, None
        for p in range(0, len(r) - k + 1, 30):
            h = nthash_canonical(r[p:p + k])
            if h in seeds:
                hit = True
                cid, cp = seeds[h]
                a = seqs[cid][cp:cp + k]
                if a.upper() == r[p:p + k].upper():
                    anchor = (cid, cp, p, "f")
                    break
                if a.upper() == revcomp(r[p:p + k].upper()):
                    anchor = (cid, cp, p, "r")
                    break
        if anchor is None:
            st["rejected_anchors"] += int(hit)
            continue
        cid, cp, rp, o = anchor
        pile = piles[cid]
        voted = False
        for idx, b in enumerate(r):
            t = cp - rp + idx if o == "f" else cp + k - 1 + rp - idx
            if 0 <= t < len(pile):
                code = "ACGT".index(b) if o == "f" else 3 - "ACGT".index(b)
                pile[t][code] = min(pile[t][code] + 1, 65535)
                st["fwd_pileup_bases" if o == "f" else "rev_pileup_bases"] += 1
                voted = True
        if voted:
            st["fwd_mapped_reads" if o == "f" else "rev_mapped_reads"] += 1
    out = []
    for i, s in enumerate(seqs):
        s = list(s)
        if eligible[i]:
            for p, c in enumerate(piles[i]):
                tot = sum(c) % 65536
                if tot >= 5:
                    m = max(range(4), key=lambda b: (c[b], -b))
                    if c[m] / tot > 0.75 and s[p].upper() != "ACGT"[m]:
                        s[p] = "ACGT"[m]
                        st["bases_modified"] += 1
        out.append("".join(s))
    return out, st
