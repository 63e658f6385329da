"""Generates the Stockham kernel instances compiled into libsip_fft.so.

Every leaf length L (a transform that fits one workgroup's LDS) gets one kernel per direction,
with its radix sequence and threads-per-transform fixed at compile time. The leaf set is every
2,3,5-smooth length up to 16384 (powers of two included), and lengths carrying one factor 7, 11
or 13 up to 4096; longer or other lengths are planned on the host as several leaf passes
(four-step) or Bluestein.

Usage: python gen_instances.py OUTDIR NPARTS   (writes OUTDIR/stockham_part_<k>.hip and
OUTDIR/stockham_parts.inc). The output is a pure function of this file.
"""
from __future__ import annotations

import os
import sys

MAX_SMOOTH = 16384
MAX_POW2 = 16384
MAX_PRIME_FACTOR_LEN = 4096


def smooth_lengths():
    out = set()
    for a in range(15):
        for b in range(10):
            for c in range(7):
                n = 2 ** a * 3 ** b * 5 ** c
                if n <= MAX_SMOOTH:
                    out.add(n)
    for a in range(15):
        if 2 ** a <= MAX_POW2:
            out.add(2 ** a)
    base = sorted(x for x in out if x <= MAX_PRIME_FACTOR_LEN)
    for p in (7, 11, 13):
        for m in base:
            if m * p <= MAX_PRIME_FACTOR_LEN:
                out.add(m * p)
    out.discard(1)
    return sorted(out)


def factor(n):
    f = {}
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            f[p] = f.get(p, 0) + 1
            n //= p
    assert n == 1, n
    return f


# Radix sequences chosen by measurement where the cost model's first choice measured slower: tuning
# builds took the second and third sequence by cost (same pass count) for every length above 1024
# where one exists (118 lengths), then the fourth and the best with one pass more (165 lengths),
# each timed as C2C rows of 128 MiB on MI355X, best of two runs per build; kept where the
# alternative was faster by more than 5 % (run-to-run spread: median 0.6 %, at most 3.4 % in the
# first round, 6.2 % in the second, whose kept gains are all above 10 %). Old / new microseconds in
# the comments.
SEQ_OVERRIDE = {
    1120: [16, 10, 7],       # 51.6 / 46.6
    1200: [20, 12, 5],       # 49.6 / 46.6
    1280: [16, 10, 8],       # 46.0 / 40.9
    1320: [15, 11, 8],       # 54.6 / 50.9
    1760: [16, 11, 10],      # 52.3 / 46.5
    1920: [16, 12, 10],      # 51.0 / 47.7
    2100: [10, 10, 7, 3],    # 62.9 / 56.0
    2500: [10, 10, 5, 5],    # 54.3 / 50.4
    2520: [12, 7, 6, 5],     # 62.7 / 56.0
    2916: [12, 9, 9, 3],     # 56.6 / 53.2
    3000: [15, 8, 5, 5],     # 60.2 / 54.0
    3300: [15, 11, 5, 4],    # 58.7 / 55.2
    5760: [16, 10, 9, 4],    # 55.1 / 50.4
    5832: [12, 9, 9, 6],     # 63.6 / 53.1
    6000: [20, 12, 5, 5],    # 66.9 / 56.1
    6144: [16, 16, 6, 4],    # 47.6 / 43.2
    6480: [16, 9, 9, 5],     # 62.9 / 54.5
    6750: [15, 15, 6, 5],    # 63.0 / 57.3
    6912: [16, 9, 8, 6],     # 51.0 / 47.7
    9600: [20, 20, 6, 4],    # 50.7 / 46.9
    11520: [16, 16, 9, 5],   # 66.6 / 54.6
    13122: [9, 9, 9, 6, 3],  # 78.5 / 73.3
    13824: [16, 16, 9, 6],   # 84.9 / 72.4
    # a second round (the fourth sequence by cost, and the best one with one pass more)
    2160: [12, 12, 5, 3],    # 51.6 / 46.5
    14400: [20, 12, 12, 5],  # 66.2 / 57.1
    15360: [20, 12, 4, 4, 4],  # 78.7 / 69.2
    # leaves up to 1024, timed in both four-step roles (n = 256 L, 128 MiB), as rows and as 2D L x L
    720: [12, 12, 5],        # four-step 268 / 237, rows 51.9 / 45.2, 2D 120 / 101
    800: [20, 8, 5],         # four-step 273 / 257, rows 46.1 / 48.8, 2D 124 / 109
    # long-row leaves (SIP_GEN_BIG_RANK tuning builds; C2C rows, 256 MiB, from 256 rows up)
    19440: [16, 15, 9, 9],   # 187 / 148
    20480: [16, 16, 10, 8],  # 158 / 145
    20736: [16, 16, 9, 9],   # 158 / 148
    21600: [16, 15, 15, 6],  # 214 / 163
    25920: [16, 9, 9, 5, 4],  # 206 / 193
    30000: [10, 10, 10, 6, 5],  # 258 / 215
}
# (1000 = 10 x 10 x 10 was faster in the four-step roles, 305 / 283 us, but its 100 threads per
# transform halve the column tiles to 8 transforms, half-line HBM runs: 2D 1000^2 gained nothing.)

# Radices with a register butterfly (butterflies.h): prime, power-of-two and two-factor composites.
RADICES = [20, 16, 15, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2]
MAX_ELEMS = 24   # complex values per thread in one pass


def _sequences(n, maxr=20, depth=0):
    if n == 1:
        yield []
        return
    if depth > 7:
        return
    for r in RADICES:
        if r <= maxr and n % r == 0:
            for rest in _sequences(n // r, r, depth + 1):
                yield [r] + rest


def _cost(n, seq):
    """Passes (each an LDS round trip) plus idle thread-rounds of passes whose butterfly count is
    not a multiple of the thread count, plus the unused lanes of a partial last wave."""
    t = threads_per_transform(n, seq)
    waste, elems = 0.0, 0
    for r in seq:
        nbf = n // r
        nb = -(-nbf // t)
        waste += (nb * t - nbf) / nbf
        elems = max(elems, nb * r)
    if elems > MAX_ELEMS:
        return None
    waves = -(-t // 64)
    partial = (waves * 64 - t) / (waves * 64) if t >= 64 else 0.0
    return len(seq) + 0.3 * waste + 0.2 * partial


def radix_sequence(n):
    """Radix passes for length n. Powers of two: the exponent split as evenly as possible into
    radices <= 16. Other lengths: the sequence over RADICES (composites included) with the lowest
    _cost, ties to the larger leading radices. Largest radices first."""
    if n & (n - 1):
        if n in SEQ_OVERRIDE:
            return list(SEQ_OVERRIDE[n])
        ranked = []
        for seq in _sequences(n):
            c = _cost(n, seq)
            if c is not None:
                ranked.append((c, seq))
        if ranked:
            ranked.sort(key=lambda cs: cs[0])  # stable: ties keep the larger leading radices
            # tuning builds (SIP_GEN_SEQ_RANK=k) take the k-th sequence by cost among those with as
            # few passes as the first
            npass = len(ranked[0][1])
            # tuning builds vary the lengths in SIP_GEN_SEQ_RANGE ("lo,hi", default 1025..16384)
            lo, hi = (int(v) for v in os.environ.get("SIP_GEN_SEQ_RANGE", "1025,16384").split(","))
            tuned = lo <= n <= hi
            if tuned and os.environ.get("SIP_GEN_SEQ_EXTRA") == "1":
                # the best sequence with one pass more, where one exists
                more = [cs for cs in ranked if len(cs[1]) == npass + 1]
                if more:
                    return more[0][1]
            same = [cs for cs in ranked if len(cs[1]) == npass]
            rank = int(os.environ.get("SIP_GEN_SEQ_RANK", "0")) if tuned else 0
            return same[min(rank, len(same) - 1)][1]
    f = factor(n)
    seq = []
    a = f.get(2, 0)
    if a:
        npass = (a + 3) // 4
        bits = [a // npass + (1 if i < a % npass else 0) for i in range(npass)]
        seq += [2 ** b for b in sorted(bits, reverse=True)]
    for p in (13, 11, 7):
        seq += [p] * f.get(p, 0)
    seq += [5] * f.get(5, 0)
    seq += [3] * f.get(3, 0)
    # keep the biggest radix first
    seq.sort(key=lambda r: -r)
    return seq


def threads_per_transform(n, seq):
    """One butterfly of the largest radix per thread, within [1, 1024]; shrink when a smaller
    radix would need more than 32 elements per thread."""
    rmax = max(seq)
    t = max(1, n // rmax)
    while t > 1024:
        t //= 2
    return t


def elems_per_thread(n, seq, t):
    return max(((n // r + t - 1) // t) * r for r in seq)


# 2,3,5-smooth lengths above MAX_SMOOTH up to BIG_ROWS_MAX get only the single-plane contiguous-row
# kernel (one transform per workgroup of T <= 1024 threads; its plane, 4 (L + L/16 + 1) bytes, fits
# the 160 KiB LDS up to 32768): a one-pass leaf where the two-plane tile would not fit. Their radix
# sequences hold at most BIG_MAX_ELEMS values per thread (register budget of four waves per SIMD:
# 77-112 VGPRs, no scratch, with radix-15 and -20 passes as well); lengths with no such sequence
# (31104, 31250, 32000, 32400) stay four-step splits.
BIG_ROWS_MAX = 32768
BIG_RADICES = [20, 16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2]
BIG_MAX_ELEMS = 32


def big_lengths():
    out = []
    for a in range(16):
        for b in range(11):
            for c in range(8):
                n = 2 ** a * 3 ** b * 5 ** c
                if MAX_SMOOTH < n <= BIG_ROWS_MAX:
                    out.append(n)
    return sorted(out)


def big_threads(n, seq):
    """One butterfly of the largest radix per thread, at most 1024 threads."""
    return min(1024, n // max(seq))


def big_radix_sequence(n):
    """Fewest passes over BIG_RADICES within BIG_MAX_ELEMS values per thread, then the fewest idle
    butterfly rounds; largest radices first."""
    if n in SEQ_OVERRIDE:
        return list(SEQ_OVERRIDE[n])
    found = []

    def rec(m, maxr, seq):
        if m == 1:
            t = big_threads(n, seq)
            if elems_per_thread(n, seq, t) > BIG_MAX_ELEMS:
                return
            waste = sum((-(-(n // r) // t) * t - n // r) / (n // r) for r in seq)
            found.append(((len(seq), waste), list(seq)))
            return
        if len(seq) >= 8:
            return
        for r in BIG_RADICES:
            if r <= maxr and m % r == 0:
                rec(m // r, r, seq + [r])

    rec(n, 1 << 30, [])
    if not found:
        return None  # no sequence fits; the length stays a four-step split
    found.sort(key=lambda ks: ks[0])  # stable: ties keep the larger leading radices
    # tuning builds (SIP_GEN_BIG_RANK=k): the k-th sequence among those with the fewest passes
    same = [ks for ks in found if ks[0][0] == found[0][0][0]]
    rank = int(os.environ.get("SIP_GEN_BIG_RANK", "0"))
    return same[min(rank, len(same) - 1)][1]


def instances():
    res = []
    for n in smooth_lengths():
        seq = radix_sequence(n)
        t = threads_per_transform(n, seq)
        e = elems_per_thread(n, seq, t)
        assert e <= 40, (n, seq, t, e)
        assert all(r in RADICES for r in seq), (n, seq)
        res.append((n, t, seq))
    for n in big_lengths():
        seq = big_radix_sequence(n)
        if seq is None:
            continue
        t = big_threads(n, seq)
        assert t >= 32 and elems_per_thread(n, seq, t) <= BIG_MAX_ELEMS, (n, seq, t)
        res.append((n, t, seq))
    return res


# Lengths whose contiguous C2C rows run faster on the two-plane tile than on the single-plane rows
# kernel (no rows kernel is emitted for them, so the planner and the tuner keep the tile). Measured
# on MI355X, 256 MiB of rows, single / two planes: 12000 208 / 179 us, 15000 189 / 164, 15625
# 288 / 240, 16200 233 / 201, 14580 190 / 168, 12150 205 / 181, 8748 141 / 134, 12960 173 / 164,
# 15552 169 / 158; every other length from 4320 to 16384 is faster on one plane (up to 1.7x).
ROWS_TWO_PLANE = {8748, 12000, 12150, 12960, 14580, 15000, 15552, 15625, 16200}

# longest leaf used for column tiles (the host's kColumnLeafMax): these lengths also get the
# single-plane column-tile kernel
COLUMN_LEAF_MAX = 1024
# half lengths (with 32 or more threads per transform) of the real transforms that also get
# single-plane R2C / C2R row kernels: every one up to 8192, and those above 8192 where the single
# plane measured faster than the two-plane fused kernel on MI355X (R2C + C2R of 256 MiB of reals,
# one box: 9216 155 / 211 us, 10000 161 / 226, 16384 131 / 173; it lost for most mixed lengths from
# 12000 up, e.g. 15000 241 / 218, 13824 225 / 212)
REAL_ROWS_MAX = 8192
REAL_ROWS_LONG = {8640, 9000, 9216, 9375, 9600, 9720, 10000, 10125, 10240, 10368, 10800, 10935, 11250, 11664,
                  12288, 16384}
# lengths up to COLUMN_LEAF_MAX with two or more passes and at least this many threads per
# transform (64-byte runs per row read) also get the rows-in / columns-out kernel of the second
# four-step stage
RC_MIN_THREADS = 8


def main():
    outdir = sys.argv[1]
    nparts = int(sys.argv[2])
    os.makedirs(outdir, exist_ok=True)
    inst = instances()
    parts = [[] for _ in range(nparts)]
    # balance by estimated code size (passes x length)
    load = [0] * nparts
    for item in sorted(inst, key=lambda x: -len(x[2]) * x[0]):
        k = load.index(min(load))
        parts[k].append(item)
        load[k] += len(item[2]) * max(64, item[0] // 8)
    for k, items in enumerate(parts):
        items.sort()
        lines = [
            "// Generated by gen_instances.py: Stockham kernel instances (do not edit).",
            '#include "stockham_device.h"',
            "namespace sip {",
        ]
        tab = []
        for n, t, seq in items:
            rs = ", ".join(str(r) for r in seq)
            radix = ", ".join(str(r) for r in seq)
            if n > MAX_SMOOTH:
                rf = f"stockham_rows_kernel<{n}, {t}, -1, {rs}>"
                ri = f"stockham_rows_kernel<{n}, {t}, 1, {rs}>"
                lines.append(f"template __global__ void {rf}(const StockhamArgs);")
                lines.append(f"template __global__ void {ri}(const StockhamArgs);")
                tab.append(f"    {{{n}, {t}, {len(seq)}, {{{radix}}}, nullptr, nullptr, (const void *)&{rf}, "
                           f"(const void *)&{ri}, {{nullptr, nullptr, nullptr, nullptr}}, "
                           f"{{nullptr, nullptr, nullptr, nullptr}}, {{nullptr, nullptr}}}},")
                continue
            kf = f"stockham_kernel<{n}, {t}, -1, {rs}>"
            ki = f"stockham_kernel<{n}, {t}, 1, {rs}>"
            lines.append(f"template __global__ void {kf}(const StockhamArgs);")
            lines.append(f"template __global__ void {ki}(const StockhamArgs);")
            if t >= 16 and n not in ROWS_TWO_PLANE:
                rf = f"stockham_rows_kernel<{n}, {t}, -1, {rs}>"
                ri = f"stockham_rows_kernel<{n}, {t}, 1, {rs}>"
                lines.append(f"template __global__ void {rf}(const StockhamArgs);")
                lines.append(f"template __global__ void {ri}(const StockhamArgs);")
                rows = f"(const void *)&{rf}, (const void *)&{ri}"
            else:
                rows = "nullptr, nullptr"
            if n <= COLUMN_LEAF_MAX:
                names = [f"stockham_cols_kernel<{n}, {t}, {sg}, {tw4}, {rs}>" for tw4 in ("false", "true")
                         for sg in (-1, 1)]
                for nm in names:
                    lines.append(f"template __global__ void {nm}(const StockhamArgs);")
                cols = "{" + ", ".join(f"(const void *)&{nm}" for nm in names) + "}"
            else:
                cols = "{nullptr, nullptr, nullptr, nullptr}"
            if t >= 32 and (n <= REAL_ROWS_MAX or n in REAL_ROWS_LONG):
                names = [f"stockham_real_rows_kernel<{n}, {t}, {sg}, {m}, {rs}>" for m in (1, 2) for sg in (-1, 1)]
                for nm in names:
                    lines.append(f"template __global__ void {nm}(const StockhamArgs);")
                real = "{" + ", ".join(f"(const void *)&{nm}" for nm in names) + "}"
            else:
                real = "{nullptr, nullptr, nullptr, nullptr}"
            if n <= COLUMN_LEAF_MAX and len(seq) >= 2 and t >= RC_MIN_THREADS:
                names = [f"stockham_rc_kernel<{n}, {t}, {sg}, {rs}>" for sg in (-1, 1)]
                for nm in names:
                    lines.append(f"template __global__ void {nm}(const StockhamArgs);")
                rc = "{" + ", ".join(f"(const void *)&{nm}" for nm in names) + "}"
            else:
                rc = "{nullptr, nullptr}"
            tab.append(f"    {{{n}, {t}, {len(seq)}, {{{radix}}}, (const void *)&{kf}, (const void *)&{ki}, {rows}, {cols}, "
                       f"{real}, {rc}}},")
        lines.append(f"extern const StockhamInstance g_stockham_part_{k}[] = {{")
        lines += tab
        lines.append("};")
        lines.append(f"extern const int g_stockham_part_{k}_count = {len(items)};")
        lines.append("}  // namespace sip")
        path = os.path.join(outdir, f"stockham_part_{k}.hip")
        text = "\n".join(lines) + "\n"
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as fh:
                fh.write(text)
    inc = ["// Generated by gen_instances.py (do not edit)."]
    for k in range(nparts):
        inc.append(f"SIP_STOCKHAM_PART({k})")
    text = "\n".join(inc) + "\n"
    path = os.path.join(outdir, "stockham_parts.inc")
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as fh:
            fh.write(text)
    print(f"{len(inst)} leaf lengths in {nparts} parts")


if __name__ == "__main__":
    main()
