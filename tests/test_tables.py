"""Coefficient tables (no GPU needed): every pass twiddle, real-transform root, four-step twiddle and
Bluestein chirp the planner generates is the FP32 value nearest to e^{s 2 pi i m / n}, i.e. within
0.5 ulp of a long-double evaluation of the definition.

The reference builds its W / T / S and mixed-radix tables in FP64 and stores FP32
(core/fft/fftcore/fft_core_common_func.cpp:174-316); this library generates in long double and
rounds once. The expected values here come from numpy's long double cos / sin of the angle reduced
to the first octant-free form 2 pi m / n, with the quarter points exact (cos(pi/2) = 0 exactly), and
the table layouts from the plan description (radix sequence per pass, four-step split).
"""
import numpy as np
import pytest

import sip_amd as S

FWD, INV = S.ASCEND_FFT_FORWARD, S.ASCEND_FFT_INVERSE
LD = np.longdouble
TWO_PI = LD("6.283185307179586476925286766559005768")


def root_ld(m, n, sign):
    """e^{sign 2 pi i m / n} in long double; exact at multiples of a quarter turn."""
    m = np.asarray(m, dtype=np.int64) % n
    a = TWO_PI * m.astype(LD) / LD(n)
    c, s = np.cos(a), np.sin(a)
    q = (4 * m) % n == 0
    quarter = (4 * m) // n
    c = np.where(q, np.select([quarter == 0, quarter == 2], [LD(1), LD(-1)], LD(0)), c)
    s = np.where(q, np.select([quarter == 1, quarter == 3], [LD(1), LD(-1)], LD(0)), s)
    return c, sign * s


def half_ulp(x):
    """Half the FP32 spacing of the binade holding each exact value x (0 for x == 0)."""
    ax = np.abs(x).astype(np.float64)
    e = np.floor(np.log2(np.where(ax > 0, ax, 1.0)))
    e = np.maximum(e, -126)
    return np.where(ax > 0, np.ldexp(1.0, (e - 24).astype(int)), 0.0)


def assert_within_half_ulp(got, re, im, what):
    for part, exact in (("re", re), ("im", im)):
        g = (got.real if part == "re" else got.imag).astype(LD)
        err = np.abs(g - exact).astype(np.float64)
        bound = half_ulp(exact) * (1 + 1e-6)
        bad = np.nonzero(err > bound)[0]
        assert bad.size == 0, (f"{what} {part}: {bad.size} entries beyond 0.5 ulp, first at {bad[0]}: "
                               f"{g[bad[0]]!r} vs {exact[bad[0]]!r}")


def pass_table_expect(L, radices, sign):
    """Layout of the pass twiddles: per pass of radix R after NS = product of the earlier radices,
    rows k < NS, slots W_L^{k r (L / (NS R))} for r = 1, 2, 4, .. < R, padded to an even count."""
    ms, pads = [], []
    ns = 1
    for R in radices:
        step = L // (ns * R)
        nl = 0
        while (1 << nl) < R:
            nl += 1
        slots = (nl + 1) & ~1
        for k in range(ns):
            for sl in range(slots):
                ms.append(k * (1 << sl) * step if sl < nl else 0)
                pads.append(sl >= nl)
        ns *= R
    return np.array(ms, dtype=np.int64), np.array(pads)


@pytest.mark.parametrize("n", [4096, 1024, 1536, 2000, 3000, 8192, 16384, 360, 7 * 512, 13 * 243, 11 * 360, 243,
                               3125])
@pytest.mark.parametrize("direction", [FWD, INV])
def test_pass_twiddles_correctly_rounded(n, direction):
    d = S.describe_dry([n], S.ASCEND_FFT_C2C, direction, 8)
    st = d["stages"][0]
    assert st["L"] == n
    got = S.plan_table_dry([n], S.ASCEND_FFT_C2C, direction, 8, 0, 0)
    sign = -1 if direction == FWD else 1
    ms, pads = pass_table_expect(n, st["radices"], sign)
    if ms.size == 0:
        return
    assert got.size == ms.size, (got.size, ms.size)
    assert np.all(got[pads] == 0)
    re, im = root_ld(ms[~pads], n, sign)
    assert_within_half_ulp(got[~pads], re, im, f"twiddles n={n}")


@pytest.mark.parametrize("n", [8192, 2000, 16384, 12000])
def test_real_roots_correctly_rounded(n):
    got = S.plan_table_dry([n], S.ASCEND_FFT_R2C, FWD, 4, 0, 1)
    L = n // 2
    assert got.size == L + 1
    re, im = root_ld(np.arange(L + 1), 2 * L, -1)
    assert_within_half_ulp(got, re, im, f"real roots n={n}")


@pytest.mark.parametrize("n", [65536, 1 << 20, 36000, 43200])
def test_four_step_twiddles_correctly_rounded(n):
    d = S.describe_dry([n], S.ASCEND_FFT_C2C, FWD, 2)
    idx = next(i for i, s in enumerate(d["stages"]) if s.get("tw4", 0) == n)
    lo = S.plan_table_dry([n], S.ASCEND_FFT_C2C, FWD, 2, idx, 2)
    hi = S.plan_table_dry([n], S.ASCEND_FFT_C2C, FWD, 2, idx, 3)
    shift = int(np.log2(lo.size))
    assert lo.size == 1 << shift and hi.size == -(-n // lo.size)
    re, im = root_ld(np.arange(lo.size), n, -1)
    assert_within_half_ulp(lo, re, im, f"four-step low n={n}")
    re, im = root_ld(np.arange(hi.size) << shift, n, -1)
    assert_within_half_ulp(hi, re, im, f"four-step high n={n}")
    # the two levels compose every W^m, m < n, to within a few ulp of the definition
    m = np.random.default_rng(1).integers(0, n, 4096)
    w = lo[m & (lo.size - 1)].astype(np.complex128) * hi[m >> shift].astype(np.complex128)
    re, im = root_ld(m, n, -1)
    assert np.max(np.abs(w - (re.astype(np.float64) + 1j * im.astype(np.float64)))) < 4 * 2.0 ** -24


@pytest.mark.parametrize("n", [17, 997, 6007])
def test_bluestein_chirp_correctly_rounded(n):
    d = S.describe_dry([n], S.ASCEND_FFT_C2C, FWD, 2)
    assert d["stages"][0]["kind"] == "aux"
    got = S.plan_table_dry([n], S.ASCEND_FFT_C2C, FWD, 2, 0, 4)
    assert got.size == n
    k = np.arange(n, dtype=np.int64)
    re, im = root_ld((k * k) % (2 * n), 2 * n, -1)   # e^{-i pi k^2 / n}
    assert_within_half_ulp(got, re, im, f"chirp n={n}")


def test_half_ulp_check_catches_a_misrounded_entry():
    """The check itself: one entry moved by one ulp fails it; the correctly rounded table passes."""
    re, im = root_ld(np.arange(64), 64, -1)
    t = (re.astype(np.float32) + 1j * im.astype(np.float32)).astype(np.complex64)
    assert_within_half_ulp(t, re, im, "exact")
    bad = t.copy()
    bad[5] = np.complex64(np.nextafter(np.float32(t[5].real), np.float32(2)) + 1j * t[5].imag)
    with pytest.raises(AssertionError):
        assert_within_half_ulp(bad, re, im, "perturbed")
