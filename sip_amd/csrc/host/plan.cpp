// Planner: lays an asdFft* transform out as Stockham / auxiliary kernel stages.
//
// Reference behaviour kept: the set of transforms (1D horizontal / vertical, 2D, 3D; C2C, R2C,
// C2R), their sign and normalisation, the shapes they accept, and the workspace contract
// (caller-allocated device memory of GetWorkspaceSize bytes, needed only by multi-stage plans;
// core/fft/fft_api.cpp:699-733, 989-1045). The stage layout is this library's own:
//   * an axis whose length is a compiled leaf runs as one strided launch, in place when it is not
//     the first stage (rows for the last axis; column tiles, walked with the unit-stride axis
//     fastest, for the others) -- no transposes between axes (the reference inserts AsStrided
//     transposes, fft_api.cpp:478-483, 527-697);
//   * a longer axis runs as a four-step pair L = L1 * L2 (column FFTs with the W_L^{l2 k1}
//     twiddle folded into their store, then row FFTs that store transposed), recursively;
//   * a length with a prime factor above 13 runs through Bluestein's chirp-z convolution with
//     power-of-two FFTs (the reference's generic matrix core FFTCoreAny, fft_core_any.cpp:282-589,
//     plays that role there);
//   * even-length real transforms run a half-length complex FFT with the split post- / pre-process
//     fused into the same launch when the half length is a leaf (reference A/B tables,
//     core/fft/fftcore/fft_core_common_func.cpp:569-673).
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <unordered_map>

namespace sip {

// ====================================================================== tables (host, FP64 -> FP32)

static float2 root_f(int64_t m, int64_t n, int sign)
{
    // e^{sign 2 pi i m / n} in long double, exact at the quarter points, rounded once.
    m %= n;
    if (m < 0) m += n;
    long double c, s;
    if ((4 * m) % n == 0) {
        const int64_t q = (4 * m) / n;
        c = (q == 0) ? 1.0L : (q == 2) ? -1.0L : 0.0L;
        s = (q == 1) ? 1.0L : (q == 3) ? -1.0L : 0.0L;
    } else {
        const long double a = 6.283185307179586476925286766559005768L * (long double)m / (long double)n;
        c = cosl(a);
        s = sinl(a);
    }
    float2 r;
    r.x = (float)c;  // one rounding, long double -> FP32
    r.y = (float)(sign * s);
    return r;
}

// Pass twiddles of a radix sequence of length L. Plain passes: k-major [k][W^1, W^2, W^4, ..] over
// the powers of two below R, padded to an even count (the kernel multiplies the other r together);
// each pass table stays 16-byte aligned. mfma: the radix-16 passes run on the matrix cores, whose
// lanes take every W^{k r}, r = 0..15, of a pass from a [k][16] table.
static TableRef twiddles_for(int64_t L, const std::vector<int> &radices, int sign, bool mfma)
{
    TableRef t;
    t.key = (mfma ? "twm:" : "twe:") + std::to_string(L);
    for (int r : radices) t.key += "," + std::to_string(r);
    t.key += ":" + std::to_string(sign);
    t.gen = [L, radices, sign, mfma]() {
        std::vector<float2> v;
        int64_t ns = 1;
        for (int rr : radices) {
            const int64_t R = rr;
            const int64_t step = L / (ns * R);
            if (mfma && R == 16) {
                for (int64_t k = 0; k < ns; k++)
                    for (int64_t r = 0; r < 16; r++) v.push_back(root_f(k * r * step, L, sign));
                ns *= R;
                continue;
            }
            int nl = 0;
            while ((int64_t(1) << nl) < R) nl++;
            const int slots = (nl + 1) & ~1;
            for (int64_t k = 0; k < ns; k++)
                for (int sl = 0; sl < slots; sl++) {
                    const int64_t r = int64_t(1) << sl;
                    v.push_back(sl < nl ? root_f(k * r * step, L, sign) : float2{0.f, 0.f});
                }
            ns *= R;
        }
        if (v.empty()) v.push_back(root_f(0, 1, 1));
        return v;
    };
    return t;
}

static TableRef pass_twiddles(const StockhamInstance *in, int sign, bool mfma = false)
{
    return twiddles_for(in->L, std::vector<int>(in->radix, in->radix + in->npass), sign, mfma);
}

TableRef pass_twiddles_seq(int64_t L, const std::vector<int> &radices, int sign)
{
    return twiddles_for(L, radices, sign, false);
}

int64_t pass_twiddle_rows(int R)
{
    int nl = 0;
    while ((1 << nl) < R) nl++;
    return (nl + 1) & ~1;
}

static TableRef real_roots(int64_t L, int sign)
{
    // w^k = e^{sign 2 pi i k / (2L)}, k = 0..L
    TableRef t;
    t.key = "rtw:" + std::to_string(L) + ":" + std::to_string(sign);
    t.gen = [L, sign]() {
        std::vector<float2> v((size_t)L + 1);
        for (int64_t k = 0; k <= L; k++) v[(size_t)k] = root_f(k, 2 * L, sign);
        return v;
    };
    return t;
}

static int tw4_shift_for(int64_t ntot)
{
    int bits = 0;
    while ((int64_t(1) << bits) < ntot) bits++;
    return (bits + 1) / 2;
}

static void four_step_tables(int64_t ntot, int sign, TableRef &lo, TableRef &hi)
{
    const int shift = tw4_shift_for(ntot);
    lo.key = "tw4lo:" + std::to_string(ntot) + ":" + std::to_string(sign);
    lo.gen = [ntot, sign, shift]() {
        std::vector<float2> v((size_t)1 << shift);
        for (int64_t m = 0; m < (int64_t(1) << shift); m++) v[(size_t)m] = root_f(m, ntot, sign);
        return v;
    };
    hi.key = "tw4hi:" + std::to_string(ntot) + ":" + std::to_string(sign);
    hi.gen = [ntot, sign, shift]() {
        const int64_t nh = (ntot + (int64_t(1) << shift) - 1) >> shift;
        std::vector<float2> v((size_t)nh);
        for (int64_t m = 0; m < nh; m++) v[(size_t)m] = root_f(m << shift, ntot, sign);
        return v;
    };
}

// In-place FP64 radix-2 FFT (host; plan-time table generation only).
static void fft_pow2_f64(std::vector<double> &re, std::vector<double> &im, int sign)
{
    const size_t n = re.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const long double ang = sign * 6.283185307179586476925286766559005768L / (long double)len;
        std::vector<double> wr(len / 2), wi(len / 2);
        for (size_t k = 0; k < len / 2; k++) {
            wr[k] = (double)cosl(ang * (long double)k);
            wi[k] = (double)sinl(ang * (long double)k);
        }
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; k++) {
                const double ur = re[i + k], ui = im[i + k];
                const double xr = re[i + k + len / 2], xi = im[i + k + len / 2];
                const double vr = xr * wr[k] - xi * wi[k], vi = xr * wi[k] + xi * wr[k];
                re[i + k] = ur + vr;
                im[i + k] = ui + vi;
                re[i + k + len / 2] = ur - vr;
                im[i + k + len / 2] = ui - vi;
            }
    }
}

static TableRef bluestein_chirp(int64_t N, int sign)
{
    // c[n] = e^{sign i pi n^2 / N}, with n^2 reduced mod 2N exactly
    TableRef t;
    t.key = "bchirp:" + std::to_string(N) + ":" + std::to_string(sign);
    t.gen = [N, sign]() {
        std::vector<float2> v((size_t)N);
        for (int64_t n = 0; n < N; n++) {
            const int64_t m = (int64_t)(((__int128)n * n) % (2 * N));
            v[(size_t)n] = root_f(m, 2 * N, sign);
        }
        return v;
    };
    return t;
}

static TableRef bluestein_kernel(int64_t N, int64_t M, int sign)
{
    // D = FFT_M(d) / M with d[m] = conj c[|m|] for |m| < N (circularly wrapped), in FP64.
    TableRef t;
    t.key = "bkern:" + std::to_string(N) + ":" + std::to_string(M) + ":" + std::to_string(sign);
    t.gen = [N, M, sign]() {
        std::vector<double> re((size_t)M, 0.0), im((size_t)M, 0.0);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int64_t n = 0; n < N; n++) {
            const int64_t m2 = (int64_t)(((__int128)n * n) % (2 * N));
            const long double a = pi * (long double)m2 / (long double)N;
            const double cr = (double)cosl(a), ci = (double)(-sign * sinl(a));
            re[(size_t)n] = cr;
            im[(size_t)n] = ci;
            if (n > 0) {
                re[(size_t)(M - n)] = cr;
                im[(size_t)(M - n)] = ci;
            }
        }
        fft_pow2_f64(re, im, -1);
        std::vector<float2> v((size_t)M);
        for (int64_t k = 0; k < M; k++) {
            v[(size_t)k].x = (float)(re[(size_t)k] / (double)M);
            v[(size_t)k].y = (float)(im[(size_t)k] / (double)M);
        }
        return v;
    };
    return t;
}

// ====================================================================== layouts

namespace {

struct Dim {
    int64_t n, s;
    int tab = -1;  // row-table id: offset = table[index] instead of index * s
};

// A length-L axis of an array: element j of transform f sits at base(f) + j * es, with the
// transforms enumerated by `dims` (fastest first).
struct Lay {
    BufRef buf;
    int64_t es = 1;
    std::vector<Dim> dims;
};

std::vector<Dim> normalize(const std::vector<Dim> &in)
{
    std::vector<Dim> d;
    for (const Dim &x : in)
        if (x.n != 1 || x.tab >= 0) d.push_back(x);
    std::vector<Dim> m;
    for (const Dim &x : d) {
        if (!m.empty() && x.tab < 0 && m.back().tab < 0 && x.s == m.back().n * m.back().s) {
            m.back().n *= x.n;
        } else {
            m.push_back(x);
        }
    }
    if (m.empty()) m.push_back({1, 0});
    return m;
}

bool has_table(const std::vector<Dim> &d)
{
    for (const Dim &x : d)
        if (x.tab >= 0) return true;
    return false;
}

int64_t count(const std::vector<Dim> &d)
{
    int64_t c = 1;
    for (const Dim &x : d) c *= x.n;
    return c;
}

// Insert the dim for a split axis so that dims stay ordered by stride (unit stride first).
size_t insert_pos(const std::vector<Dim> &dims, int64_t stride)
{
    size_t pos = 0;  // gathered (table) dimensions count as the outermost
    while (pos < dims.size() && dims[pos].tab < 0 && dims[pos].s < stride) pos++;
    return pos;
}

std::vector<Dim> with_dim(std::vector<Dim> dims, size_t pos, Dim d)
{
    dims.insert(dims.begin() + (std::ptrdiff_t)pos, d);
    return dims;
}

TileDesc to_tile(const Lay &l, bool *ok, int *tab_id)
{
    std::vector<Dim> d = normalize(l.dims);
    TileDesc t{};
    *tab_id = -1;
    if (d.size() > 3) {
        *ok = false;
        return t;
    }
    while (d.size() < 3) d.push_back({1, 0});
    t.n0 = d[0].n;
    t.n1 = d[1].n;
    t.s0 = d[0].s;
    t.s1 = d[1].s;
    t.s2 = d[2].s;
    t.es = l.es;
    t.tab = nullptr;
    t.tab_dim = -1;
    for (int i = 0; i < 3; i++) {
        if (d[(size_t)i].tab >= 0) {
            if (*tab_id >= 0) *ok = false;  // one gathered dimension per tile
            *tab_id = d[(size_t)i].tab;
            t.tab_dim = i;
        }
    }
    t.unit_f = (d[0].s == 1 && d[0].n > 1 && l.es != 1 && d[0].tab < 0) ? 1 : 0;
    // gathered rows keep 8-byte accesses (row offsets need not be even)
    t.vec2 = (*tab_id < 0 && l.es == 1 && (d[0].s % 2 == 0) && (d[1].s % 2 == 0) && (d[2].s % 2 == 0)) ? 1 : 0;
    t.div_n0 = make_fastdiv((uint32_t)t.n0);
    t.div_n1 = make_fastdiv((uint32_t)t.n1);
    return t;
}

AuxRows to_rows(const Lay &l, bool *ok, int *tab_id)
{
    std::vector<Dim> d = normalize(l.dims);
    AuxRows r{};
    *tab_id = -1;
    if (d.size() > 2) {
        *ok = false;
        return r;
    }
    while (d.size() < 2) d.push_back({1, 0});
    r.n0 = d[0].n;
    r.s0 = d[0].s;
    r.s1 = d[1].s;
    r.es = l.es;
    r.div_n0 = make_fastdiv((uint32_t)r.n0);
    r.tab = nullptr;
    r.tab_dim = -1;
    for (int i = 0; i < 2; i++) {
        if (d[(size_t)i].tab >= 0) {
            if (*tab_id >= 0) *ok = false;  // one gathered dimension per side
            *tab_id = d[(size_t)i].tab;
            r.tab_dim = i;
        }
    }
    return r;
}

bool env_flag(const char *name, bool dflt)
{
    const char *e = std::getenv(name);
    if (!e || !e[0]) return dflt;
    return e[0] != '0';
}

constexpr int64_t kColumnLeafMax = 1024;
constexpr int64_t kRowLeafMax = 16384;
// single-pass leaves above kRowLeafMax: transforms per launch from which they replace the split
// (the plan's own choice, else 256; the tuning switch SIP_FFT_BIG_LEAF_ROWS overrides both)
static int64_t big_leaf_min_rows(const Plan &p)
{
    const char *e = std::getenv("SIP_FFT_BIG_LEAF_ROWS");
    if (e && e[0]) return std::atoll(e);
    return p.big_leaf_rows > 0 ? p.big_leaf_rows : 256;
}
constexpr int64_t kFillWorkgroups = 512;  // two workgroups per CU of MI355X (256 CUs)

struct Builder {
    Plan &p;
    int64_t ws_count = 0;
    // C2C_SEP plans: the user's input and output (BUF_IN / BUF_OUT) are two float planes each
    bool sep = false;
    explicit Builder(Plan &pl) : p(pl) {}

    bool is_user(const BufRef &b) const { return sep && (b.kind == BUF_IN || b.kind == BUF_OUT); }

    BufRef new_ws(size_t bytes)
    {
        const size_t align = 256;
        size_t off = (p.ws_bytes + align - 1) / align * align;
        p.ws_offsets.push_back(off);
        p.ws_bytes = off + (bytes + align - 1) / align * align;
        BufRef b;
        b.kind = BUF_WS;
        b.ws_index = (int)p.ws_offsets.size() - 1;
        return b;
    }

    static bool is_column(const Lay &l)
    {
        std::vector<Dim> d = normalize(l.dims);
        return l.es != 1 && d[0].s == 1 && d[0].n > 1;
    }

    bool leaf_ok(int64_t L, const Lay &in, const Lay &out) const
    {
        const StockhamInstance *inst = find_instance(L);
        if (!inst) return false;
        if (is_column(in) || is_column(out)) return L <= kColumnLeafMax;
        if (L <= kRowLeafMax) return true;
        // Lengths above 16384 (2,3,5-smooth, up to 32768): one workgroup of up to 1024 threads per
        // transform on the single-plane rows kernel (plane of 4 (L + L/16 + 1) bytes <= 136 KiB),
        // for contiguous interleaved FP32 rows on both sides, from 256 transforms up (one
        // workgroup per CU); everything else takes a four-step split. Measured on MI355X, one launch
        // against the two four-step launches: 32768 x 4096 0.54 / 0.81 ms, 20000 x 6710 0.83 /
        // 1.39 ms, 30375 x 4418 1.11 / 1.55 ms; with few transforms the split's parallelism wins
        // (32768 x 16 23 / 14 us, 20000 x 64 25 / 20 us).
        return inst->rows_fwd != nullptr && in.es == 1 && out.es == 1 && !is_user(in.buf) && !is_user(out.buf) &&
               p.precision == PREC_FP32 && p.big_leaf && count(normalize(in.dims)) >= big_leaf_min_rows(p) &&
               env_flag("SIP_FFT_BIG_LEAF", true);
    }

    void add_stockham(int64_t L, int sign, const Lay &in, const Lay &out, int io_mode, int64_t in_len,
                      int64_t out_len, int64_t tw4_ntot, int64_t col_div, const std::string &note)
    {
        const StockhamInstance *inst = find_instance(L);
        // precision form: C2C launches with radix-16 stages run them on the matrix cores
        const bool in_sep = is_user(in.buf), out_sep = is_user(out.buf);
        if ((in_sep || out_sep) && io_mode != IO_C2C)
            throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "split planes on a real-transform stage");
        const StockhamInstance *mf = (p.precision != PREC_FP32 && io_mode == IO_C2C && !in_sep && !out_sep)
                                         ? find_mfma_instance(L, p.precision) : nullptr;
        if (mf) inst = mf;
        if (!inst) throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "no kernel for length " + std::to_string(L));
        bool ok = true;
        Stage st;
        st.kind = StageKind::Stockham;
        st.in = in.buf;
        st.out = out.buf;
        st.inst = inst;
        st.sign = sign;
        StockhamArgs &a = st.sa;
        std::memset(&a, 0, sizeof(a));
        a.tin = to_tile(in, &ok, &st.in_tab);
        a.tout = to_tile(out, &ok, &st.out_tab);
        if (!ok) throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "layout has more than 3 batch dimensions");
        const int64_t nfft = count(normalize(in.dims));
        if (nfft != count(normalize(out.dims)))
            throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "stage in/out transform counts differ");
        if (nfft >= (int64_t(1) << 31))
            throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "more than 2^31 transforms in one launch");
        a.nfft = nfft;
        a.io_mode = io_mode;
        a.in_len = (int32_t)in_len;
        a.out_len = (int32_t)out_len;
        const int64_t maxidx = (io_mode == IO_C2R_EVEN) ? L : L - 1;
        int lp = (int)(maxidx + (maxidx >> 4)) + 1;
        if ((lp & 1) == 0) lp++;
        a.LP = lp;
        const int T = inst->T;
        const bool column = a.tin.unit_f || a.tout.unit_f;
        const size_t per = (size_t)2 * lp * 4 + 24;
        int F;
        if (column) {
            // Tile width: a multiple of 16 transforms (each row run covers whole 128-byte lines),
            // about 512 threads per workgroup, at least 16 wide even when that takes the whole
            // LDS (1024 x 16: 947 us against 1126 us at width 8 per 1 GiB 2D pass on MI355X),
            // at most 64.
            int want = p.tune.col_F;
            if (want <= 0) {
                if (const char *e = std::getenv("SIP_FFT_COL_F")) want = std::atoi(e);
            }
            // Mixed lengths: 16 wide (whole lines, small LDS). Measured on MI355X, four-step
            // C2C rows: 43200 x 1056 459 -> 401 us, 36000 x 864 350 -> 277 us, 27000 x 604
            // 168 -> 154 us; power-of-two columns keep the wider tiles (the 256-row stage B of
            // 64000 x 635: 360 us at 32 wide, 393 us at 16).
            const bool pow2 = (L & (L - 1)) == 0;
            F = want > 0 ? want : (pow2 ? std::min(64, std::max(16, (512 / T) & ~15)) : 16);
            if (F >= 16) F &= ~15;
            while (F > 1 && ((size_t)F * per > 160 * 1024 || F * T > 1024)) F /= 2;
        } else {
            F = std::max(1, 256 / T);
            while (F > 1 && ((size_t)F * per > 48 * 1024 || F * T > 1024)) F--;
            int f = p.tune.row_F;
            if (f <= 0) {
                if (const char *e = std::getenv("SIP_FFT_ROW_F")) f = std::atoi(e);
            }
            if (f >= 1 && f * T <= 1024 && (size_t)f * per <= 160 * 1024) {
                F = f;
            } else {
                // Small batches: fewer transforms per workgroup until the grid has two workgroups
                // per CU (256 CUs on MI355X), down to one wave per workgroup (1024 x 256: 256
                // workgroups of one transform instead of 64 of four).
                while (F > 1 && ((F + 1) / 2) * T >= 64 && (nfft + F - 1) / F < kFillWorkgroups) F = (F + 1) / 2;
            }
        }
        // column tiles of the precision form: single-plane tiles of 16, 32 or 64 adjacent
        // transforms whose radix-16 passes take whole 16-transform blocks (K = 32 instructions)
        const bool mf_cols = mf != nullptr && column && a.tin.unit_f && a.tout.unit_f && tw4_ntot == 0 &&
                             !in_sep && !out_sep && inst->cols[0] != nullptr && F >= 16 && F % 16 == 0 &&
                             F * T <= 1024 && (size_t)F * lp * sizeof(float) <= 160 * 1024 &&
                             env_flag("SIP_FFT_MF_COLS", true);
        if (mf && !mf_cols) {
            // every wave owns 4 blocks of 16 butterflies: F * T must be a multiple of 64
            while ((F * T) % 64 != 0) F++;
            while (F > 1 && ((size_t)F * per > 80 * 1024 || F * T > 1024) && ((F / 2) * T) % 64 == 0) F /= 2;
            if (F * T > 1024 || (size_t)F * per > 160 * 1024)
                throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "no MFMA tile fits length " + std::to_string(L));
        } else if (!mf_cols) {
            if (F > nfft) F = (int)nfft;
            if (F < 1) F = 1;
        }
        a.F = F;
        a.div_in_len = make_fastdiv((uint32_t)in_len);
        a.div_out_len = make_fastdiv((uint32_t)out_len);
        a.div_F = make_fastdiv((uint32_t)F);
        a.div_in_vec = make_fastdiv((uint32_t)std::max<int64_t>(1, in_len / 2));
        a.div_out_vec = make_fastdiv((uint32_t)std::max<int64_t>(1, out_len / 2));
        a.div_col = make_fastdiv((uint32_t)std::max<int64_t>(1, col_div));
        if (in_len % 2) a.tin.vec2 = 0;
        if (out_len % 2) a.tout.vec2 = 0;
        if (io_mode == IO_R2C_FULL) a.tin.vec2 = 0;
        if (io_mode == IO_C2R_FULL) a.tout.vec2 = 0;
        // First / last radix pass straight from / to HBM for contiguous complex rows, when a
        // transform has enough threads for its accesses to coalesce.
        a.direct_in = (io_mode == IO_C2C || io_mode == IO_R2C_EVEN) && a.tin.es == 1 && !a.tin.unit_f &&
                      (T >= 16 || mf != nullptr);
        a.direct_out = (io_mode == IO_C2C || io_mode == IO_C2R_EVEN) && a.tout.es == 1 && !a.tout.unit_f &&
                       (T >= 16 || mf != nullptr) && tw4_ntot == 0;
        // split-plane user buffers go through the LDS-staged load / store
        a.in_sep = in_sep ? 1 : 0;
        a.out_sep = out_sep ? 1 : 0;
        if (in_sep) {
            a.direct_in = 0;
            a.tin.vec2 = 0;
        }
        if (out_sep) {
            a.direct_out = 0;
            a.tout.vec2 = 0;
        }
        const bool rows = io_mode == IO_C2C && a.direct_in && a.direct_out && inst->rows_fwd != nullptr;
        // column tiles (C2C, strided on both sides): first / last pass straight from / to HBM with
        // lanes walking adjacent columns
        a.col_direct = (io_mode == IO_C2C && a.tin.unit_f && a.tout.unit_f && (!mf || mf_cols) && !in_sep && !out_sep &&
                        env_flag("SIP_FFT_COL_DIRECT", true)) ? 1 : 0;
        // Single-plane exchange (half the LDS, twice the resident workgroups) by default for rows
        // with 32 or more threads per transform. Measured on MI355X, 1 GiB of C2C rows: 4096
        // 373 us against 397 us, 1024 370 / 395, 8192 428 / 497, 16384 436 / 646; at 16 threads
        // (length 256) the two-plane kernel stays ahead (371 / 381). Launch order beats
        // XCD-contiguous tile order (4096 x 16384: 0.194 against 0.200 ms); the latter stays
        // available as the tuning switch SIP_FFT_XCD=1.
        a.split_lds = rows && (p.tune.split_lds >= 0 ? p.tune.split_lds == 1
                                                     : env_flag("SIP_FFT_SPLIT_LDS", T >= 32 || mf != nullptr));
        if (L > kRowLeafMax) {
            // single-pass leaves above 16384 exist only as the single-plane rows kernel
            if (!rows)
                throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "length " + std::to_string(L) + " leaf off contiguous rows");
            a.split_lds = 1;
        }
        // Column tiles with direct HBM passes: single plane too (a 1024 x 16 tile needs 70 KiB
        // instead of 139 KiB, so two workgroups fit on a CU). Measured on MI355X: 2D 1024^2 x 128
        // 892 -> 863 us, 1024^2 x 8 52 -> 48 us, 512^2 x 256 445 -> 422 us, 2^20 x 64 606 -> 566 us;
        // with 16-wide mixed-length tiles and the derived four-step twiddles, mixed lengths gain
        // too: four-step 43200 x 1056 387-399 -> 346 us, 64000 x 635 370 -> 337 us, 2D 1000^2 x 64
        // 528-535 -> 500 us, 3D 240^3 x 8 1155 -> 1086 us.
        {
            const char *e = std::getenv("SIP_FFT_SPLIT_COLS");
            const int mode = e ? std::atoi(e) : 2;  // 0 never, 1 power-of-two lengths, 2 every length
            if (a.col_direct && inst->cols[0] != nullptr && (mode == 2 || (mode == 1 && (L & (L - 1)) == 0)))
                a.split_lds = 1;
        }
        // Fused real transforms on contiguous rows: single plane as well (pair pre-process read
        // straight from HBM for C2R, post-process through the plane for R2C). Measured on MI355X:
        // 8192 x 4096 R2C 58.1 -> 47.5 us, C2R 58.8 -> 46.7 us; 12000 x 2048 77 / 82 -> 68 / 67 us;
        // 2000 x 8192 40 / 38 -> 36 / 34 us; 3000 and 6000 R2C unchanged.
        const bool real_rows = (io_mode == IO_R2C_EVEN && a.direct_in && a.tout.es == 1 && !a.tout.unit_f) ||
                               (io_mode == IO_C2R_EVEN && a.direct_out && a.tin.es == 1 && !a.tin.unit_f);
        if (real_rows && inst->real_rows[0] != nullptr && env_flag("SIP_FFT_SPLIT_REAL", true)) a.split_lds = 1;
        // Rows in, columns out (the second four-step stage: contiguous rows of the intermediate,
        // transposed store): one plane, the first pass straight from HBM along each row, the last
        // pass straight to HBM across adjacent transforms (instead of the two-plane tile with
        // LDS-staged store).
        st.rc = io_mode == IO_C2C && !mf && !in_sep && !out_sep && a.tin.es == 1 && !a.tin.unit_f &&
                a.tout.unit_f && tw4_ntot == 0 && inst->rc[0] != nullptr &&
                (p.tune.rc >= 0 ? p.tune.rc == 1 : env_flag("SIP_FFT_RC", true));
        if (st.rc) {
            a.split_lds = 0;
            a.col_direct = 0;
        }
        a.xcd_remap = p.tune.xcd >= 0 ? p.tune.xcd == 1 : env_flag("SIP_FFT_XCD", false);
        // matrix passes read [k][16] twiddle rows; the single-plane rows kernel (K = 32) reads the
        // interleaved operand table, the two-plane tile (K = 16) the blocked one
        const bool mf2 = mf != nullptr && a.split_lds;
        st.tw = pass_twiddles(inst, sign, mf != nullptr);
        if (mf) st.mfw = mf2 ? mfma2_dft_table(p.precision, sign) : mfma_dft_table(p.precision, sign);
        if (io_mode == IO_R2C_EVEN || io_mode == IO_C2R_EVEN) st.rtw = real_roots(L, sign);
        if (tw4_ntot) {
            four_step_tables(tw4_ntot, sign, st.tw4lo, st.tw4hi);
            a.tw4_ntot = tw4_ntot;
            const int shift = tw4_shift_for(tw4_ntot);
            a.tw4_shift = shift;
            a.tw4_mask = (int64_t(1) << shift) - 1;
        }
        a.col_fixed = ((F * T) % F) == 0 ? 1 : 0;  // always true: block = F * T
        st.grid = dim3((unsigned)((nfft + F - 1) / F));
        st.block = dim3((unsigned)(F * T));
        st.lds = (a.split_lds || st.rc) ? (size_t)F * lp * sizeof(float)
                                        : (size_t)2 * F * lp * sizeof(float) + (size_t)3 * F * sizeof(int64_t);
        if (const char *e = std::getenv("SIP_FFT_LDS_MIN_KB")) {
            // occupancy throttle for tuning: allocate at least this much LDS per workgroup
            const size_t m = (size_t)std::atoi(e) * 1024;
            if (m > st.lds && m <= 160 * 1024) st.lds = m;
        }
        st.note = note;
        p.stages.push_back(st);
    }

    void add_aux(int mode, const Lay &in, const Lay &out, int64_t n, int64_t out_len, TableRef tab, float scale,
                 const std::string &note)
    {
        bool ok = true;
        Stage st;
        st.kind = StageKind::Aux;
        st.in = in.buf;
        st.out = out.buf;
        AuxArgs &x = st.xa;
        std::memset(&x, 0, sizeof(x));
        x.mode = mode;
        x.rin = to_rows(in, &ok, &st.in_tab);
        x.rout = to_rows(out, &ok, &st.out_tab);
        if (!ok) throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "aux layout has more than 2 batch dimensions");
        x.rows = count(normalize(in.dims));
        x.n = n;
        x.out_len = out_len;
        x.scale = scale;
        x.in_sep = is_user(in.buf) ? 1 : 0;
        x.out_sep = is_user(out.buf) ? 1 : 0;
        if ((x.in_sep && mode != AUX_CPLX_TRUNC && mode != AUX_BLUE_PRE) ||
            (x.out_sep && mode != AUX_CPLX_TRUNC && mode != AUX_BLUE_POST))
            throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "split planes on an unsupported auxiliary stage");
        st.tab0 = tab;
        st.note = note;
        p.stages.push_back(st);
    }

    // Fused 2D tiles: both axes of `slices` contiguous X x Y slices in one launch (in -> out).
    void add_tile2d(const Stockham2dInstance *t, int sign, const BufRef &in, const BufRef &out, int64_t slices,
                    const std::string &note)
    {
        if (slices >= (int64_t(1) << 31))
            throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "more than 2^31 slices in one launch");
        Stage st;
        st.kind = StageKind::Stockham;
        st.in = in;
        st.out = out;
        st.inst = find_instance(t->Y);
        st.inst2d = t;
        st.sign = sign;
        std::memset(&st.sa, 0, sizeof(st.sa));
        st.sa.nfft = slices;
        st.sa.io_mode = IO_C2C;
        st.sa.F = 1;
        st.sa.in_len = st.sa.out_len = t->X * t->Y;
        st.grid = dim3((unsigned)slices);
        st.block = dim3((unsigned)t->threads);
        st.lds = (size_t)t->lds_bytes;
        st.tw = pass_twiddles(find_instance(t->Y), sign);
        st.tw2 = pass_twiddles(find_instance(t->X), sign);
        st.note = note;
        p.stages.push_back(st);
    }

    // Two-leaf split L = L1 * L2 for a four-step pair. Stage B stores transposed (a column tile),
    // so both factors must be column leaves. Among the pairs within 4x of balanced the choice is, in
    // this order:
    //  1. HBM line alignment: stage A reads and writes runs of 16 adjacent columns at a stride of L2
    //     elements and stage B writes them at a stride of L1, so a factor that is a multiple of 16
    //     (8) keeps those runs on 128-byte lines (half lines); the smaller of the two factors'
    //     alignments counts, and below 8 elements all count alike;
    //  2. fewer radix passes in total;
    //  3. balance (|log L1 - log L2|), then the smaller L1.
    // Measured on MI355X over every 2,3,5-smooth length from 32805 to 65536 (128 MiB each), against
    // the most balanced pair: 62208 = 144 x 432 100 us against 243 x 256 128 us, 64000 = 160 x 400
    // against 250 x 256 (1.24x), 55296 1.20x, 34560 1.19x, 41472 1.16x, 34992 1.14x, 38400 1.13x;
    // the balance cap and the smaller-L1 tie-break keep the rest where the most balanced pair was.
    static int line_align(int64_t n)
    {
        int a = 1;
        while (a < 16 && n % (2 * a) == 0) a *= 2;
        return a < 8 ? 0 : a;
    }

    static bool leaf_split(int64_t L, int64_t &L1, int64_t &L2)
    {
        L1 = L2 = 0;
        int best_align = -1, best_pass = 0;
        double best_bal = 0;
        for (int pass = 0; pass < 2 && L1 == 0; pass++) {
            // second round without the balance cap (lengths with no pair within 4x)
            const double cap = pass == 0 ? std::log(4.0) + 1e-9 : 1e30;
            for (int64_t a = 2; a <= kColumnLeafMax && a < L; a++) {
                if (L % a) continue;
                const int64_t b = L / a;
                const StockhamInstance *ia = find_instance(a), *ib = find_instance(b);
                if (b > kColumnLeafMax || !ia || !ib) continue;
                const double bal = std::fabs(std::log((double)a) - std::log((double)b));
                if (bal > cap) continue;
                const int align = std::min(line_align(a), line_align(b));
                const int passes = ia->npass + ib->npass;
                bool better;
                if (align != best_align)
                    better = align > best_align;
                else if (passes != best_pass)
                    better = passes < best_pass;
                else
                    better = bal < best_bal - 1e-9;  // ties keep the smaller L1 (found first)
                if (better) {
                    best_align = align;
                    best_pass = passes;
                    best_bal = bal;
                    L1 = a;
                    L2 = b;
                }
            }
        }
        return L1 != 0;
    }

    // Can the axis be done with leaves (possibly split recursively)?
    static bool decomposable(int64_t L, int depth = 0)
    {
        if (L <= kRowLeafMax && find_instance(L)) return true;
        if (depth > 3) return false;
        int64_t a, b;
        if (leaf_split(L, a, b)) return true;
        for (int64_t l1 = std::min(L, kColumnLeafMax); l1 >= 2; l1--)
            if (L % l1 == 0 && find_instance(l1) && decomposable(L / l1, depth + 1)) return true;
        return false;
    }

    // Complex transform of length L along the axis described by in/out (same transform order).
    // `in_scratch`: the input buffer may be overwritten.
    void axis(int64_t L, int sign, const Lay &in, const Lay &out, bool in_scratch, int depth = 0)
    {
        if (L == 1) {
            if (in.buf.kind != out.buf.kind || in.buf.ws_index != out.buf.ws_index)
                add_aux(AUX_CPLX_TRUNC, in, out, 1, 1, TableRef(), 1.f, "copy (length 1)");
            return;
        }
        // a tuned split (split_l1) of a length above 16384 takes precedence over its single-pass leaf
        const bool tuned_split = depth == 0 && L > kRowLeafMax && p.tune.split_l1 > 1;
        if (!tuned_split && leaf_ok(L, in, out)) {
            add_stockham(L, sign, in, out, IO_C2C, L, L, 0, 1, "leaf " + std::to_string(L));
            return;
        }
        if (!decomposable(L)) {
            bluestein(L, sign, in, out);
            return;
        }
        // Split L = L1 * L2: L1 column FFTs (+ twiddle), then L2 row FFTs with transposed store.
        int64_t L1 = 0, L2 = 0;
        int64_t forced = p.tune.split_l1;
        if (forced <= 1) {
            if (const char *e = std::getenv("SIP_FFT_SPLIT_L1")) forced = std::atoll(e);  // tuning switch
        }
        if (depth == 0 && forced > 1 && L % forced == 0 && forced <= kColumnLeafMax && find_instance(forced) &&
            decomposable(L / forced, depth + 1)) {
            L1 = forced;
            L2 = L / L1;
        } else if (!leaf_split(L, L1, L2)) {
            for (int64_t l1 = std::min(L, kColumnLeafMax); l1 >= 2; l1--)
                if (L % l1 == 0 && find_instance(l1) && decomposable(L / l1, depth + 1)) {
                    L1 = l1;
                    L2 = L / l1;
                    break;
                }
        }
        if (!L1) throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "cannot split length " + std::to_string(L));
        // mid layout: same structure as the input; in place when the input may be overwritten and
        // the final store does not go back into that same buffer (stage B is not in-place safe)
        Lay mid = in;
        const bool same_buf = in.buf.kind == out.buf.kind && (in.buf.kind != BUF_WS || in.buf.ws_index == out.buf.ws_index);
        if (has_table(in.dims)) {
            // gathered rows: the intermediate is a compact [rows][L] array
            const int64_t rows = count(normalize(in.dims));
            mid = Lay{new_ws((size_t)rows * L * 8), 1, {}};
            int64_t acc = L;
            for (const Dim &d : in.dims) {
                mid.dims.push_back({d.n, acc});
                acc *= d.n;
            }
        } else if (!in_scratch || same_buf) {
            int64_t extent = L * in.es;
            for (const Dim &d : in.dims) extent = std::max(extent, d.s * d.n);
            mid.buf = new_ws((size_t)extent * 8);
        }
        // stage A: element k1 stride es*L2, transforms (l2, batch) with l2 stride es
        const size_t posA = insert_pos(in.dims, in.es);
        Lay ain{in.buf, in.es * L2, with_dim(in.dims, posA, {L2, in.es})};
        Lay aout{mid.buf, mid.es * L2, with_dim(mid.dims, posA, {L2, mid.es})};
        add_stockham(L1, sign, ain, aout, IO_C2C, L1, L1, L, in.es,
                     "four-step A: " + std::to_string(L1) + " columns x twiddle W_" + std::to_string(L));
        // stage B: rows l2 (stride es) of each k1, stored at k1 + L1*k2
        const size_t posB = insert_pos(out.dims, out.es);
        Lay bin{mid.buf, mid.es, with_dim(mid.dims, posB, {L1, mid.es * L2})};
        Lay bout{out.buf, out.es * L1, with_dim(out.dims, posB, {L1, out.es})};
        axis(L2, sign, bin, bout, true, depth + 1);
    }

    void bluestein(int64_t N, int sign, const Lay &in, const Lay &out)
    {
        int64_t M = 1;
        while (M < 2 * N - 1) M <<= 1;
        if (M > (int64_t(1) << 26))
            throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "Bluestein length too large");
        const int64_t rows = count(normalize(in.dims));
        BufRef w = new_ws((size_t)rows * M * 8);
        Lay wl{w, 1, {{rows, M}}};
        TableRef chirp = bluestein_chirp(N, sign);
        add_aux(AUX_BLUE_PRE, Lay{in.buf, in.es, in.dims}, wl, N, M, chirp, 1.f, "bluestein chirp " + std::to_string(N));
        axis(M, -1, wl, wl, true);
        add_aux(AUX_BLUE_MUL, wl, wl, M, M, bluestein_kernel(N, M, sign), 1.f, "bluestein kernel");
        axis(M, 1, wl, wl, true);
        add_aux(AUX_BLUE_POST, wl, Lay{out.buf, out.es, out.dims}, N, N, chirp, 1.f, "bluestein post");
    }

    // Real-to-complex transform of the last axis (n reals -> n/2+1 bins per row).
    void r2c_axis(int64_t n, int sign, int64_t rows, const BufRef &src, const BufRef &dst)
    {
        const int64_t h = n / 2 + 1;
        if (n % 2 == 0 && n >= 2) {
            const int64_t L = n / 2;
            Lay in{src, 1, {{rows, L}}};
            Lay out{dst, 1, {{rows, h}}};
            if (L == 1 || (find_instance(L) && L <= kRowLeafMax)) {
                if (L == 1) {
                    // n = 2: X0 = x0 + x1, X1 = x0 - x1 via the post-process with Z = x0 + i x1
                    BufRef w = new_ws((size_t)rows * 8);
                    Lay wl{w, 1, {{rows, 1}}};
                    add_aux(AUX_CPLX_TRUNC, in, wl, 1, 1, TableRef(), 1.f, "r2c n=2 copy");
                    add_aux(AUX_R2C_POST, wl, out, 1, 2, real_roots(1, sign), 1.f, "r2c n=2 post");
                } else {
                    add_stockham(L, sign, in, out, IO_R2C_EVEN, L, h, 0, 1, "r2c even fused " + std::to_string(n));
                }
                return;
            }
            BufRef w = new_ws((size_t)rows * L * 8);
            Lay wl{w, 1, {{rows, L}}};
            axis(L, sign, in, wl, false);
            add_aux(AUX_R2C_POST, wl, out, L, h, real_roots(L, sign), 1.f, "r2c post");
            return;
        }
        if (find_instance(n) && n <= kRowLeafMax) {
            Lay in{src, 1, {{rows, n}}};
            Lay out{dst, 1, {{rows, h}}};
            add_stockham(n, sign, in, out, IO_R2C_FULL, n, h, 0, 1, "r2c full " + std::to_string(n));
            return;
        }
        BufRef w = new_ws((size_t)rows * n * 8);
        Lay rin{src, 1, {{rows, n}}};
        Lay wl{w, 1, {{rows, n}}};
        add_aux(AUX_REAL_TO_CPLX, rin, wl, n, n, TableRef(), 1.f, "r2c widen");
        axis(n, sign, wl, wl, true);
        add_aux(AUX_CPLX_TRUNC, wl, Lay{dst, 1, {{rows, h}}}, n, h, TableRef(), 1.f, "r2c keep half");
    }

    // Complex-to-real transform of the last axis (n/2+1 bins -> n reals per row).
    void c2r_axis(int64_t n, int sign, int64_t rows, const BufRef &src, const BufRef &dst)
    {
        const int64_t h = n / 2 + 1;
        if (n % 2 == 0 && n >= 4) {
            const int64_t L = n / 2;
            Lay in{src, 1, {{rows, h}}};
            Lay out{dst, 1, {{rows, L}}};
            if (find_instance(L) && L <= kRowLeafMax) {
                add_stockham(L, sign, in, out, IO_C2R_EVEN, h, L, 0, 1, "c2r even fused " + std::to_string(n));
                return;
            }
            BufRef w = new_ws((size_t)rows * L * 8);
            Lay wl{w, 1, {{rows, L}}};
            add_aux(AUX_C2R_PRE, in, wl, L, L, real_roots(L, sign), 1.f, "c2r pre");
            axis(L, sign, wl, out, true);
            return;
        }
        if (n > 1 && find_instance(n) && n <= kRowLeafMax) {
            Lay in{src, 1, {{rows, h}}};
            Lay out{dst, 1, {{rows, n}}};
            add_stockham(n, sign, in, out, IO_C2R_FULL, h, n, 0, 1, "c2r full " + std::to_string(n));
            return;
        }
        BufRef w = new_ws((size_t)rows * n * 8);
        Lay wl{w, 1, {{rows, n}}};
        add_aux(AUX_HERM_EXTEND, Lay{src, 1, {{rows, h}}}, wl, n, n, TableRef(), 1.f, "c2r hermitian extend");
        axis(n, sign, wl, wl, true);
        add_aux(AUX_CPLX_REAL, wl, Lay{dst, 1, {{rows, n}}}, n, n, TableRef(), 1.f, "c2r real part");
    }
};

}  // namespace

// ====================================================================== plan build

void plan_build(Plan &p)
{
    p.stages.clear();
    p.ws_offsets.clear();
    p.ws_bytes = 0;
    Builder b(p);
    b.sep = p.fftType == kTypeC2CSep;
    const int sign = p.sign();
    const int nd = (int)p.sizes.size();
    BufRef IN, OUT;
    IN.kind = BUF_IN;
    OUT.kind = BUF_OUT;
    const int64_t B = p.batch;

    if (p.istft.active) {
        // spectra X[c][k][t] (element k of frame (c, t) at c*n*T + k*T + t): inverse DFTs straight
        // from that layout (column tiles over t, no transpose) into frames F[c][t][j], then the
        // windowed overlap-add (reference: transpose + C2C + unfold-backward + envelope division,
        // core/fft/istft.cpp:91-104, core/fft/fftcore/fft_core_istft_any.cpp:292-380)
        const Plan::Istft &q = p.istft;
        BufRef fr = b.new_ws((size_t)q.channel * q.frames * q.nfft * 8);
        Lay in{IN, q.frames, {{q.frames, 1}, {q.channel, q.nfft * q.frames}}};
        Lay fl{fr, 1, {{q.frames, q.nfft}, {q.channel, q.frames * q.nfft}}};
        b.axis(q.nfft, +1, in, fl, false);
        b.add_aux(AUX_ISTFT_OLA, Lay{fr, 1, {{q.channel, q.frames * q.nfft}}}, Lay{OUT, 1, {{q.channel, q.out_len}}},
                  q.nfft, q.out_len, TableRef(), 1.f, "istft overlap-add");
        AuxArgs &x = p.stages.back().xa;
        x.hop = q.hop;
        x.frames = q.frames;
        x.start = q.nfft / 2;
        return;
    }
    if (nd == 1) {
        const int64_t N = p.sizes[0];
        if (p.fftType == kTypeC2C || p.fftType == kTypeC2CSep) {
            if (p.dimType == kDimVertical) {
                // (N, batch): axis 0 with element stride batch
                Lay in{IN, B, {{B, 1}}};
                Lay out{OUT, B, {{B, 1}}};
                b.axis(N, sign, in, out, false);
            } else {
                Lay in{IN, 1, {{B, N}}};
                Lay out{OUT, 1, {{B, N}}};
                b.axis(N, sign, in, out, false);
            }
        } else if (p.fftType == kTypeR2C) {
            b.r2c_axis(N, sign, B, IN, OUT);
        } else if (p.fftType == kTypeC2R) {
            b.c2r_axis(N, sign, B, IN, OUT);
        } else {
            throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "fft type not supported");
        }
        return;
    }

    // 2D / 3D: array [B][d0][d1]([d2]), last axis fastest.
    std::vector<int64_t> d = p.sizes;
    if (p.fftType == kTypeC2C || p.fftType == kTypeC2CSep) {
        // the last two axes of 32..128 each: one fused 2D tile launch (in -> out), then the first
        // axis of a 3D transform in place on out
        const Stockham2dInstance *t2 = (!b.sep && p.precision == PREC_FP32 && env_flag("SIP_FFT_FUSED2D", true))
                                           ? find_2d_instance(d[nd - 2], d[nd - 1]) : nullptr;
        if (t2) {
            int64_t slices = B;
            for (int i = 0; i < nd - 2; i++) slices *= d[i];
            b.add_tile2d(t2, sign, IN, OUT, slices,
                         "fused 2D " + std::to_string(d[nd - 2]) + "x" + std::to_string(d[nd - 1]));
            for (int ax = nd - 3; ax >= 0; ax--) {
                int64_t inner = 1, outer = B;
                for (int i = ax + 1; i < nd; i++) inner *= d[i];
                for (int i = 0; i < ax; i++) outer *= d[i];
                Lay l{OUT, inner, {{inner, 1}, {outer, d[ax] * inner}}};
                b.axis(d[ax], sign, l, l, true);
            }
            return;
        }
        // last axis: rows in -> out, then the other axes in place on out
        int64_t rows = B;
        for (int i = 0; i < nd - 1; i++) rows *= d[i];
        const int64_t nl = d[nd - 1];
        b.axis(nl, sign, Lay{IN, 1, {{rows, nl}}}, Lay{OUT, 1, {{rows, nl}}}, false);
        for (int ax = nd - 2; ax >= 0; ax--) {
            int64_t inner = 1, outer = B;
            for (int i = ax + 1; i < nd; i++) inner *= d[i];
            for (int i = 0; i < ax; i++) outer *= d[i];
            Lay l{OUT, inner, {{inner, 1}, {outer, d[ax] * inner}}};
            b.axis(d[ax], sign, l, l, true);
        }
        return;
    }
    if (p.fftType == kTypeR2C) {
        const int64_t nl = d[nd - 1], h = nl / 2 + 1;
        int64_t rows = B;
        for (int i = 0; i < nd - 1; i++) rows *= d[i];
        b.r2c_axis(nl, sign, rows, IN, OUT);
        for (int ax = nd - 2; ax >= 0; ax--) {
            int64_t inner = h, outer = B;
            for (int i = ax + 1; i < nd - 1; i++) inner *= d[i];
            for (int i = 0; i < ax; i++) outer *= d[i];
            Lay l{OUT, inner, {{inner, 1}, {outer, d[ax] * inner}}};
            b.axis(d[ax], sign, l, l, true);
        }
        return;
    }
    if (p.fftType == kTypeC2R) {
        const int64_t nl = d[nd - 1], h = nl / 2 + 1;
        int64_t rows = B;
        for (int i = 0; i < nd - 1; i++) rows *= d[i];
        // complex axes first, from the input into a workspace copy (the input is not modified)
        BufRef w = b.new_ws((size_t)rows * h * 8);
        bool first = true;
        for (int ax = 0; ax <= nd - 2; ax++) {
            int64_t inner = h, outer = B;
            for (int i = ax + 1; i < nd - 1; i++) inner *= d[i];
            for (int i = 0; i < ax; i++) outer *= d[i];
            Lay lin{first ? IN : w, inner, {{inner, 1}, {outer, d[ax] * inner}}};
            Lay lout{w, inner, {{inner, 1}, {outer, d[ax] * inner}}};
            b.axis(d[ax], sign, lin, lout, !first);
            first = false;
        }
        b.c2r_axis(nl, sign, rows, w, OUT);
        return;
    }
    throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "fft type not supported");
}

void plan_build_gathered_c2c(Plan &p, int64_t n, const std::vector<int64_t> &in_rows,
                             const std::vector<int64_t> &out_rows, void *in_base, void *out_base)
{
    plan_build_gathered(p, kTypeC2C, n, in_rows, out_rows, in_base, out_base);
}

bool gathered_supported(int type, int64_t n)
{
    if (type == kTypeC2C) return n >= 1;
    if (type == kTypeR2C) {
        if (n % 2 == 0 && n >= 4) return find_instance(n / 2) && n / 2 <= kRowLeafMax;
        return n >= 3 && find_instance(n) && n <= kRowLeafMax;
    }
    return false;
}

void plan_build_gathered(Plan &p, int type, int64_t n, const std::vector<int64_t> &in_rows,
                         const std::vector<int64_t> &out_rows, void *in_base, void *out_base)
{
    if (in_rows.size() != out_rows.size() || in_rows.empty())
        throw SipError(AsdSip::ACL_ERROR_INVALID_PARAM, "gathered plan needs matching non-empty row lists");
    if (!gathered_supported(type, n))
        throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "gathered rows not supported for this type / length");
    p.stages.clear();
    p.ws_offsets.clear();
    p.ws_bytes = 0;
    p.fftType = type;
    p.sizes = {n};
    p.batch = (int64_t)in_rows.size();
    p.abs_ptrs = {in_base, out_base};
    Builder b(p);
    BufRef bi, bo;
    bi.kind = BUF_ABS;
    bi.ws_index = 0;
    bo.kind = BUF_ABS;
    bo.ws_index = 1;
    const int64_t rows = (int64_t)in_rows.size();
    Lay in{bi, 1, {Dim{rows, 0, 0}}};
    Lay out{bo, 1, {Dim{rows, 0, 1}}};
    if (type == kTypeC2C) {
        p.row_tables = {in_rows, out_rows};
        b.axis(n, p.sign(), in, out, false);
        return;
    }
    // R2C: even n reads each real row as n/2 complex values (row offsets given in floats)
    const int64_t h = n / 2 + 1;
    if (n % 2 == 0) {
        std::vector<int64_t> cin(in_rows.size());
        for (size_t i = 0; i < in_rows.size(); i++) {
            if (in_rows[i] % 2) throw SipError(AsdSip::ACL_ERROR_INVALID_PARAM, "R2C rows must start at even floats");
            cin[i] = in_rows[i] / 2;
        }
        p.row_tables = {cin, out_rows};
        b.add_stockham(n / 2, p.sign(), in, out, IO_R2C_EVEN, n / 2, h, 0, 1, "gathered r2c even " + std::to_string(n));
    } else {
        p.row_tables = {in_rows, out_rows};
        b.add_stockham(n, p.sign(), in, out, IO_R2C_FULL, n, h, 0, 1, "gathered r2c full " + std::to_string(n));
    }
}

Plan::~Plan()
{
    for (void *d : row_tables_dev)
        if (d) (void)hipFree(d);
}

// ====================================================================== device side

static const void *stage_kernel(const Stage &s)
{
    if (s.inst2d) return s.sign < 0 ? s.inst2d->fwd : s.inst2d->inv;
    if (s.rc) return s.inst->rc[s.sign < 0 ? 0 : 1];
    if (s.sa.split_lds && s.sa.col_direct) return s.inst->cols[(s.sa.tw4_ntot != 0 ? 2 : 0) + (s.sign < 0 ? 0 : 1)];
    if (s.sa.split_lds && s.sa.io_mode == IO_R2C_EVEN) return s.inst->real_rows[s.sign < 0 ? 0 : 1];
    if (s.sa.split_lds && s.sa.io_mode == IO_C2R_EVEN) return s.inst->real_rows[s.sign < 0 ? 2 : 3];
    if (s.sa.split_lds) return s.sign < 0 ? s.inst->rows_fwd : s.inst->rows_inv;
    return s.sign < 0 ? s.inst->fwd : s.inst->inv;
}

bool plan_launchable(const Plan &p)
{
    // gfx950: at most 1024 threads and 160 KiB of LDS (static + dynamic) per workgroup
    constexpr size_t kLdsPerWorkgroup = 160 * 1024;
    for (const Stage &s : p.stages) {
        if (s.kind != StageKind::Stockham) continue;
        hipFuncAttributes attr;
        if (hipFuncGetAttributes(&attr, stage_kernel(s)) != hipSuccess) {
            (void)hipGetLastError();  // the query's own failure, raised just above
            return false;
        }
        if ((int)(s.block.x * s.block.y * s.block.z) > attr.maxThreadsPerBlock) return false;
        if (s.lds + attr.sharedSizeBytes > kLdsPerWorkgroup) return false;
    }
    return true;
}

void plan_materialize_external(Plan &p, const std::vector<const int64_t *> &dev_tables)
{
    // row tables already on the device (caller-owned, e.g. uploaded asynchronously)
    p.row_tables_dev.clear();
    p.external_tables = dev_tables;
    plan_materialize(p);
}

// Streaming (nontemporal) HBM accesses per launch. Stores: the last pass of a stage whose output no
// later stage of the plan reads (single-plane rows, the matrix-core rows, and column tiles under
// the conditions below). Loads: the first pass
// of a stage that reads 512 MiB or more (single-plane rows, column tiles, rows-in / columns-out),
// far beyond the 256 MB last-level cache. Measured on MI355X (one box, in-process A/B against plain
// accesses): nontemporal stores 4096 x 16384 C2C 192 -> 184 us and the batching workload 9.2 ->
// 9.0 ms, neutral elsewhere; nontemporal loads on top 186 -> 177 us, 1024 x 65536 186 -> 180 us,
// 2D 1024^2 x 128 855 -> 837 us, 3D 128^3 x 64 801 -> 786 us, 32768 x 4096 540 -> 532 us, but
// slower on inputs that the last-level cache still holds (2000 x 8192 45 -> 56 us, R2C 8192 x 4096
// 48 -> 52 us, four-step 43200 x 1056 341 -> 359 us, batching workload 8.9 -> 9.3 ms), hence the
// size threshold.
void plan_mark_streaming(Plan &p)
{
    const bool on_out = env_flag("SIP_FFT_NT_OUT", true), on_in = env_flag("SIP_FFT_NT_IN", true);
    constexpr int64_t kNtInBytes = int64_t(512) << 20;
    for (size_t i = 0; i < p.stages.size(); i++) {
        Stage &s = p.stages[i];
        if (s.kind != StageKind::Stockham) continue;
        bool read_later = false;
        for (size_t j = i + 1; j < p.stages.size(); j++) {
            const BufRef &in = p.stages[j].in;
            if (in.kind == s.out.kind && (in.kind == BUF_IN || in.kind == BUF_OUT || in.ws_index == s.out.ws_index))
                read_later = true;
        }
        // Nontemporal accesses only where every wave covers whole 128-byte lines: a run of T
        // consecutive complex values per radix slot, at multiples of 16 elements (2000 x 8192 with
        // T = 100 wrote 7 % more bytes with nontemporal stores); column tiles 16 transforms wide.
        const StockhamInstance *in = s.inst;
        const bool rows_aligned_in = in && in->T % 16 == 0 && (in->L / in->radix[0]) % 16 == 0;
        const bool rows_aligned_out = in && in->T % 16 == 0 && (in->L / in->radix[in->npass - 1]) % 16 == 0;
        const bool cols_in = s.sa.col_direct && s.sa.split_lds;
        const bool aligned_in = cols_in ? s.sa.F % 16 == 0 : rows_aligned_in;
        // column outputs (column tiles, rows-in / columns-out): runs of F adjacent transforms at a
        // stride of tout.es elements; only for outputs of 512 MiB or more of plain (not gathered)
        // plans: measured on MI355X, 65536 x 2048 823 -> 774 us, 2D 1024^2 x 128 847 -> 808 us,
        // 512^2 x 256 394 -> 352 us, 3D 128^3 x 64 778 -> 758 us, but the batching form's graph
        // 8.9 -> 9.5 ms with them on its groups' final stores
        const bool cols_out = s.sa.tout.unit_f != 0;
        const bool gathered = !p.row_tables.empty() || !p.external_tables.empty();
        const int64_t out_bytes = s.sa.nfft * (int64_t)s.sa.out_len * 8;
        const bool aligned_out = cols_out ? (s.sa.F % 16 == 0 && s.sa.tout.es % 16 == 0 && s.sa.tw4_ntot == 0 &&
                                             !gathered && out_bytes >= kNtInBytes)
                                          : rows_aligned_out;
        s.sa.nt_out = (on_out && !read_later && aligned_out) ? 1 : 0;
        const int64_t in_bytes = s.sa.nfft * (int64_t)s.sa.in_len * 8;
        s.sa.nt_in = (on_in && s.sa.io_mode == IO_C2C && in_bytes >= kNtInBytes && aligned_in) ? 1 : 0;
    }
}

void plan_materialize(Plan &p)
{
    SIP_HIP_CHECK(hipGetDevice(&p.device));
    plan_mark_streaming(p);
    if (!p.external_tables.empty()) {
        for (Stage &s : p.stages) {
            const int64_t *ti = s.in_tab >= 0 ? p.external_tables[(size_t)s.in_tab] : nullptr;
            const int64_t *to = s.out_tab >= 0 ? p.external_tables[(size_t)s.out_tab] : nullptr;
            if (s.kind == StageKind::Stockham) {
                if (ti) s.sa.tin.tab = ti;
                if (to) s.sa.tout.tab = to;
            } else {
                if (ti) s.xa.rin.tab = ti;
                if (to) s.xa.rout.tab = to;
            }
        }
    } else if (p.row_tables_dev.size() != p.row_tables.size()) {
        for (void *d : p.row_tables_dev)
            if (d) (void)hipFree(d);
        p.row_tables_dev.assign(p.row_tables.size(), nullptr);
        for (size_t i = 0; i < p.row_tables.size(); i++) {
            const size_t bytes = p.row_tables[i].size() * sizeof(int64_t);
            SIP_HIP_CHECK(hipMalloc(&p.row_tables_dev[i], std::max<size_t>(bytes, 8)));
            SIP_HIP_CHECK(hipMemcpy(p.row_tables_dev[i], p.row_tables[i].data(), bytes, hipMemcpyHostToDevice));
        }
    }
    for (Stage &s : p.stages) {
        if (s.kind == StageKind::Stockham) {
            if (p.external_tables.empty()) {
                if (s.in_tab >= 0) s.sa.tin.tab = static_cast<const int64_t *>(p.row_tables_dev[(size_t)s.in_tab]);
                if (s.out_tab >= 0) s.sa.tout.tab = static_cast<const int64_t *>(p.row_tables_dev[(size_t)s.out_tab]);
            }
            s.tw.dev = coeff_get(s.tw);
            if (s.rtw.used()) s.rtw.dev = coeff_get(s.rtw);
            if (s.tw4lo.used()) {
                s.tw4lo.dev = coeff_get(s.tw4lo);
                s.tw4hi.dev = coeff_get(s.tw4hi);
            }
            s.sa.tw = static_cast<const float2 *>(s.tw.dev->ptr);
            if (s.tw2.used()) {
                s.tw2.dev = coeff_get(s.tw2);
                s.sa.tw2 = static_cast<const float2 *>(s.tw2.dev->ptr);
            }
            if (s.mfw.used()) {
                s.mfw.dev = coeff_get(s.mfw);
                s.sa.mfma_w = static_cast<const unsigned short *>(s.mfw.dev->ptr);
            }
            s.sa.rtw = s.rtw.dev ? static_cast<const float2 *>(s.rtw.dev->ptr) : nullptr;
            s.sa.tw4_lo = s.tw4lo.dev ? static_cast<const float2 *>(s.tw4lo.dev->ptr) : nullptr;
            s.sa.tw4_hi = s.tw4hi.dev ? static_cast<const float2 *>(s.tw4hi.dev->ptr) : nullptr;
            const void *fn = stage_kernel(s);
            if (s.lds > 64 * 1024)
                SIP_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s.lds));
        } else {
            if (p.external_tables.empty()) {
                if (s.in_tab >= 0) s.xa.rin.tab = static_cast<const int64_t *>(p.row_tables_dev[(size_t)s.in_tab]);
                if (s.out_tab >= 0) s.xa.rout.tab = static_cast<const int64_t *>(p.row_tables_dev[(size_t)s.out_tab]);
            }
            if (s.tab0.used()) {
                s.tab0.dev = coeff_get(s.tab0);
                s.xa.tab0 = static_cast<const float2 *>(s.tab0.dev->ptr);
            }
        }
    }
}

static void *resolve(const Plan &p, const BufRef &b, const void *in, void *out)
{
    if (b.kind == BUF_IN) return const_cast<void *>(in);
    if (b.kind == BUF_OUT) return out;
    if (b.kind == BUF_ABS) return p.abs_ptrs[(size_t)b.ws_index];
    return static_cast<char *>(p.workspace) + p.ws_offsets[(size_t)b.ws_index];
}

void plan_execute(Plan &p, const void *in, void *out) { plan_execute_sep(p, in, nullptr, out, nullptr); }

void plan_execute_istft(Plan &p, const void *in, const float *window, void *out)
{
    for (Stage &s : p.stages)
        if (s.kind == StageKind::Aux && s.xa.mode == AUX_ISTFT_OLA) s.xa.win = window;
    plan_execute(p, in, out);
}

void plan_execute_sep(Plan &p, const void *in, const void *in_im, void *out, void *out_im)
{
    // imaginary planes of the user buffers (split-plane plans only)
    auto im_of = [&](const BufRef &b) -> float * {
        if (b.kind == BUF_IN) return static_cast<float *>(const_cast<void *>(in_im));
        if (b.kind == BUF_OUT) return static_cast<float *>(out_im);
        return nullptr;
    };
    for (Stage &s : p.stages) {
        void *src = resolve(p, s.in, in, out);
        void *dst = resolve(p, s.out, in, out);
        if (s.kind == StageKind::Stockham) {
            StockhamArgs a = s.sa;
            a.in = src;
            a.out = dst;
            a.in_im = im_of(s.in);
            a.out_im = im_of(s.out);
            if ((a.in_sep && !a.in_im) || (a.out_sep && !a.out_im))
                throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "split-plane plan executed without imaginary planes");
            if (a.tin.vec2 && (reinterpret_cast<uintptr_t>(src) & 15)) a.tin.vec2 = 0;
            if (a.tout.vec2 && (reinterpret_cast<uintptr_t>(dst) & 15)) a.tout.vec2 = 0;
            void *args[] = {&a};
            const void *fn = stage_kernel(s);
            SIP_HIP_CHECK(hipLaunchKernel(fn, s.grid, s.block, args, s.lds, p.stream));
        } else {
            AuxArgs a = s.xa;
            a.in = src;
            a.out = dst;
            a.in_im = im_of(s.in);
            a.out_im = im_of(s.out);
            if ((a.in_sep && !a.in_im) || (a.out_sep && !a.out_im))
                throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "split-plane plan executed without imaginary planes");
            SIP_HIP_CHECK(aux_launch(a, p.stream));
        }
    }
}

// ====================================================================== description

static const char *buf_name(const BufRef &b)
{
    return b.kind == BUF_IN ? "in" : b.kind == BUF_OUT ? "out" : "ws";
}

std::string plan_describe(const Plan &p)
{
    std::ostringstream os;
    os << "{\"type\":" << p.fftType << ",\"direction\":" << p.direction << ",\"sizes\":[";
    for (size_t i = 0; i < p.sizes.size(); i++) os << (i ? "," : "") << p.sizes[i];
    os << "],\"batch\":" << p.batch << ",\"dimType\":" << p.dimType << ",\"workspace\":" << p.ws_bytes;
    os << ",\"stages\":[";
    for (size_t i = 0; i < p.stages.size(); i++) {
        const Stage &s = p.stages[i];
        if (i) os << ",";
        os << "{\"kind\":\"" << (s.kind == StageKind::Stockham ? "stockham" : "aux") << "\",\"in\":\""
           << buf_name(s.in) << "\",\"out\":\"" << buf_name(s.out) << "\",\"note\":\"" << s.note << "\"";
        if (s.kind == StageKind::Stockham) {
            os << ",\"L\":" << s.inst->L << ",\"T\":" << s.inst->T << ",\"F\":" << s.sa.F << ",\"radices\":[";
            for (int r = 0; r < s.inst->npass; r++) os << (r ? "," : "") << s.inst->radix[r];
            os << "],\"nfft\":" << s.sa.nfft << ",\"io\":" << s.sa.io_mode << ",\"lds\":" << s.lds
               << ",\"grid\":" << s.grid.x << ",\"block\":" << s.block.x << ",\"unit_f_in\":" << s.sa.tin.unit_f
               << ",\"unit_f_out\":" << s.sa.tout.unit_f << ",\"tw4\":" << s.sa.tw4_ntot
               << ",\"direct_in\":" << s.sa.direct_in << ",\"direct_out\":" << s.sa.direct_out
               << ",\"split_lds\":" << s.sa.split_lds << ",\"col_direct\":" << s.sa.col_direct << ",\"rc\":" << (s.rc ? 1 : 0)
               << ",\"in_sep\":" << s.sa.in_sep << ",\"out_sep\":" << s.sa.out_sep << ",\"mfma\":" << (s.mfw.used() ? 1 : 0)
               << ",\"xcd_remap\":" << s.sa.xcd_remap << ",\"nt_out\":" << s.sa.nt_out << ",\"nt_in\":" << s.sa.nt_in;
            if (s.inst2d) os << ",\"tile2d\":[" << s.inst2d->X << "," << s.inst2d->Y << "]";
        } else {
            os << ",\"mode\":" << s.xa.mode << ",\"rows\":" << s.xa.rows << ",\"n\":" << s.xa.n
               << ",\"out_len\":" << s.xa.out_len;
        }
        os << "}";
    }
    os << "]}";
    return os.str();
}

}  // namespace sip
