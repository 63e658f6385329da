// Restatement of the reference's plan rules: which core and step list the reference would build
// for a given asdFft* call. Test infrastructure only (built into oracle/libref_rules.so by
// oracle/Makefile and called from the CPU tests); nothing here is part of libsip_fft.so or drives
// its kernels.
//
// Sources: ChooseRadix / Support / deDuplicates / orderedFactorize (core/fft/fft_api.cpp:66-110,
// core/fft/utils/include/utils/fft_common_func.h:39-62), getC2CCore / getR2CCore / getC2RCore /
// getCore (fft_api.cpp:169-245, 383-432), SelectCore::configs + FindConfig
// (core/fft/fftcore/select_core.cpp:155-200), init2DSteps / init3DSteps / Fft2dSupportVertical /
// Fft2dSupportFusing / addFFTC2CFirstTwoDimsStep (fft_api.cpp:112-167, 527-697), FFTCoreB::InitRadix
// (core/fft/fftcore/fft_core_b.cpp:126-168), FindTwoRadix / InitMixRadixShort
// (core/fft/fftcore/fft_core_common_func.cpp:113-144, 767-826; RADIX_ADDR_START = 128 from
// fft_core_common_func.h:43), the R2C/C2R half length (fft_r2c_core.h:20-24) and the ISTFT step
// list (core/fft/istft.cpp:91-104).
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace refrules {
namespace {

// enum values of the reference's include/fft_api.h:26-44
constexpr int kTypeC2C = 0x10, kTypeC2R = 0x11, kTypeR2C = 0x12, kTypeC2CSep = 0x30;
constexpr int kDirInverse = 0x11;
constexpr int kDimVertical = 0x11;

constexpr int kRadix2 = 2, kRadixMix = -2, kRadixAny = -1;

std::vector<int64_t> ordered_factorize(int64_t size)
{
    std::vector<int64_t> f;
    if (size <= 0) return f;
    int64_t bound = (int64_t)std::sqrt((double)size);
    for (int64_t factor = 2; factor <= bound;) {
        if (size % factor == 0) {
            f.push_back(factor);
            size /= factor;
            bound = (int64_t)std::sqrt((double)size);
        } else {
            factor++;
        }
    }
    if (size != 1) f.push_back(size);
    return f;
}

std::vector<int64_t> dedup(const std::vector<int64_t> &v)
{
    std::vector<int64_t> u;
    for (int64_t x : v)
        if (u.empty() || u.back() != x) u.push_back(x);
    return u;
}

bool supports(const std::vector<int64_t> &u, const std::vector<int64_t> &set)
{
    for (int64_t f : u)
        if (std::find(set.begin(), set.end(), f) == set.end()) return false;
    return true;
}

int choose_radix(int type, int64_t n)
{
    static const std::vector<int64_t> r2 = {2};
    static const std::vector<int64_t> rmix = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47};
    auto u = dedup(ordered_factorize(n));
    if (type == kTypeC2C && supports(u, r2)) return kRadix2;
    if (supports(u, rmix)) return kRadixMix;
    return kRadixAny;
}

// SelectCore::FindConfig (select_core.cpp:155-200): batch range -> list of (sizes, core)
std::string find_config(int64_t batch, int64_t n)
{
    struct Entry {
        int64_t lo, hi;
        std::vector<std::pair<std::vector<int64_t>, const char *>> cores;
    };
    static const std::vector<Entry> configs = {
        {8, 96, {{{256, 512, 1024}, "Dft"}, {{2048, 4096, 8192, 16384}, "FftB"},
                 {{32768, 65536, 131072, 262144, 524288, 1048576, 2097152, 4194304}, "FftN"}}},
        {96, 1024, {{{256, 512}, "Dft"}, {{1024, 2048, 4096, 8192, 16384}, "FftB"}}},
        {1024, 65537, {{{256}, "Dft"}, {{512, 1024, 2048, 4096, 8192, 16384}, "FftB"}}},
    };
    const uint32_t b = (uint32_t)batch, v = (uint32_t)n;
    for (const Entry &e : configs) {
        if (b >= e.lo && b < e.hi) {
            for (auto &c : e.cores)
                if (std::find(c.first.begin(), c.first.end(), (int64_t)v) != c.first.end()) return c.second;
            return "";
        }
    }
    return "";
}

std::vector<int64_t> find_two_radix(std::vector<int64_t> &factors, int64_t n)
{
    if (n == 0) n = 1;
    int64_t n1 = factors[0];
    int64_t minN1 = n1 == 0 ? 1 : n1;
    float minRatio = (float)std::max(n / std::pow(n1, 2), std::pow(n1, 2) / n);
    const int64_t fLen = (int64_t)factors.size();
    for (int64_t i = 0; i < (int64_t(1) << fLen); i++) {
        n1 = 1;
        for (int64_t j = 0; j < fLen; j++)
            if ((uint64_t(i) & (uint64_t(1) << j)) != 0) n1 *= factors[(size_t)j];
        const float ratio = (float)std::max(n / std::pow(n1, 2), std::pow(n1, 2) / n);
        if (ratio < minRatio) {
            minN1 = n1;
            minRatio = ratio;
        }
    }
    std::vector<int64_t> r = {minN1, n / minN1};
    std::sort(r.begin(), r.end());
    return r;
}

std::vector<int64_t> mix_radix_short(int64_t fftN)
{
    if (fftN <= 128) return {fftN};
    std::vector<int64_t> factors;
    int64_t n = fftN;
    for (int64_t p = 2; p <= n; p++) {
        while (n % p == 0) {
            factors.push_back(p);
            n /= p;
        }
        if (n == 1) break;
    }
    std::vector<int64_t> radix = find_two_radix(factors, fftN);
    if (radix[1] > 128) {
        for (size_t i = 0; i < factors.size(); i++) {
            const int64_t t = factors[i];
            if (t == 0) continue;
            factors.erase(factors.begin() + (std::ptrdiff_t)i);
            std::vector<int64_t> nr = find_two_radix(factors, fftN / t);
            factors.insert(factors.begin() + (std::ptrdiff_t)i, t);
            if (nr[1] <= 128) {
                nr.insert(nr.begin(), t);
                radix = nr;
                break;
            }
        }
    }
    return radix;
}

std::string radix_str(const std::vector<int64_t> &r)
{
    std::ostringstream os;
    os << "[";
    for (size_t i = 0; i < r.size(); i++) os << (i ? "," : "") << r[i];
    os << "]";
    return os.str();
}

std::vector<int64_t> fftb_radix(int64_t n)
{
    switch (n) {
        case 256: return {16, 16};
        case 512: return {16, 32};
        case 1024: return {16, 64};
        case 2048: return {32, 64};
        case 4096: case 8192: case 16384: return {64, 64};
        case 32768: return {32, 32, 32};
        case 65536: return {32, 32, 64};
        case 131072: return {32, 64, 64};
        case 262144: return {64, 64, 64};
        default: return {};
    }
}

std::string core_step(int type, int64_t n, int64_t stride, int64_t batch, bool forward)
{
    std::ostringstream os;
    std::string core;
    if (type == kTypeC2CSep) {
        // split planes: FFT-B for forward 256 / 512, one-GEMM DFT otherwise (fft_api.cpp:394-399)
        core = ((n == 256 || n == 512) && forward) ? "FftBSep" : "DftSep";
    } else if (type == kTypeC2C && stride > 1) {
        core = "FftStride";
    } else {
        if (type == kTypeC2C) core = find_config(batch, n);
        if (core.empty()) {
            const int radix = choose_radix(type, n);
            if (type == kTypeC2C) {
                if (n <= 256) core = "Dft";
                else if (radix == kRadix2) core = n < 32768 ? "FftB" : "FftN";
                else if (radix == kRadixMix) core = "FftMix";
            } else if (type == kTypeR2C) {
                if (n <= 1024) core = "DftR2C";
                else if (radix == kRadixMix) core = "FftR2C";
            } else if (type == kTypeC2R) {
                if (n <= 1024) core = "DftC2R";
                else if (radix == kRadixMix) core = "FftC2R";
            }
            if (core.empty()) core = "Any";
        }
    }
    os << core << "(N=" << n << ",batch=" << batch;
    if (stride > 1) os << ",stride=" << stride;
    if (core == "FftB") os << ",radix=" << radix_str(fftb_radix(n));
    if (core == "FftMix" && n < 32768) os << ",radix=" << radix_str(mix_radix_short(n));
    if ((core == "FftR2C" || core == "FftC2R")) {
        const int64_t c2c = (n % 2) ? n : n / 2;
        if (c2c < 32768) os << ",c2c=" << c2c << ",radix=" << radix_str(mix_radix_short(c2c));
    }
    os << ")";
    return os.str();
}

std::string transpose_step(int a0, int a1, const std::vector<int64_t> &dims)
{
    std::ostringstream os;
    os << "Transpose(" << a0 << "," << a1 << ",dims=" << radix_str(dims) << ")";
    return os.str();
}

bool support_vertical(int64_t batch, int64_t x, int64_t y, int type)
{
    if (batch != 1) return false;
    const int64_t h = y / 2 + 1;
    if (type == kTypeC2R || type == kTypeR2C) {
        if (h % 128 != 0) return false;
    } else if (type == kTypeC2C) {
        if (y % 128 != 0) return false;
    }
    if (!supports(dedup(ordered_factorize(x)), {2})) return false;
    return x >= 256 && x <= 65536;
}

bool support_fusing(int64_t x, int64_t y, int rx, int ry, int type)
{
    return type == kTypeC2C && rx == kRadix2 && ry == kRadix2 && x >= 32 && y >= 32 && x <= 128 && y <= 128;
}

}  // namespace

std::vector<std::string> reference_istft_steps(int64_t channel, int64_t nfft, int64_t frames, int64_t hop)
{
    // core/fft/istft.cpp:91-104: transpose (channel, n_fft, frames) -> (channel, frames, n_fft),
    // the inverse C2C of n_fft over channel * frames rows, then the ISTFT overlap-add core
    std::vector<std::string> s;
    s.push_back(transpose_step(1, 2, {channel, nfft, frames}));
    s.push_back(core_step(kTypeC2C, nfft, 1, channel * frames, false));
    s.push_back("IstftAny(n_fft=" + std::to_string(nfft) + ",hop=" + std::to_string(hop) +
                ",frames=" + std::to_string(frames) + ",channel=" + std::to_string(channel) + ")");
    return s;
}

std::vector<std::string> reference_steps(int type, int direction, const std::vector<int64_t> &sizes, int64_t batch,
                                         int dimType)
{
    const bool fwd = direction != kDirInverse;
    std::vector<std::string> s;
    if (sizes.size() == 1) {
        const int64_t stride = (dimType == kDimVertical) ? batch : 1;
        s.push_back(core_step(type, sizes[0], stride, batch, fwd));
        return s;
    }
    if (sizes.size() == 2 && type == kTypeC2CSep) return s;  // init2DSteps adds no step for C2C_SEP
    if (sizes.size() == 3 && type == kTypeC2CSep) {
        // one fused 3D split-plane DFT step (fft_api.cpp:451-458, 657-660)
        s.push_back("DddSep(X=" + std::to_string(sizes[0]) + ",Y=" + std::to_string(sizes[1]) + ",Z=" +
                    std::to_string(sizes[2]) + ",batch=" + std::to_string(batch) + ")");
        return s;
    }
    if (sizes.size() == 2) {
        const int64_t X = sizes[0], Y = sizes[1];
        const int rx = choose_radix(type, X), ry = choose_radix(type, Y);
        const int64_t h = Y / 2 + 1;
        if (type == kTypeC2R) {
            if (support_vertical(batch, X, Y, type)) {
                s.push_back(core_step(kTypeC2C, X, h, batch, fwd));
                s.push_back(core_step(kTypeC2R, Y, 1, batch * X, fwd));
            } else {
                s.push_back(transpose_step(-1, -2, {batch, X, h}));
                s.push_back(core_step(kTypeC2C, X, 1, batch * h, fwd));
                s.push_back(transpose_step(-1, -2, {batch, h, X}));
                s.push_back(core_step(kTypeC2R, Y, 1, batch * X, fwd));
            }
        } else if (type == kTypeR2C) {
            if (support_vertical(batch, X, Y, type)) {
                s.push_back(core_step(kTypeR2C, Y, 1, batch * X, fwd));
                s.push_back(core_step(kTypeC2C, X, h, batch, fwd));
            } else {
                s.push_back(core_step(kTypeR2C, Y, 1, batch * X, fwd));
                s.push_back(transpose_step(-1, -2, {batch, X, h}));
                s.push_back(core_step(kTypeC2C, X, 1, batch * h, fwd));
                s.push_back(transpose_step(-1, -2, {batch, h, X}));
            }
        } else if (type == kTypeC2C) {
            if (support_fusing(X, Y, rx, ry, type)) {
                s.push_back("Dd(X=" + std::to_string(X) + ",Y=" + std::to_string(Y) + ",batch=" + std::to_string(batch) + ")");
            } else if (support_vertical(batch, X, Y, type)) {
                s.push_back(core_step(kTypeC2C, Y, 1, batch * X, fwd));
                s.push_back(core_step(kTypeC2C, X, Y, batch, fwd));
            } else {
                s.push_back(core_step(kTypeC2C, Y, 1, batch * X, fwd));
                s.push_back(transpose_step(-1, -2, {batch, X, Y}));
                s.push_back(core_step(kTypeC2C, X, 1, batch * Y, fwd));
                s.push_back(transpose_step(-1, -2, {batch, Y, X}));
            }
        }
        return s;
    }
    // 3D
    const int64_t X = sizes[0], Y = sizes[1], Z = sizes[2];
    auto first_two = [&](int64_t zlen) {
        bool strideTag = false;
        if (support_vertical(batch, X, Y * zlen, kTypeC2C)) {
            strideTag = true;
            s.push_back(core_step(kTypeC2C, X, Y * zlen, batch, fwd));
        } else {
            s.push_back(transpose_step(-1, -3, {batch, X, Y, zlen}));
            s.push_back(core_step(kTypeC2C, X, 1, batch * Y * zlen, fwd));
        }
        if (X == 1 && support_vertical(batch, Y, zlen, kTypeC2C)) {
            s.push_back(core_step(kTypeC2C, Y, zlen, batch * X, fwd));
        } else if (strideTag) {
            s.push_back(transpose_step(-1, -2, {batch, X, Y, zlen}));
            s.push_back(core_step(kTypeC2C, Y, 1, batch * X * zlen, fwd));
            s.push_back(transpose_step(-1, -2, {batch, X, zlen, Y}));
        } else {
            s.push_back(transpose_step(-1, -2, {batch, zlen, Y, X}));
            s.push_back(core_step(kTypeC2C, Y, 1, batch * X * zlen, fwd));
            s.push_back(transpose_step(-1, -2, {batch, zlen, X, Y}));
            s.push_back(transpose_step(-1, -3, {batch, zlen, Y, X}));
        }
    };
    if (type == kTypeC2R) {
        first_two(Z / 2 + 1);
        s.push_back(core_step(kTypeC2R, Z, 1, batch * X * Y, fwd));
    } else if (type == kTypeR2C) {
        s.push_back(core_step(kTypeR2C, Z, 1, batch * X * Y, fwd));
        first_two(Z / 2 + 1);
    } else if (type == kTypeC2C) {
        const int ry = choose_radix(type, Y), rz = choose_radix(type, Z);
        if (support_fusing(Y, Z, ry, rz, type)) {
            s.push_back("Dd(X=" + std::to_string(Y) + ",Y=" + std::to_string(Z) + ",batch=" + std::to_string(batch * X) + ")");
            if (support_vertical(batch, X, Y * Z, type)) {
                s.push_back(core_step(kTypeC2C, X, Y * Z, batch, fwd));
            } else {
                s.push_back(transpose_step(-1, -3, {batch, X, Y, Z}));
                s.push_back(core_step(kTypeC2C, X, 1, batch * Y * Z, fwd));
                s.push_back(transpose_step(-1, -3, {batch, Z, Y, X}));
            }
        } else {
            first_two(Z);
            s.push_back(core_step(kTypeC2C, Z, 1, batch * X * Y, fwd));
        }
    }
    return s;
}

}  // namespace refrules

// C entry points for the tests: the step list as a JSON array of strings. Returns 0, or 1 when
// cap is too small (needed holds the size either way).
static int write_json(const std::vector<std::string> &steps, char *buf, size_t cap, size_t *needed)
{
    std::string j = "[";
    for (size_t i = 0; i < steps.size(); i++) j += (i ? ",\"" : "\"") + steps[i] + "\"";
    j += "]";
    if (needed) *needed = j.size() + 1;
    if (!buf || cap < j.size() + 1) return 1;
    std::memcpy(buf, j.c_str(), j.size() + 1);
    return 0;
}

extern "C" int ref_steps_json(int ndim, const int64_t *dims, int type, int direction, int64_t batch, int dim_type,
                              char *buf, size_t cap, size_t *needed)
{
    std::vector<int64_t> sizes(dims, dims + ndim);
    return write_json(refrules::reference_steps(type, direction, sizes, batch, dim_type), buf, cap, needed);
}

extern "C" int ref_istft_steps_json(int64_t channel, int64_t nfft, int64_t frames, int64_t hop, char *buf, size_t cap,
                                    size_t *needed)
{
    return write_json(refrules::reference_istft_steps(channel, nfft, frames, hop), buf, cap, needed);
}
