// Host side of the persistent multi-length Stockham kernel (kernels/persist.hip): radix sequence and
// thread shape per length group, the tile list (largest groups first, as given), device
// descriptors and the launch. No counterpart in the reference's host code; the reference's
// bulk-per-core loop (ops/fft/common/include/kernel/fft_all_common.cceh:102-119) is the nearest.
#include "persist.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <sstream>

namespace sip {

namespace {

bool persist_radix_ok(int r)
{
    for (int x : kPersistRadices)
        if (x == r) return true;
    return false;
}

bool smooth235(int64_t n)
{
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

}  // namespace

bool persist_shape(int64_t n, const std::vector<int> &radices, int &T, int &F)
{
    if (n < 2 || n > kPersistMaxL || radices.empty() || (int)radices.size() > kPersistMaxPass) return false;
    int64_t prod = 1, tmin = 1;
    for (int r : radices) {
        if (!persist_radix_ok(r)) return false;
        prod *= r;
        const int64_t nbf = n / r;
        const int64_t mx = persist_max_nb(r);
        tmin = std::max(tmin, (nbf + mx - 1) / mx);
    }
    if (prod != n || tmin > kPersistThreads) return false;
    // the most threads in use per workgroup (F * T), then the fewest threads per transform
    int bestT = 0, bestUse = -1;
    for (int t = (int)tmin; t <= kPersistThreads; t++) {
        const int f = kPersistThreads / t;
        if (f * t > bestUse) {
            bestUse = f * t;
            bestT = t;
        }
    }
    T = bestT;
    F = kPersistThreads / bestT;
    return true;
}

namespace {

// Radix sequences over kPersistRadices with product n, fewest passes first; each multiset in
// non-increasing order and then reversed.
std::vector<std::vector<int>> persist_sequences(int64_t n)
{
    std::vector<std::vector<int>> sets, out;
    std::vector<int> cur;
    std::function<void(int64_t, int)> rec = [&](int64_t m, int maxr) {
        if (m == 1) {
            sets.push_back(cur);
            return;
        }
        if ((int)cur.size() >= kPersistMaxPass) return;
        for (int i = (int)(sizeof(kPersistRadices) / sizeof(int)) - 1; i >= 0; i--) {
            const int r = kPersistRadices[i];
            if (r > maxr || m % r) continue;
            cur.push_back(r);
            rec(m / r, r);
            cur.pop_back();
        }
    };
    rec(n, 1 << 30);
    std::stable_sort(sets.begin(), sets.end(), [](const std::vector<int> &a, const std::vector<int> &b) {
        return a.size() < b.size();
    });
    for (const auto &sq : sets) {
        out.push_back(sq);
        std::vector<int> rev(sq.rbegin(), sq.rend());
        if (rev != sq) out.push_back(rev);
    }
    return out;
}

bool feasible(int64_t n, const std::vector<int> &r)
{
    int T, F;
    return persist_shape(n, r, T, F);
}

}  // namespace

std::vector<int> persist_default_radices(int64_t n)
{
    if (n < 2 || n > kPersistMaxL || !smooth235(n)) return {};
    // the leaf kernel's sequence when the kernel can hold it, else the first that fits
    if (const StockhamInstance *in = find_instance(n)) {
        std::vector<int> r(in->radix, in->radix + in->npass);
        if (feasible(n, r)) return r;
    }
    for (const auto &sq : persist_sequences(n))
        if (feasible(n, sq)) return sq;
    return {};
}

std::vector<std::vector<int>> persist_candidate_radices(int64_t n, size_t max_count)
{
    std::vector<std::vector<int>> out;
    const std::vector<int> def = persist_default_radices(n);
    if (def.empty()) return out;
    out.push_back(def);
    for (const auto &sq : persist_sequences(n)) {
        if (out.size() >= max_count) break;
        if (std::find(out.begin(), out.end(), sq) == out.end() && feasible(n, sq)) out.push_back(sq);
    }
    return out;
}

PersistLaunch::PersistLaunch(const std::vector<PersistSpec> &specs, const void *in_base, void *out_base, int sign)
    : sign_(sign)
{
    if (specs.empty()) throw SipError(AsdSip::ACL_ERROR_INVALID_PARAM, "persistent launch without groups");
    int64_t ntiles = 0, nrows = 0;
    for (const PersistSpec &sp : specs) {
        std::vector<int> rad = sp.radices.empty() ? persist_default_radices(sp.n) : sp.radices;
        int T = 0, F = 0;
        if (rad.empty() || !persist_shape(sp.n, rad, T, F) || sp.in_rows.size() != sp.out_rows.size())
            throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT,
                           "length " + std::to_string(sp.n) + " has no persistent-kernel shape");
        PersistGroup g;
        std::memset(&g, 0, sizeof(g));
        g.L = (int32_t)sp.n;
        g.T = T;
        g.F = F;
        const int64_t maxidx = sp.n - 1;
        int lp = (int)(maxidx + (maxidx >> 4)) + 1;
        if ((lp & 1) == 0) lp++;
        g.LP = lp;
        g.npass = (int32_t)rad.size();
        int64_t ns = 1, off = 0;
        for (size_t p = 0; p < rad.size(); p++) {
            g.radix[p] = rad[p];
            g.ns[p] = (int32_t)ns;
            g.nbf[p] = (int32_t)(sp.n / rad[p]);
            g.tw_off[p] = (int32_t)off;
            g.div_ns[p] = make_fastdiv((uint32_t)ns);
            off += ns * pass_twiddle_rows(rad[p]);
            ns *= rad[p];
        }
        g.div_T = make_fastdiv((uint32_t)T);
        g.nrows = (int64_t)sp.in_rows.size();
        g.tile0 = ntiles;
        ntiles += (g.nrows + F - 1) / F;
        nrows += g.nrows;
        TablePtr tw = coeff_get(pass_twiddles_seq(sp.n, rad, sign));
        g.tw = static_cast<const float2 *>(tw->ptr);
        tables_.push_back(tw);
        radices_.push_back(rad);
        lds_ = std::max(lds_, (size_t)2 * F * lp * sizeof(float));
        host_.push_back(g);
    }
    if (ntiles >= (int64_t(1) << 31)) throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "too many tiles");
    // device image: groups | tile -> group | input rows | output rows
    const size_t gbytes = (host_.size() * sizeof(PersistGroup) + 255) / 256 * 256;
    const size_t tbytes = ((size_t)ntiles * sizeof(int32_t) + 255) / 256 * 256;
    const size_t rbytes = (size_t)nrows * sizeof(int64_t);
    std::vector<char> img(gbytes + tbytes + 2 * rbytes);
    SIP_HIP_CHECK(hipMalloc(&dev_, img.size()));
    char *d = static_cast<char *>(dev_);
    int32_t *tg = reinterpret_cast<int32_t *>(img.data() + gbytes);
    int64_t *ri = reinterpret_cast<int64_t *>(img.data() + gbytes + tbytes);
    int64_t *ro = reinterpret_cast<int64_t *>(img.data() + gbytes + tbytes + rbytes);
    int64_t row = 0;
    for (size_t i = 0; i < host_.size(); i++) {
        PersistGroup &g = host_[i];
        const PersistSpec &sp = specs[i];
        g.in_rows = reinterpret_cast<const int64_t *>(d + gbytes + tbytes) + row;
        g.out_rows = reinterpret_cast<const int64_t *>(d + gbytes + tbytes + rbytes) + row;
        std::copy(sp.in_rows.begin(), sp.in_rows.end(), ri + row);
        std::copy(sp.out_rows.begin(), sp.out_rows.end(), ro + row);
        row += g.nrows;
        const int64_t nt = (g.nrows + g.F - 1) / g.F;
        for (int64_t t = 0; t < nt; t++) tg[g.tile0 + t] = (int32_t)i;
    }
    std::memcpy(img.data(), host_.data(), host_.size() * sizeof(PersistGroup));
    const hipError_t ce = hipMemcpy(dev_, img.data(), img.size(), hipMemcpyHostToDevice);
    if (ce != hipSuccess) {
        (void)hipFree(dev_);  // the destructor does not run when the constructor throws
        dev_ = nullptr;
        SIP_HIP_CHECK(ce);
    }
    args_.in = static_cast<const float2 *>(in_base);
    args_.out = static_cast<float2 *>(out_base);
    args_.groups = reinterpret_cast<const PersistGroup *>(d);
    args_.tile_group = reinterpret_cast<const int32_t *>(d + gbytes);
    args_.ntiles = ntiles;
    // a fixed grid: every workgroup the device holds at once, walking the tiles with a stride
    const void *fn = g_persist_kernel[sign < 0 ? 0 : 1];
    if (lds_ > 64 * 1024) SIP_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_));
    int dev = 0, cus = 0, per_cu = 0;
    SIP_HIP_CHECK(hipGetDevice(&dev));
    SIP_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SIP_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPersistThreads, lds_));
    const int64_t resident = (int64_t)std::max(1, cus) * std::max(1, per_cu);
    grid_ = (unsigned)std::max<int64_t>(1, std::min(ntiles, resident));
}

PersistLaunch::~PersistLaunch()
{
    if (dev_) (void)hipFree(dev_);
}

void PersistLaunch::launch(hipStream_t stream) const
{
    PersistArgs a = args_;
    void *kargs[] = {&a};
    SIP_HIP_CHECK(hipLaunchKernel(g_persist_kernel[sign_ < 0 ? 0 : 1], dim3(grid_), dim3(kPersistThreads), kargs, lds_,
                                  stream));
}

std::string PersistLaunch::describe() const
{
    std::ostringstream os;
    os << "{\"groups\":" << host_.size() << ",\"tiles\":" << args_.ntiles << ",\"grid\":" << grid_ << ",\"lds\":" << lds_
       << ",\"group_list\":[";
    for (size_t i = 0; i < host_.size(); i++) {
        const PersistGroup &g = host_[i];
        os << (i ? "," : "") << "{\"n\":" << g.L << ",\"rows\":" << g.nrows << ",\"T\":" << g.T << ",\"F\":" << g.F
           << ",\"radices\":[";
        for (size_t p = 0; p < radices_[i].size(); p++) os << (p ? "," : "") << radices_[i][p];
        os << "]}";
    }
    os << "]}";
    return os.str();
}

}  // namespace sip
