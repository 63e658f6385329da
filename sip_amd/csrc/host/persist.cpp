// Host side of the persistent multi-length Stockham kernel (kernels/persist.hip): radix sequence and
// thread shape per length group, the tile list (largest groups first, as given), device
// descriptors and the launch. No counterpart in the reference's host code; the reference's
// bulk-per-core loop (ops/fft/common/include/kernel/fft_all_common.cceh:102-119) is the nearest.
#include "persist.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

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

namespace {

// Workgroups of the persistent kernel one CU holds at `lds` bytes of LDS (cached per device and size).
int persist_per_cu(const void *fn, size_t lds)
{
    static std::mutex mu;
    static std::map<std::pair<int, size_t>, int> cache;
    int dev = 0;
    SIP_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, lds});
    if (it != cache.end()) return it->second;
    if (lds > 64 * 1024) SIP_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    SIP_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPersistThreads, lds));
    cache[{dev, lds}] = std::max(1, per_cu);
    return std::max(1, per_cu);
}

}  // namespace

namespace {

// Group descriptor of one (device, n, radix sequence, sign) without its rows (a serving loop builds
// images of the same lengths over and over). The twiddle table is referenced weakly: it lives as
// long as an image (or a plan) uses it, and is fetched again from the coefficient cache after that.
struct GroupTemplate {
    PersistGroup g;
    std::vector<int> radices;
    std::weak_ptr<DeviceTable> tw_weak;
};

struct GroupRef {
    PersistGroup g;
    std::vector<int> radices;
    TablePtr tw;
};

GroupRef group_template(int64_t n, const std::vector<int> &requested, int sign)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int64_t, std::vector<int>, int>, GroupTemplate> cache;
    int dev = 0;
    SIP_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(dev, n, requested, sign);
    auto it = cache.find(key);
    if (it != cache.end()) {
        GroupTemplate &t = it->second;
        TablePtr tw = t.tw_weak.lock();
        if (!tw) {
            tw = coeff_get(pass_twiddles_seq(n, t.radices, sign));
            t.tw_weak = tw;
            t.g.tw = static_cast<const float2 *>(tw->ptr);
        }
        return GroupRef{t.g, t.radices, tw};
    }
    std::vector<int> rad = requested.empty() ? persist_default_radices(n) : requested;
    int T = 0, F = 0;
    if (rad.empty() || !persist_shape(n, rad, T, F))
        throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "length " + std::to_string(n) + " has no persistent-kernel shape");
    GroupTemplate t;
    PersistGroup &g = t.g;
    std::memset(&g, 0, sizeof(g));
    g.L = (int32_t)n;
    g.T = T;
    g.F = F;
    const int64_t maxidx = n - 1;
    int lp = (int)(maxidx + (maxidx >> 4)) + 1;
    if ((lp & 1) == 0) lp++;
    g.LP = lp;
    g.npass = (int32_t)rad.size();
    int64_t ns = 1, off = 0;
    for (size_t p = 0; p < rad.size(); p++) {
        g.radix[p] = rad[p];
        g.ns[p] = (int32_t)ns;
        g.nbf[p] = (int32_t)(n / rad[p]);
        g.tw_off[p] = (int32_t)off;
        g.div_ns[p] = make_fastdiv((uint32_t)ns);
        off += ns * pass_twiddle_rows(rad[p]);
        ns *= rad[p];
    }
    g.div_T = make_fastdiv((uint32_t)T);
    TablePtr tw = coeff_get(pass_twiddles_seq(n, rad, sign));
    g.tw = static_cast<const float2 *>(tw->ptr);
    t.radices = rad;
    t.tw_weak = tw;
    cache.emplace(key, t);
    return GroupRef{t.g, t.radices, tw};
}

int cu_count()
{
    static std::mutex mu;
    static std::map<int, int> cache;
    int dev = 0;
    SIP_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int cus = 0;
    SIP_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cache[dev] = std::max(1, cus);
    return cache[dev];
}

}  // namespace

void persist_build_image(const std::vector<PersistSpec> &specs, const void *in_base, void *out_base, int sign,
                         PersistImage &img)
{
    if (specs.empty()) throw SipError(AsdSip::ACL_ERROR_INVALID_PARAM, "persistent launch without groups");
    img = PersistImage{};
    img.sign = sign;
    int64_t ntiles = 0, nrows = 0;
    for (const PersistSpec &sp : specs) {
        if (sp.in_rows.size() != sp.out_rows.size())
            throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "persistent group with unequal row lists");
        const GroupRef tpl = group_template(sp.n, sp.radices, sign);
        PersistGroup g = tpl.g;
        g.nrows = (int64_t)sp.in_rows.size();
        g.tile0 = ntiles;
        ntiles += (g.nrows + g.F - 1) / g.F;
        nrows += g.nrows;
        img.tables.push_back(tpl.tw);
        img.radices.push_back(tpl.radices);
        img.lds = std::max(img.lds, (size_t)2 * g.F * g.LP * sizeof(float));
        img.groups.push_back(g);
    }
    if (ntiles >= (int64_t(1) << 31)) throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "too many tiles");
    // image: groups | tile -> group | input rows | output rows
    img.gbytes = (img.groups.size() * sizeof(PersistGroup) + 255) / 256 * 256;
    img.tbytes = ((size_t)ntiles * sizeof(int32_t) + 255) / 256 * 256;
    img.rbytes = (size_t)nrows * sizeof(int64_t);
    img.bytes.assign(img.gbytes + img.tbytes + 2 * img.rbytes, 0);
    int32_t *tg = reinterpret_cast<int32_t *>(img.bytes.data() + img.gbytes);
    int64_t *ri = reinterpret_cast<int64_t *>(img.bytes.data() + img.gbytes + img.tbytes);
    int64_t *ro = reinterpret_cast<int64_t *>(img.bytes.data() + img.gbytes + img.tbytes + img.rbytes);
    int64_t row = 0;
    for (size_t i = 0; i < img.groups.size(); i++) {
        PersistGroup &g = img.groups[i];
        const PersistSpec &sp = specs[i];
        std::copy(sp.in_rows.begin(), sp.in_rows.end(), ri + row);
        std::copy(sp.out_rows.begin(), sp.out_rows.end(), ro + row);
        g.in_rows = reinterpret_cast<const int64_t *>((uintptr_t)row);  // element index until placed
        row += g.nrows;
        const int64_t nt = (g.nrows + g.F - 1) / g.F;
        for (int64_t t = 0; t < nt; t++) tg[g.tile0 + t] = (int32_t)i;
    }
    img.args.in = static_cast<const float2 *>(in_base);
    img.args.out = static_cast<float2 *>(out_base);
    img.args.ntiles = ntiles;
    // a fixed grid: every workgroup the device holds at once, walking the tiles with a stride
    const int64_t resident = (int64_t)cu_count() * persist_per_cu(g_persist_kernel[sign < 0 ? 0 : 1], img.lds);
    img.grid = (unsigned)std::max<int64_t>(1, std::min(ntiles, resident));
}

void persist_place(PersistImage &img, void *dev_base)
{
    char *d = static_cast<char *>(dev_base);
    std::vector<PersistGroup> g = img.groups;
    for (PersistGroup &x : g) {
        const int64_t row = (int64_t)reinterpret_cast<uintptr_t>(x.in_rows);
        x.in_rows = reinterpret_cast<const int64_t *>(d + img.gbytes + img.tbytes) + row;
        x.out_rows = reinterpret_cast<const int64_t *>(d + img.gbytes + img.tbytes + img.rbytes) + row;
    }
    std::memcpy(img.bytes.data(), g.data(), g.size() * sizeof(PersistGroup));
    img.args.groups = reinterpret_cast<const PersistGroup *>(d);
    img.args.tile_group = reinterpret_cast<const int32_t *>(d + img.gbytes);
}

void persist_launch(const PersistImage &img, hipStream_t stream)
{
    PersistArgs a = img.args;
    void *kargs[] = {&a};
    SIP_HIP_CHECK(hipLaunchKernel(g_persist_kernel[img.sign < 0 ? 0 : 1], dim3(img.grid), dim3(kPersistThreads), kargs,
                                  img.lds, stream));
}

PersistLaunch::PersistLaunch(const std::vector<PersistSpec> &specs, const void *in_base, void *out_base, int sign)
{
    persist_build_image(specs, in_base, out_base, sign, img_);
    SIP_HIP_CHECK(hipMalloc(&dev_, img_.bytes.size()));
    persist_place(img_, dev_);
    const hipError_t ce = hipMemcpy(dev_, img_.bytes.data(), img_.bytes.size(), hipMemcpyHostToDevice);
    if (ce != hipSuccess) {
        (void)hipFree(dev_);  // the destructor does not run when the constructor throws
        dev_ = nullptr;
        SIP_HIP_CHECK(ce);
    }
}

PersistLaunch::~PersistLaunch()
{
    if (dev_) (void)hipFree(dev_);
}

void PersistLaunch::launch(hipStream_t stream) const { persist_launch(img_, stream); }

std::string PersistLaunch::describe() const
{
    std::ostringstream os;
    os << "{\"groups\":" << img_.groups.size() << ",\"tiles\":" << img_.args.ntiles << ",\"grid\":" << img_.grid
       << ",\"lds\":" << img_.lds << ",\"group_list\":[";
    for (size_t i = 0; i < img_.groups.size(); i++) {
        const PersistGroup &g = img_.groups[i];
        os << (i ? "," : "") << "{\"n\":" << g.L << ",\"rows\":" << g.nrows << ",\"T\":" << g.T << ",\"F\":" << g.F
           << ",\"radices\":[";
        for (size_t p = 0; p < img_.radices[i].size(); p++) os << (p ? "," : "") << img_.radices[i][p];
        os << "]}";
    }
    os << "]}";
    return os.str();
}

}  // namespace sip
