// Device-resident coefficient cache (twiddles, real-transform roots, four-step and Bluestein
// tables). Counterpart of the reference's FFTensorCache (core/fft/utils/fftensor_cache.cpp:55-92):
// tables are keyed by their content description, held through weak_ptr so that a table lives as
// long as one plan uses it, created on first use with one host->device copy at plan time, and
// kept per device.
#include <map>
#include <mutex>

#include "plan.h"

namespace sip {

DeviceTable::~DeviceTable()
{
    if (ptr) {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != device && device >= 0) {
            (void)hipSetDevice(device);
            (void)hipFree(ptr);
            (void)hipSetDevice(cur);
        } else {
            (void)hipFree(ptr);
        }
    }
}

static std::mutex g_cache_mu;
static std::map<std::pair<int, std::string>, std::weak_ptr<DeviceTable>> g_cache;

TablePtr coeff_get(const TableRef &ref)
{
    int dev = 0;
    SIP_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto key = std::make_pair(dev, ref.key);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
        if (TablePtr live = it->second.lock()) return live;
    }
    // entries whose last plan is gone are dropped here (tables are created at plan time only), so
    // a long-lived process that sees many shapes keeps a map of live tables only
    for (auto e = g_cache.begin(); e != g_cache.end();) {
        if (e->second.expired())
            e = g_cache.erase(e);
        else
            ++e;
    }
    std::vector<float2> host = ref.gen();
    auto t = std::make_shared<DeviceTable>();
    t->bytes = host.size() * sizeof(float2);
    t->device = dev;
    SIP_HIP_CHECK(hipMalloc(&t->ptr, t->bytes));
    SIP_HIP_CHECK(hipMemcpy(t->ptr, host.data(), t->bytes, hipMemcpyHostToDevice));
    g_cache[key] = t;
    SIP_LOG(DEBUG, "coefficient table " << ref.key << " (" << t->bytes << " B) uploaded");
    return t;
}

size_t coeff_cache_live_tables()
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    size_t n = 0;
    for (auto &kv : g_cache)
        if (!kv.second.expired()) n++;
    return n;
}

void coeff_cache_stats(size_t *entries, size_t *live, size_t *live_bytes)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    size_t n = 0, b = 0;
    for (auto &kv : g_cache)
        if (TablePtr t = kv.second.lock()) {
            n++;
            b += t->bytes;
        }
    if (entries) *entries = g_cache.size();
    if (live) *live = n;
    if (live_bytes) *live_bytes = b;
}

}  // namespace sip

extern "C" int32_t sipFftCoeffCacheStats(size_t *entries, size_t *live_tables, size_t *live_bytes)
{
    sip::coeff_cache_stats(entries, live_tables, live_bytes);
    return AsdSip::ACL_SUCCESS;
}
