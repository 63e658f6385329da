// Batching form: many C2C FFT problems of uneven lengths executed as one hipGraph launch.
//
// Problem i is batches[i] transforms of length lengths[i], stored contiguously at in_ptrs[i]
// (written to out_ptrs[i]). Problems of the same length are grouped into one gathered-row plan
// (plan_build_gathered_c2c): a single Stockham launch (or four-step pair) walks the rows of every
// member through a device row-offset table, so no data is copied or packed. The groups are spread
// over a few parallel chains (longest-processing-time first) and captured once into a hipGraph;
// each exec is one graph launch. An optional autotuner times the runtime tile-shape choices of every
// group (transforms per workgroup, LDS layout, tile order, four-step split) on scratch rows of the
// group's shape and keeps the fastest, caching decisions in a JSON file keyed by length, row count,
// direction and device. With flag bit 2 every short length group (2,3,5-smooth, up to 4096) instead
// joins one persistent multi-length launch (kernels/persist.hip) whose radix sequence per group is
// runtime data, chosen by the tuner among candidate orders when autotuning. This layer has no
// counterpart in the reference; it is an extension API (sipFftBatch*).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <unordered_map>

#include "persist.h"
#include "plan.h"

namespace sip {
namespace {

// Rows from which a group of length 16384 < n <= 32768 runs as one single-pass launch (plain plans:
// 256). Few rows leave a lone launch short of the chip, but the graph's concurrent chains fill it:
// measured on MI355X, the batching workload's graph 8.60 -> 8.48 ms against 256 (16 rows: 8.49).
constexpr int64_t kBatchLeafRows = 64;

struct Group {
    int64_t n = 0;
    std::vector<int> members;
    int64_t rows = 0;
    double cost = 0;  // estimated bytes moved
    std::unique_ptr<Plan> plan;
    int chain = 0;
    float tuned_us = -1.f;
};

struct BatchPlan {
    int direction = kDirForward;
    std::vector<int64_t> lengths, batches;
    std::vector<void *> ins, outs;
    std::vector<Group> groups;
    // short length groups run by one persistent multi-length launch (flag bit 2), on chain
    // persist_chain ahead of that chain's groups
    std::unique_ptr<PersistLaunch> persist;
    int persist_chain = 0;
    std::vector<float> persist_tuned_us;
    int nchains = 1;
    std::vector<size_t> chain_ws;      // workspace bytes per chain
    std::vector<size_t> chain_ws_off;  // offset of each chain's region
    size_t ws_bytes = 0;
    void *workspace = nullptr;
    hipStream_t stream = nullptr;
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipStream_t> side;  // capture-only side streams
    std::vector<hipEvent_t> events;
    hipStream_t own = nullptr;      // capture stream when the caller's stream is the null stream
    hipEvent_t own_in = nullptr, own_out = nullptr;
    std::mutex mu;
    bool ready = false;

    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        if (own_in) (void)hipEventDestroy(own_in);
        if (own_out) (void)hipEventDestroy(own_out);
        if (own) (void)hipStreamDestroy(own);
    }

    // The null stream cannot be captured: graphs are captured on (and launched into) a private
    // stream ordered after / before the caller's null-stream work with events.
    hipStream_t work_stream() const { return stream ? stream : own; }
};

std::mutex g_batch_mu;
std::unordered_map<void *, std::unique_ptr<BatchPlan>> g_batches;

BatchPlan *lookup_batch(void *h)
{
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto it = g_batches.find(h);
    return it == g_batches.end() ? nullptr : it->second.get();
}

// ------------------------------------------------------------------ tuning cache (JSON lines)

std::string cache_path()
{
    if (const char *e = std::getenv("SIP_FFT_TUNE_CACHE")) return e;
    const char *home = std::getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.sip_fft_tune.jsonl";
}

std::string device_name()
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    return prop.gcnArchName;
}

std::string tune_key(int64_t n, int64_t rows, int direction)
{
    int64_t bucket = 1;
    while (bucket < rows) bucket <<= 1;
    std::ostringstream os;
    os << "c2c.2:" << n << ":" << bucket << ":" << direction << ":" << device_name();
    return os.str();
}

std::mutex g_cache_mu;
std::map<std::string, Tune> g_tune_cache;
std::map<std::string, std::vector<int>> g_radix_cache;  // persistent-launch radix sequences
bool g_cache_loaded = false;

// {"key":"...","radices":[a,b,...]}
bool parse_radix_line(const std::string &line, std::string &key, std::vector<int> &rad)
{
    const std::string kk = "{\"key\":\"", rr = "\"radices\":[";
    if (line.compare(0, kk.size(), kk) != 0) return false;
    const size_t ke = line.find('"', kk.size());
    const size_t rp = line.find(rr);
    if (ke == std::string::npos || rp == std::string::npos) return false;
    key = line.substr(kk.size(), ke - kk.size());
    rad.clear();
    size_t i = rp + rr.size();
    while (i < line.size() && line[i] != ']') {
        char *end = nullptr;
        const long v = std::strtol(line.c_str() + i, &end, 10);
        if (end == line.c_str() + i) return false;
        rad.push_back((int)v);
        i = (size_t)(end - line.c_str());
        if (i < line.size() && line[i] == ',') i++;
    }
    return !rad.empty();
}

void load_cache()
{
    if (g_cache_loaded) return;
    g_cache_loaded = true;
    std::ifstream f(cache_path());
    std::string line;
    while (std::getline(f, line)) {
        char key[256];
        Tune t;
        long long l1 = 0;
        if (std::sscanf(line.c_str(),
                        "{\"key\":\"%255[^\"]\",\"row_F\":%d,\"split_lds\":%d,\"xcd\":%d,\"split_l1\":%lld,\"col_F\":%d,"
                        "\"rc\":%d}",
                        key, &t.row_F, &t.split_lds, &t.xcd, &l1, &t.col_F, &t.rc) == 7) {
            t.split_l1 = l1;
            g_tune_cache[key] = t;
        } else {
            std::string k;
            std::vector<int> rad;
            if (parse_radix_line(line, k, rad)) g_radix_cache[k] = rad;
        }
    }
}

void store_cache(const std::string &key, const Tune &t)
{
    g_tune_cache[key] = t;
    std::ofstream f(cache_path(), std::ios::app);
    if (!f) return;  // an unwritable cache only costs re-tuning
    f << "{\"key\":\"" << key << "\",\"row_F\":" << t.row_F << ",\"split_lds\":" << t.split_lds << ",\"xcd\":" << t.xcd
      << ",\"split_l1\":" << t.split_l1 << ",\"col_F\":" << t.col_F << ",\"rc\":" << t.rc << "}\n";
}

// ------------------------------------------------------------------ group plans

std::unique_ptr<Plan> build_group_plan(const BatchPlan &bp, const Group &g, const Tune &t)
{
    auto p = std::make_unique<Plan>();
    p->direction = bp.direction;
    p->tune = t;
    p->big_leaf_rows = kBatchLeafRows;
    // element offsets relative to the lowest member pointer of each side
    uintptr_t in_lo = UINTPTR_MAX, out_lo = UINTPTR_MAX;
    for (int i : g.members) {
        in_lo = std::min(in_lo, reinterpret_cast<uintptr_t>(bp.ins[(size_t)i]));
        out_lo = std::min(out_lo, reinterpret_cast<uintptr_t>(bp.outs[(size_t)i]));
    }
    std::vector<int64_t> in_rows, out_rows;
    in_rows.reserve((size_t)g.rows);
    out_rows.reserve((size_t)g.rows);
    for (int i : g.members) {
        const int64_t ib = (int64_t)((reinterpret_cast<uintptr_t>(bp.ins[(size_t)i]) - in_lo) / 8);
        const int64_t ob = (int64_t)((reinterpret_cast<uintptr_t>(bp.outs[(size_t)i]) - out_lo) / 8);
        for (int64_t r = 0; r < bp.batches[(size_t)i]; r++) {
            in_rows.push_back(ib + r * g.n);
            out_rows.push_back(ob + r * g.n);
        }
    }
    plan_build_gathered_c2c(*p, g.n, in_rows, out_rows, reinterpret_cast<void *>(in_lo),
                            reinterpret_cast<void *>(out_lo));
    plan_materialize(*p);
    p->initialized = true;
    return p;
}

// The plan the tuner times for a group: the same length, row count and tile choices, on compact
// scratch rows (row r at element r * n of scratch_in / scratch_out). The caller's buffers are never
// touched at plan time (an in-place problem would otherwise be transformed by every timing run).
std::unique_ptr<Plan> build_tuning_plan(const BatchPlan &bp, const Group &g, const Tune &t, void *scratch_in,
                                        void *scratch_out)
{
    auto p = std::make_unique<Plan>();
    p->direction = bp.direction;
    p->tune = t;
    p->big_leaf_rows = kBatchLeafRows;
    std::vector<int64_t> rows((size_t)g.rows);
    for (int64_t r = 0; r < g.rows; r++) rows[(size_t)r] = r * g.n;
    plan_build_gathered_c2c(*p, g.n, rows, rows, scratch_in, scratch_out);
    if (!plan_launchable(*p)) throw SipError(AsdSip::ACL_ERROR_API_NOT_SUPPORT, "candidate exceeds launch limits");
    plan_materialize(*p);
    p->initialized = true;
    return p;
}

// Candidate tile-shape choices for one group.
std::vector<Tune> candidates(int64_t n)
{
    std::vector<Tune> c;
    c.push_back(Tune{});
    const StockhamInstance *inst = find_instance(n);
    if (inst && n <= 16384) {
        for (int sl : {0, 1})
            for (int x : {0, 1}) {
                if (sl == 1 && !inst->rows_fwd) continue;
                Tune t;
                t.split_lds = sl;
                t.xcd = x;
                c.push_back(t);
            }
        const int T = inst->T;
        for (int f : {1, 2, 4, 8, 16}) {
            if (f * T > 1024 || f * T < 64) continue;
            Tune t;
            t.row_F = f;
            t.split_lds = 0;
            c.push_back(t);
        }
    } else {
        // four-step splits n = a * b: the 8 most balanced ones (the default plan takes the most
        // balanced pair of column leaves), each with three column-tile widths and the second stage on
        // the rows-in / columns-out kernel or on the two-plane tile
        std::vector<std::pair<double, int64_t>> splits;
        for (int64_t a = 2; a <= 1024 && a < n; a++) {
            if (n % a || !find_instance(a)) continue;
            const int64_t b = n / a;
            if (b > 16384) continue;
            if (a * 16 < b || b * 16 < a) continue;  // keep splits within 16x of balanced
            splits.emplace_back(std::fabs(std::log((double)a) - std::log((double)b)), a);
        }
        std::sort(splits.begin(), splits.end());
        if (splits.size() > 8) splits.resize(8);
        for (const auto &sp : splits)
            for (int cf : {0, 16, 32})
                for (int rc : {0, 1}) {
                    Tune t;
                    t.split_l1 = sp.second;
                    t.col_F = cf;
                    t.rc = rc;
                    c.push_back(t);
                }
    }
    return c;
}

float time_plan(Plan &p, hipStream_t s, int reps)
{
    hipEvent_t e0, e1;
    SIP_HIP_CHECK(hipEventCreate(&e0));
    SIP_HIP_CHECK(hipEventCreate(&e1));
    p.stream = s;
    plan_execute(p, nullptr, nullptr);  // warm-up
    SIP_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) plan_execute(p, nullptr, nullptr);
    SIP_HIP_CHECK(hipEventRecord(e1, s));
    SIP_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    SIP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms * 1000.f / reps;
}

struct TuneScratch {
    void *in = nullptr, *out = nullptr, *ws = nullptr;
    size_t ws_bytes = 0;
};

void tune_group(BatchPlan &bp, Group &g, const TuneScratch &sc)
{
    const std::string key = tune_key(g.n, g.rows, bp.direction);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        load_cache();
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            g.plan = build_group_plan(bp, g, it->second);
            return;
        }
    }
    std::vector<Tune> cands = candidates(g.n);
    // A length above 16384 that the default plan runs as one single-pass launch keeps it. Alone,
    // the four-step split can time faster (its two launches fill the chip where a few hundred
    // one-workgroup transforms do not), but inside the graph the concurrent chains fill the chip
    // and the split's second HBM pass costs more: measured on MI355X, the batching workload
    // 8.92 -> 8.56 ms with the 11 such groups the timing had split kept single-pass.
    if (g.n > 16384) {
        try {
            if (build_tuning_plan(bp, g, Tune{}, sc.in, sc.out)->stages.size() == 1) cands.assign(1, Tune{});
        } catch (const SipError &) {
        }
    }
    Tune best;
    float best_us = 1e30f;
    for (const Tune &t : cands) {
        std::unique_ptr<Plan> p;
        try {
            p = build_tuning_plan(bp, g, t, sc.in, sc.out);
        } catch (const SipError &) {
            continue;  // not plannable or beyond the device's launch limits: never launched
        }
        if (p->ws_bytes > sc.ws_bytes) continue;
        p->workspace = sc.ws;
        const float us = time_plan(*p, bp.stream, 5);  // null stream is fine for timing
        if (us < best_us) {
            best_us = us;
            best = t;
        }
    }
    g.tuned_us = best_us;
    g.plan = build_group_plan(bp, g, best);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    store_cache(key, best);
}

// Radix order of one short group in the persistent launch: each candidate sequence
// (persist_candidate_radices) runs the group alone on scratch rows, the fastest is kept and cached.
std::vector<int> tune_persist(BatchPlan &bp, const Group &g, const TuneScratch &sc, float &best_us)
{
    const std::string key = "persist:" + tune_key(g.n, g.rows, bp.direction);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        load_cache();
        auto it = g_radix_cache.find(key);
        if (it != g_radix_cache.end()) return it->second;
    }
    PersistSpec sp;
    sp.n = g.n;
    sp.in_rows.resize((size_t)g.rows);
    for (int64_t r = 0; r < g.rows; r++) sp.in_rows[(size_t)r] = r * g.n;
    sp.out_rows = sp.in_rows;
    std::vector<int> best;
    best_us = 1e30f;
    hipEvent_t e0, e1;
    SIP_HIP_CHECK(hipEventCreate(&e0));
    SIP_HIP_CHECK(hipEventCreate(&e1));
    try {
        for (const std::vector<int> &cand : persist_candidate_radices(g.n, 6)) {
            sp.radices = cand;
            PersistLaunch pl({sp}, sc.in, sc.out, bp.direction == kDirInverse ? 1 : -1);
            pl.launch(bp.stream);  // warm-up
            SIP_HIP_CHECK(hipEventRecord(e0, bp.stream));
            for (int r = 0; r < 5; r++) pl.launch(bp.stream);
            SIP_HIP_CHECK(hipEventRecord(e1, bp.stream));
            SIP_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            SIP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms * 200.f < best_us) {
                best_us = ms * 200.f;  // us per launch over 5 launches
                best = cand;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best.empty()) return best;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_radix_cache[key] = best;
    std::ofstream f(cache_path(), std::ios::app);
    if (f) {
        f << "{\"key\":\"" << key << "\",\"radices\":[";
        for (size_t i = 0; i < best.size(); i++) f << (i ? "," : "") << best[i];
        f << "]}\n";
    }
    return best;
}

// The persistent launch over the short groups (largest first): rows at element offsets from the
// lowest input / output pointer of those groups' problems.
std::unique_ptr<PersistLaunch> build_persist(const BatchPlan &bp, const std::vector<Group> &pg,
                                             const std::vector<std::vector<int>> &radices)
{
    uintptr_t in_lo = UINTPTR_MAX, out_lo = UINTPTR_MAX;
    for (const Group &g : pg)
        for (int i : g.members) {
            in_lo = std::min(in_lo, reinterpret_cast<uintptr_t>(bp.ins[(size_t)i]));
            out_lo = std::min(out_lo, reinterpret_cast<uintptr_t>(bp.outs[(size_t)i]));
        }
    std::vector<PersistSpec> specs(pg.size());
    for (size_t k = 0; k < pg.size(); k++) {
        const Group &g = pg[k];
        PersistSpec &sp = specs[k];
        sp.n = g.n;
        sp.radices = radices[k];
        sp.in_rows.reserve((size_t)g.rows);
        sp.out_rows.reserve((size_t)g.rows);
        for (int i : g.members) {
            const int64_t ib = (int64_t)((reinterpret_cast<uintptr_t>(bp.ins[(size_t)i]) - in_lo) / 8);
            const int64_t ob = (int64_t)((reinterpret_cast<uintptr_t>(bp.outs[(size_t)i]) - out_lo) / 8);
            for (int64_t r = 0; r < bp.batches[(size_t)i]; r++) {
                sp.in_rows.push_back(ib + r * g.n);
                sp.out_rows.push_back(ob + r * g.n);
            }
        }
    }
    return std::make_unique<PersistLaunch>(specs, reinterpret_cast<const void *>(in_lo), reinterpret_cast<void *>(out_lo),
                                           bp.direction == kDirInverse ? 1 : -1);
}

void capture(BatchPlan &bp)
{
    // Chains run concurrently: chain 0 on the plan's stream, the others on side streams forked and
    // joined with events, so the graph holds nchains independent branches.
    if (bp.exec) {
        (void)hipGraphExecDestroy(bp.exec);
        bp.exec = nullptr;
    }
    if (bp.graph) {
        (void)hipGraphDestroy(bp.graph);
        bp.graph = nullptr;
    }
    for (Group &g : bp.groups) g.plan->workspace = static_cast<char *>(bp.workspace) + bp.chain_ws_off[(size_t)g.chain];
    auto enqueue_all = [&](bool graph_mode) {
        hipStream_t main = bp.work_stream();
        if (bp.nchains > 1) {
            SIP_HIP_CHECK(hipEventRecord(bp.events[0], main));
            for (int c = 1; c < bp.nchains; c++) SIP_HIP_CHECK(hipStreamWaitEvent(bp.side[(size_t)c - 1], bp.events[0], 0));
        }
        if (bp.persist) bp.persist->launch(bp.persist_chain == 0 ? main : bp.side[(size_t)bp.persist_chain - 1]);
        for (Group &g : bp.groups) {
            g.plan->stream = g.chain == 0 ? main : bp.side[(size_t)g.chain - 1];
            plan_execute(*g.plan, nullptr, nullptr);
        }
        if (bp.nchains > 1) {
            for (int c = 1; c < bp.nchains; c++) {
                SIP_HIP_CHECK(hipEventRecord(bp.events[(size_t)c], bp.side[(size_t)c - 1]));
                SIP_HIP_CHECK(hipStreamWaitEvent(main, bp.events[(size_t)c], 0));
            }
        }
        (void)graph_mode;
    };
    if (!bp.use_graph) {
        enqueue_all(false);
        return;
    }
    hipStream_t cs = bp.work_stream();
    SIP_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_all(true);
    } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(cs, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
    }
    SIP_HIP_CHECK(hipStreamEndCapture(cs, &bp.graph));
    SIP_HIP_CHECK(hipGraphInstantiate(&bp.exec, bp.graph, nullptr, nullptr, 0));
}

template <class F>
int32_t guarded(F &&f)
{
    try {
        return f();
    } catch (const SipError &e) {
        SIP_LOG(ERROR, e.what());
        return e.status;
    } catch (const std::exception &e) {
        SIP_LOG(ERROR, e.what());
        return AsdSip::ACL_ERROR_INTERNAL_ERROR;
    } catch (...) {
        return AsdSip::ACL_ERROR_INTERNAL_ERROR;
    }
}

}  // namespace
}  // namespace sip

using sip::BatchPlan;
namespace E = AsdSip;

extern "C" {

int32_t sipFftBatchCreate(void **handle)
{
    return sip::guarded([&]() -> int32_t {
        if (!handle) return E::ACL_ERROR_INVALID_PARAM;
        auto b = std::make_unique<BatchPlan>();
        void *h = b.get();
        std::lock_guard<std::mutex> lk(sip::g_batch_mu);
        sip::g_batches.emplace(h, std::move(b));
        *handle = h;
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftBatchSetStream(void *handle, void *stream)
{
    return sip::guarded([&]() -> int32_t {
        BatchPlan *b = sip::lookup_batch(handle);
        if (!b) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(b->mu);
        b->stream = static_cast<hipStream_t>(stream);
        b->ready = false;
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftBatchMakePlan(void *handle, int32_t count, const int64_t *lengths, const int64_t *batches,
                            void *const *in_ptrs, void *const *out_ptrs, int32_t direction, int32_t flags)
{
    return sip::guarded([&]() -> int32_t {
        BatchPlan *b = sip::lookup_batch(handle);
        if (!b || count <= 0 || !lengths || !batches || !in_ptrs || !out_ptrs) return E::ACL_ERROR_INVALID_PARAM;
        if (direction != sip::kDirForward && direction != sip::kDirInverse) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(b->mu);
        b->direction = direction;
        b->lengths.assign(lengths, lengths + count);
        b->batches.assign(batches, batches + count);
        b->ins.assign(in_ptrs, in_ptrs + count);
        b->outs.assign(out_ptrs, out_ptrs + count);
        b->use_graph = (flags & 1) != 0;
        for (int i = 0; i < count; i++) {
            if (lengths[i] <= 0 || lengths[i] > sip::kMaxFftSize || batches[i] <= 0) return E::ACL_ERROR_INVALID_PARAM;
            if (!in_ptrs[i] || !out_ptrs[i]) return E::ACL_ERROR_INVALID_PARAM;
            if ((reinterpret_cast<uintptr_t>(in_ptrs[i]) | reinterpret_cast<uintptr_t>(out_ptrs[i])) & 7)
                return E::ACL_ERROR_INVALID_PARAM;
        }
        // group by length
        std::map<int64_t, sip::Group> by;
        for (int i = 0; i < count; i++) {
            sip::Group &g = by[lengths[i]];
            g.n = lengths[i];
            g.members.push_back(i);
            g.rows += batches[i];
        }
        b->groups.clear();
        b->persist.reset();
        b->persist_tuned_us.clear();
        std::vector<sip::Group> pgroups;  // short lengths for the persistent launch (flag bit 2)
        for (auto &kv : by) {
            kv.second.cost = 16.0 * (double)kv.second.n * (double)kv.second.rows;
            if ((flags & 4) && !sip::persist_default_radices(kv.second.n).empty())
                pgroups.push_back(std::move(kv.second));
            else
                b->groups.push_back(std::move(kv.second));
        }
        std::sort(b->groups.begin(), b->groups.end(), [](const sip::Group &x, const sip::Group &y) { return x.cost > y.cost; });
        std::sort(pgroups.begin(), pgroups.end(), [](const sip::Group &x, const sip::Group &y) { return x.cost > y.cost; });
        // chains: longest processing time first over the group launches and the persistent launch
        // (one item whose cost is its groups' bytes)
        double pcost = 0;
        for (const sip::Group &g : pgroups) pcost += g.cost;
        // parallel chains (tuning switch SIP_FFT_BATCH_CHAINS): measured on MI355X, the batching
        // workload's graph 8.37-8.52 ms on 4 chains, 8.17-8.24 on 8, no better on 12 or 16
        size_t max_chains = 8;
        if (const char *e = std::getenv("SIP_FFT_BATCH_CHAINS")) max_chains = (size_t)std::max(1, std::min(16, std::atoi(e)));
        b->nchains = (int)std::min<size_t>(max_chains, b->groups.size() + (pgroups.empty() ? 0 : 1));
        std::vector<double> load((size_t)b->nchains, 0.0);
        bool persist_placed = pgroups.empty();
        auto place = [&](double cost) {
            const size_t c = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            load[c] += cost + 2.0e5;  // + a fixed per-launch cost
            return (int)c;
        };
        for (sip::Group &g : b->groups) {
            if (!persist_placed && pcost >= g.cost) {
                b->persist_chain = place(pcost);
                persist_placed = true;
            }
            g.chain = place(g.cost);
        }
        if (!persist_placed) b->persist_chain = place(pcost);
        // plans (tuned or default); tuning needs a scratch workspace sized for the largest candidate
        // Tuning runs on scratch rows (input, output and workspace of the largest group), zeroed
        // once; the caller's buffers are not read or written before the first exec.
        const bool autotune = (flags & 2) != 0;
        void *scratch = nullptr;
        sip::TuneScratch sc;
        if (autotune) {
            size_t rows_bytes = 0, ws_bytes = 0;
            for (const sip::Group &g : pgroups) rows_bytes = std::max(rows_bytes, (size_t)g.n * (size_t)g.rows * 8);
            for (sip::Group &g : b->groups) {
                rows_bytes = std::max(rows_bytes, (size_t)g.n * (size_t)g.rows * 8);
                // four-step and Bluestein intermediates of the default layout (host-only plan)
                sip::Plan dry;
                dry.direction = b->direction;
                std::vector<int64_t> rr((size_t)g.rows, 0);
                sip::plan_build_gathered_c2c(dry, g.n, rr, rr, nullptr, nullptr);
                ws_bytes = std::max(ws_bytes, dry.ws_bytes);
            }
            rows_bytes = (rows_bytes + 255) / 256 * 256;
            sc.ws_bytes = std::max(ws_bytes, rows_bytes) + 4096;
            SIP_HIP_CHECK(hipMalloc(&scratch, 2 * rows_bytes + sc.ws_bytes));
            sc.in = scratch;
            sc.out = static_cast<char *>(scratch) + rows_bytes;
            sc.ws = static_cast<char *>(scratch) + 2 * rows_bytes;
            const hipError_t ze = hipMemset(scratch, 0, 2 * rows_bytes + sc.ws_bytes);
            if (ze != hipSuccess) {
                (void)hipFree(scratch);
                SIP_HIP_CHECK(ze);
            }
        }
        try {
            for (sip::Group &g : b->groups) {
                if (autotune)
                    sip::tune_group(*b, g, sc);
                else
                    g.plan = sip::build_group_plan(*b, g, sip::Tune{});
            }
            if (!pgroups.empty()) {
                std::vector<std::vector<int>> rad(pgroups.size());
                for (size_t i = 0; i < pgroups.size(); i++) {
                    float us = -1.f;
                    if (autotune) rad[i] = sip::tune_persist(*b, pgroups[i], sc, us);
                    b->persist_tuned_us.push_back(us);
                }
                b->persist = sip::build_persist(*b, pgroups, rad);
            }
        } catch (...) {
            if (scratch) {
                (void)hipDeviceSynchronize();
                (void)hipFree(scratch);
            }
            throw;
        }
        if (scratch) SIP_HIP_CHECK(hipFree(scratch));
        // workspace: chains run concurrently, groups within a chain sequentially
        b->chain_ws.assign((size_t)b->nchains, 0);
        for (sip::Group &g : b->groups)
            b->chain_ws[(size_t)g.chain] = std::max(b->chain_ws[(size_t)g.chain], g.plan->ws_bytes);
        b->chain_ws_off.assign((size_t)b->nchains, 0);
        size_t off = 0;
        for (int c = 0; c < b->nchains; c++) {
            b->chain_ws_off[(size_t)c] = off;
            off += (b->chain_ws[(size_t)c] + 255) / 256 * 256;
        }
        b->ws_bytes = off;
        b->ready = false;
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftBatchGetWorkspaceSize(void *handle, size_t *bytes)
{
    BatchPlan *b = sip::lookup_batch(handle);
    if (!b || !bytes) return E::ACL_ERROR_INVALID_PARAM;
    std::lock_guard<std::mutex> lk(b->mu);
    *bytes = b->ws_bytes;
    return E::ACL_SUCCESS;
}

int32_t sipFftBatchSetWorkspace(void *handle, void *ws)
{
    BatchPlan *b = sip::lookup_batch(handle);
    if (!b) return E::ACL_ERROR_INVALID_PARAM;
    std::lock_guard<std::mutex> lk(b->mu);
    b->workspace = ws;
    b->ready = false;
    return E::ACL_SUCCESS;
}

int32_t sipFftBatchExec(void *handle)
{
    return sip::guarded([&]() -> int32_t {
        BatchPlan *b = sip::lookup_batch(handle);
        if (!b) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(b->mu);
        if (b->groups.empty() && !b->persist) return E::ACL_ERROR_INVALID_PARAM;
        if (b->ws_bytes > 0 && !b->workspace) return E::ACL_ERROR_INTERNAL_ERROR;
        if (b->side.size() + 1 < (size_t)b->nchains) {
            for (int c = (int)b->side.size() + 1; c < b->nchains; c++) {
                hipStream_t s;
                SIP_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
                b->side.push_back(s);
            }
            while (b->events.size() < (size_t)b->nchains) {
                hipEvent_t e;
                SIP_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                b->events.push_back(e);
            }
        }
        if (!b->use_graph) {
            sip::capture(*b);  // direct launches
            return E::ACL_SUCCESS;
        }
        if (!b->stream && !b->own) {
            SIP_HIP_CHECK(hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking));
            SIP_HIP_CHECK(hipEventCreateWithFlags(&b->own_in, hipEventDisableTiming));
            SIP_HIP_CHECK(hipEventCreateWithFlags(&b->own_out, hipEventDisableTiming));
        }
        if (!b->ready) {
            sip::capture(*b);
            b->ready = true;
        }
        if (b->stream) {
            SIP_HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
        } else {
            SIP_HIP_CHECK(hipEventRecord(b->own_in, nullptr));
            SIP_HIP_CHECK(hipStreamWaitEvent(b->own, b->own_in, 0));
            SIP_HIP_CHECK(hipGraphLaunch(b->exec, b->own));
            SIP_HIP_CHECK(hipEventRecord(b->own_out, b->own));
            SIP_HIP_CHECK(hipStreamWaitEvent(nullptr, b->own_out, 0));
        }
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftBatchDescribe(void *handle, char *buf, size_t cap, size_t *needed)
{
    return sip::guarded([&]() -> int32_t {
        BatchPlan *b = sip::lookup_batch(handle);
        if (!b) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(b->mu);
        std::ostringstream os;
        os << "{\"problems\":" << b->lengths.size() << ",\"groups\":" << b->groups.size() << ",\"chains\":" << b->nchains
           << ",\"graph\":" << (b->use_graph ? 1 : 0) << ",\"workspace\":" << b->ws_bytes;
        if (b->persist) {
            os << ",\"persist_chain\":" << b->persist_chain << ",\"persist_tuned_us\":[";
            for (size_t i = 0; i < b->persist_tuned_us.size(); i++) os << (i ? "," : "") << b->persist_tuned_us[i];
            os << "],\"persist\":" << b->persist->describe();
        }
        os << ",\"group_list\":[";
        for (size_t i = 0; i < b->groups.size(); i++) {
            const sip::Group &g = b->groups[i];
            os << (i ? "," : "") << "{\"n\":" << g.n << ",\"rows\":" << g.rows << ",\"chain\":" << g.chain
               << ",\"stages\":" << g.plan->stages.size() << ",\"row_F\":" << g.plan->tune.row_F
               << ",\"split_lds\":" << g.plan->tune.split_lds << ",\"xcd\":" << g.plan->tune.xcd
               << ",\"split_l1\":" << g.plan->tune.split_l1 << ",\"col_F\":" << g.plan->tune.col_F
               << ",\"rc\":" << g.plan->tune.rc
               << ",\"tuned_us\":" << g.tuned_us << "}";
        }
        os << "]}";
        const std::string s = os.str();
        if (needed) *needed = s.size() + 1;
        if (!buf || cap < s.size() + 1) return E::ACL_ERROR_INVALID_PARAM;
        std::memcpy(buf, s.c_str(), s.size() + 1);
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftBatchDestroy(void *handle)
{
    std::unique_ptr<BatchPlan> victim;
    {
        std::lock_guard<std::mutex> lk(sip::g_batch_mu);
        auto it = sip::g_batches.find(handle);
        if (it == sip::g_batches.end()) return E::ACL_ERROR_INVALID_PARAM;
        victim = std::move(it->second);
        sip::g_batches.erase(it);
    }
    return E::ACL_SUCCESS;
}

}  // extern "C"
