// C-ABI entry points (asdFft*): handle registry, validation, tensor matching, plan build and exec.
//
// Mirrors core/fft/fft_api.cpp of the reference: the same checks in the same order with the same
// status codes (invalid handle / size / batch -> 100000, Synchronize and GetType on a bad handle
// -> 500000, missing workspace or storage -> 500000; fft_api.cpp:737-1233). Differences, all on the
// safe side: no C++ exception crosses the boundary (plan-build failures return a status instead of
// escaping, fft_api.cpp:308-320); the registry and each plan have their own locks instead of one
// global mutex (core/utils/sip_lock.cpp:13-14); tensors must be contiguous and their element count
// must match the plan (the reference checks only the trailing dims, fft_api.cpp:1047-1167, and
// would read out of bounds otherwise); exec on a plan of another type is rejected.
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "aclnn/acl_meta.h"
#include "plan.h"
#include "scratch.h"
#include "tensor.h"

using sip::Plan;
using sip::SipError;
namespace E = AsdSip;

namespace {

std::mutex g_reg_mu;
std::unordered_map<void *, std::unique_ptr<Plan>> g_plans;

Plan *lookup(void *h)
{
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_plans.find(h);
    return it == g_plans.end() ? nullptr : it->second.get();
}

template <class F>
int32_t guarded(F &&f)
{
    try {
        return f();
    } catch (const SipError &e) {
        SIP_LOG(ERROR, e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        SIP_LOG(ERROR, "out of host memory");
        return E::ACL_ERROR_INTERNAL_ERROR;
    } catch (const std::exception &e) {
        SIP_LOG(ERROR, e.what());
        return E::ACL_ERROR_INTERNAL_ERROR;
    } catch (...) {
        return E::ACL_ERROR_INTERNAL_ERROR;
    }
}

bool valid_type(int t) { return t == sip::kTypeC2C || t == sip::kTypeC2R || t == sip::kTypeR2C || t == sip::kTypeC2CSep; }
bool valid_dir(int d) { return d == sip::kDirForward || d == sip::kDirInverse; }

int32_t finish_plan(Plan &p)
{
    p.initialized = false;
    p.failed = false;
    p.stages.clear();
    try {
        sip::plan_build(p);
        sip::plan_materialize(p);
    } catch (...) {
        p.failed = true;
        p.stages.clear();
        throw;
    }
    p.initialized = true;
    return E::ACL_SUCCESS;
}

int64_t elems(const aclTensor *t)
{
    int64_t n = 1;
    for (int i = 0; i < t->ndim; i++) n *= t->dims[i];
    return n;
}

bool contiguous(const aclTensor *t)
{
    int64_t acc = 1;
    for (int i = t->ndim - 1; i >= 0; i--) {
        if (t->dims[i] != 1 && t->strides[i] != acc) return false;
        acc *= t->dims[i];
    }
    return true;
}

// Trailing-dimension check of matchC2C_ / matchR2C_ / matchC2R_ (fft_api.cpp:1047-1167) plus the
// element-count and contiguity checks this library adds. `half_last` selects the N/2+1 rule.
bool match(const Plan &p, const aclTensor *t, int dtype, bool half_last)
{
    if (t == nullptr || t->dtype != dtype || !contiguous(t)) return false;
    const int nd = (int)p.sizes.size();
    const bool vertical = nd == 1 && p.dimType == sip::kDimVertical && p.fftType == sip::kTypeC2C;
    int64_t per = 1;
    for (int i = 0; i < nd; i++) per *= (half_last && i == nd - 1) ? p.sizes[i] / 2 + 1 : p.sizes[i];
    if (elems(t) != per * p.batch) return false;
    if (vertical) return true;  // the reference skips the shape check for strided plans
    if (t->ndim < nd) return false;
    for (int i = 0; i < nd; i++) {
        int64_t want = p.sizes[nd - 1 - i];
        if (half_last && i == 0) want = want / 2 + 1;
        if (t->dims[t->ndim - 1 - i] != want) return false;
    }
    return true;
}

int32_t exec_common(Plan &p, const aclTensor *in, const aclTensor *out)
{
    if (!p.initialized) return E::ACL_ERROR_INVALID_PARAM;
    if (p.ws_bytes > 0 && p.workspace == nullptr) return E::ACL_ERROR_INTERNAL_ERROR;
    const void *src = aclGetTensorData(in);
    void *dst = aclGetTensorData(out);
    if (src == nullptr || dst == nullptr) return E::ACL_ERROR_INTERNAL_ERROR;
    int cur = 0;
    if (hipGetDevice(&cur) == hipSuccess && cur != p.device) SIP_HIP_CHECK(hipSetDevice(p.device));
    sip::plan_execute(p, src, dst);
    return E::ACL_SUCCESS;
}

}  // namespace

extern "C" {

int32_t asdFftCreate(void **handle)
{
    return guarded([&]() -> int32_t {
        if (handle == nullptr) return E::ACL_ERROR_INVALID_PARAM;
        auto p = std::make_unique<Plan>();
        void *h = p.get();
        std::lock_guard<std::mutex> lk(g_reg_mu);
        g_plans.emplace(h, std::move(p));
        *handle = h;
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftSetStream(void *handle, void *stream)
{
    return guarded([&]() -> int32_t {
        if (stream == nullptr) return E::ACL_ERROR_INVALID_PARAM;
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        p->stream = static_cast<hipStream_t>(stream);
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftMakePlan1D(void *handle, int64_t fftSize, int32_t fftType, int32_t direction, int64_t batchSize,
                         int32_t dimType)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        if (fftSize <= 0 || fftSize > sip::kMaxFftSize) return E::ACL_ERROR_INVALID_PARAM;
        if (batchSize <= 0) return E::ACL_ERROR_INVALID_PARAM;
        if (!valid_type(fftType) || !valid_dir(direction)) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        p->istft = Plan::Istft{};
        p->fftType = fftType;
        p->direction = direction;
        p->sizes = {fftSize};
        p->dimType = (dimType == sip::kDimVertical) ? sip::kDimVertical : sip::kDimHorizontal;
        p->batch = batchSize;
        p->strides = {p->dimType == sip::kDimVertical ? batchSize : 1};
        return finish_plan(*p);
    });
}

static int32_t common_check(void *handle, int64_t x, int64_t y, int64_t batch)
{
    if (!lookup(handle)) return E::ACL_ERROR_INVALID_PARAM;
    if (x <= 0 || x > sip::kMaxFftSize) return E::ACL_ERROR_INVALID_PARAM;
    if (y <= 0 || y > sip::kMaxFftSize) return E::ACL_ERROR_INVALID_PARAM;
    if (batch <= 0) return E::ACL_ERROR_INVALID_PARAM;
    return E::ACL_SUCCESS;
}

int32_t asdFftMakePlan2D(void *handle, int64_t fftSizeX, int64_t fftSizeY, int32_t fftType, int32_t direction,
                         int32_t batchSize)
{
    return guarded([&]() -> int32_t {
        int32_t st = common_check(handle, fftSizeX, fftSizeY, batchSize);
        if (st != E::ACL_SUCCESS) return st;
        if (!valid_type(fftType) || !valid_dir(direction)) return E::ACL_ERROR_INVALID_PARAM;
        Plan *p = lookup(handle);
        std::lock_guard<std::mutex> lk(p->mu);
        p->istft = Plan::Istft{};
        p->fftType = fftType;
        p->direction = direction;
        p->batch = batchSize;
        p->dimType = sip::kDimHorizontal;
        p->sizes = {fftSizeX, fftSizeY};
        p->strides = {1, 1};
        return finish_plan(*p);
    });
}

int32_t asdFftMakePlan3D(void *handle, int64_t fftSizeX, int64_t fftSizeY, int64_t fftSizeZ, int32_t fftType,
                         int32_t direction, int32_t batchSize)
{
    return guarded([&]() -> int32_t {
        int32_t st = common_check(handle, fftSizeX, fftSizeY, batchSize);
        if (st != E::ACL_SUCCESS) return st;
        if (fftSizeZ <= 0 || fftSizeZ > sip::kMaxFftSize) return E::ACL_ERROR_INVALID_PARAM;
        if (!valid_type(fftType) || !valid_dir(direction)) return E::ACL_ERROR_INVALID_PARAM;
        Plan *p = lookup(handle);
        std::lock_guard<std::mutex> lk(p->mu);
        p->istft = Plan::Istft{};
        p->fftType = fftType;
        p->direction = direction;
        p->batch = batchSize;
        p->dimType = sip::kDimHorizontal;
        p->sizes = {fftSizeX, fftSizeY, fftSizeZ};
        p->strides = {1, 1, 1};
        return finish_plan(*p);
    });
}

int32_t asdFftGetWorkspaceSize(void *handle, size_t *workspaceSize)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p || workspaceSize == nullptr) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        *workspaceSize = p->ws_bytes;
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftSetWorkspace(void *handle, void *workspace)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        p->workspace = workspace;
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftExecC2C(void *handle, const aclTensor *input, const aclTensor *output)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        if (p->fftType != sip::kTypeC2C || p->istft.active) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, input, ACL_COMPLEX64, false)) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, output, ACL_COMPLEX64, false)) return E::ACL_ERROR_INVALID_PARAM;
        return exec_common(*p, input, output);
    });
}

int32_t asdFftExecC2R(void *handle, const aclTensor *input, const aclTensor *output)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        if (p->fftType != sip::kTypeC2R) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, input, ACL_COMPLEX64, true)) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, output, ACL_FLOAT, false)) return E::ACL_ERROR_INVALID_PARAM;
        return exec_common(*p, input, output);
    });
}

int32_t asdFftExecR2C(void *handle, const aclTensor *input, const aclTensor *output)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        if (p->fftType != sip::kTypeR2C) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, input, ACL_FLOAT, false)) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, output, ACL_COMPLEX64, true)) return E::ACL_ERROR_INVALID_PARAM;
        return exec_common(*p, input, output);
    });
}

// asdFftExecC2C without a caller workspace, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. The
// plan's workspace (four-step intermediate, Bluestein rows) comes from the library's per-device
// scratch array of sip::kScratchBytes (64 MiB), whose address is looked up on the first call on a
// device; a plan needing more is refused with 100000. Nothing is allocated, and no memory is
// cleared: every stage writes all of the workspace it later reads. Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (other streams, other threads) race.
int32_t sipFftExecC2CNoAlloc(void *handle, const aclTensor *input, const aclTensor *output)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        if (p->fftType != sip::kTypeC2C || p->istft.active) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, input, ACL_COMPLEX64, false)) return E::ACL_ERROR_INVALID_PARAM;
        if (!match(*p, output, ACL_COMPLEX64, false)) return E::ACL_ERROR_INVALID_PARAM;
        if (!p->initialized) return E::ACL_ERROR_INVALID_PARAM;
        if (p->ws_bytes > sip::kScratchBytes) {
            SIP_LOG(ERROR, "sipFftExecC2CNoAlloc: the plan needs " << p->ws_bytes << " bytes of workspace, more than "
                                                                   << "the " << sip::kScratchBytes
                                                                   << "-byte scratch array");
            return E::ACL_ERROR_INVALID_PARAM;
        }
        const void *src = aclGetTensorData(input);
        void *dst = aclGetTensorData(output);
        if (src == nullptr || dst == nullptr) return E::ACL_ERROR_INTERNAL_ERROR;
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != p->device) SIP_HIP_CHECK(hipSetDevice(p->device));
        void *const saved = p->workspace;
        p->workspace = p->ws_bytes > 0 ? sip::scratch_device_ptr(p->device) : nullptr;
        try {
            sip::plan_execute(*p, src, dst);
        } catch (...) {
            p->workspace = saved;
            throw;
        }
        p->workspace = saved;
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftExecC2CSeparated(void *handle, const aclTensor *inputReal, const aclTensor *inputImag,
                               const aclTensor *outputReal, const aclTensor *outputImag)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        if (!inputReal || !inputImag || !outputReal || !outputImag) return E::ACL_ERROR_INTERNAL_ERROR;
        // Same shape rule as the reference (fft_api.cpp:1283-1342): the four planes have one shape.
        const aclTensor *ts[4] = {inputReal, inputImag, outputReal, outputImag};
        for (int i = 1; i < 4; i++) {
            if (ts[i]->ndim != ts[0]->ndim) return E::ACL_ERROR_INTERNAL_ERROR;
            for (int d = 0; d < ts[0]->ndim; d++)
                if (ts[i]->dims[d] != ts[0]->dims[d]) return E::ACL_ERROR_INTERNAL_ERROR;
        }
        std::lock_guard<std::mutex> lk(p->mu);
        // the reference runs any plan here; only a C2C_SEP plan has split-plane stages
        if (p->fftType != sip::kTypeC2CSep) return E::ACL_ERROR_INVALID_PARAM;
        for (const aclTensor *t : ts)
            if (!match(*p, t, ACL_FLOAT, false)) return E::ACL_ERROR_INVALID_PARAM;
        if (!p->initialized) return E::ACL_ERROR_INVALID_PARAM;
        if (p->ws_bytes > 0 && p->workspace == nullptr) return E::ACL_ERROR_INTERNAL_ERROR;
        void *d[4];
        for (int i = 0; i < 4; i++) {
            d[i] = aclGetTensorData(ts[i]);
            if (d[i] == nullptr) return E::ACL_ERROR_INTERNAL_ERROR;
        }
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != p->device) SIP_HIP_CHECK(hipSetDevice(p->device));
        sip::plan_execute_sep(*p, d[0], d[1], d[2], d[3]);
        return E::ACL_SUCCESS;
    });
}

// ISTFT (core/fft/istft.cpp:305-575). The reference runs one parameter set only and forces it,
// with a warning, whatever the caller passes: center = true, normalized = false, onesided = false,
// return_complex = true, length = 0 (istft.cpp:330-370); this library does the same, so a caller
// gets the reference's result. Checks and status codes follow IstftParamsCheck and
// asdFftExecIstft: input complex64 (channel, n_fft, frames) with frames > 1, 0 < hop <= win_length,
// win_length == n_fft, n_fft < 1500, hop < 1500 (100000 otherwise); window float (n_fft);
// output complex64 (channel, n_fft + hop (frames - 1) - 2 floor(n_fft / 2)).
int32_t asdFftIstftMakePlan(void *handle, const aclTensor *input, int64_t nFft, int64_t hopLengthOpt,
                            int64_t winLengthOpt, bool center, bool normalized, bool onesidedOpt, int64_t lengthOpt,
                            bool returnComplex)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        if (input == nullptr || input->dtype != ACL_COMPLEX64 || input->ndim != 3) return E::ACL_ERROR_INVALID_PARAM;
        const int64_t channel = input->dims[0], fftSize = input->dims[1], frames = input->dims[2];
        if (!center) SIP_LOG(WARN, "istft: only center = true is supported; using it");
        if (normalized) SIP_LOG(WARN, "istft: only normalized = false is supported; using it");
        if (onesidedOpt) SIP_LOG(WARN, "istft: only onesided = false is supported; using it");
        if (!returnComplex) SIP_LOG(WARN, "istft: only return_complex = true is supported; using it");
        if (lengthOpt != 0) SIP_LOG(WARN, "istft: only length = 0 is supported; using it");
        if (frames <= 1) return E::ACL_ERROR_INVALID_PARAM;
        if (nFft != fftSize) return E::ACL_ERROR_INVALID_PARAM;
        if (!(hopLengthOpt > 0 && hopLengthOpt <= winLengthOpt)) return E::ACL_ERROR_INVALID_PARAM;
        if (!(winLengthOpt > 0 && winLengthOpt == nFft)) return E::ACL_ERROR_INVALID_PARAM;
        if (nFft >= 1500 || hopLengthOpt >= 1500) return E::ACL_ERROR_INVALID_PARAM;
        if (channel <= 0) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        p->fftType = sip::kTypeC2C;
        p->direction = sip::kDirInverse;
        p->dimType = sip::kDimHorizontal;
        p->sizes = {nFft};
        p->strides = {1};
        p->batch = channel * frames;
        p->istft.active = true;
        p->istft.channel = channel;
        p->istft.nfft = nFft;
        p->istft.frames = frames;
        p->istft.hop = hopLengthOpt;
        p->istft.win_len = winLengthOpt;
        p->istft.out_len = nFft + hopLengthOpt * (frames - 1) - 2 * (nFft / 2);
        try {
            return finish_plan(*p);
        } catch (const SipError &e) {
            // the reference's MakePlan1DFft failure: the handle keeps no ISTFT description
            SIP_LOG(ERROR, "istft plan: " << e.what());
            p->istft = Plan::Istft{};
            return E::ACL_ERROR_INTERNAL_ERROR;
        }
    });
}

int32_t asdFftExecIstft(void *handle, const aclTensor *input, const aclTensor *windowOpt, const aclTensor *output)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        const Plan::Istft &q = p->istft;
        if (!q.active) return E::ACL_ERROR_INVALID_PARAM;  // the plan holds no ISTFT description
        if (input == nullptr || input->dtype != ACL_COMPLEX64 || input->ndim != 3 || !contiguous(input))
            return E::ACL_ERROR_INVALID_PARAM;
        if (input->dims[0] != q.channel || input->dims[1] != q.nfft || input->dims[2] != q.frames)
            return E::ACL_ERROR_INVALID_PARAM;
        if (output == nullptr || output->dtype != ACL_COMPLEX64 || output->ndim != 2 || !contiguous(output))
            return E::ACL_ERROR_INVALID_PARAM;
        if (output->dims[0] != q.channel || output->dims[1] != q.out_len) return E::ACL_ERROR_INVALID_PARAM;
        if (windowOpt == nullptr || windowOpt->dtype != ACL_FLOAT || windowOpt->ndim != 1 || !contiguous(windowOpt))
            return E::ACL_ERROR_INVALID_PARAM;
        if (windowOpt->dims[0] != q.win_len || windowOpt->dims[0] != q.nfft) return E::ACL_ERROR_INVALID_PARAM;
        if (!p->initialized) return E::ACL_ERROR_INVALID_PARAM;
        if (p->ws_bytes > 0 && p->workspace == nullptr) return E::ACL_ERROR_INTERNAL_ERROR;
        const void *src = aclGetTensorData(input);
        const void *win = aclGetTensorData(windowOpt);
        void *dst = aclGetTensorData(output);
        if (src == nullptr || win == nullptr || dst == nullptr) return E::ACL_ERROR_INTERNAL_ERROR;
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && cur != p->device) SIP_HIP_CHECK(hipSetDevice(p->device));
        sip::plan_execute_istft(*p, src, static_cast<const float *>(win), dst);
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftSynchronize(void *handle)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INTERNAL_ERROR;
        hipStream_t s;
        {
            std::lock_guard<std::mutex> lk(p->mu);
            s = p->stream;
        }
        SIP_HIP_CHECK(hipStreamSynchronize(s));
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftGetType(void *handle, int32_t *fftType)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INTERNAL_ERROR;
        if (!fftType) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        *fftType = p->fftType;
        return E::ACL_SUCCESS;
    });
}

int32_t asdFftDestroy(void *handle)
{
    return guarded([&]() -> int32_t {
        std::unique_ptr<Plan> victim;
        {
            std::lock_guard<std::mutex> lk(g_reg_mu);
            auto it = g_plans.find(handle);
            if (it == g_plans.end()) return E::ACL_ERROR_INVALID_PARAM;
            victim = std::move(it->second);
            g_plans.erase(it);
        }
        { std::lock_guard<std::mutex> lk(victim->mu); }
        victim.reset();
        return E::ACL_SUCCESS;
    });
}

// ------------------------------------------------------------------ extensions

static int32_t write_str(const std::string &s, char *buf, size_t cap, size_t *needed)
{
    if (needed) *needed = s.size() + 1;
    if (buf == nullptr || cap < s.size() + 1) return E::ACL_ERROR_INVALID_PARAM;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return E::ACL_SUCCESS;
}

int32_t sipFftPlanDescribe(void *handle, char *buf, size_t cap, size_t *needed)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        return write_str(sip::plan_describe(*p), buf, cap, needed);
    });
}

int32_t sipFftPlanDescribeDry(int32_t ndim, const int64_t *dims, int32_t fftType, int32_t direction,
                              int64_t batchSize, int32_t dimType, char *buf, size_t cap, size_t *needed)
{
    return guarded([&]() -> int32_t {
        if (ndim < 1 || ndim > 3 || dims == nullptr || batchSize <= 0) return E::ACL_ERROR_INVALID_PARAM;
        if (!valid_type(fftType) || !valid_dir(direction)) return E::ACL_ERROR_INVALID_PARAM;
        Plan p;
        p.fftType = fftType;
        p.direction = direction;
        p.batch = batchSize;
        p.dimType = (ndim == 1 && dimType == sip::kDimVertical) ? sip::kDimVertical : sip::kDimHorizontal;
        for (int i = 0; i < ndim; i++) {
            if (dims[i] <= 0 || dims[i] > sip::kMaxFftSize) return E::ACL_ERROR_INVALID_PARAM;
            p.sizes.push_back(dims[i]);
        }
        sip::plan_build(p);
        sip::plan_mark_streaming(p);
        return write_str(sip::plan_describe(p), buf, cap, needed);
    });
}

int32_t sipFftPlanTableDry(int32_t ndim, const int64_t *dims, int32_t fftType, int32_t direction, int64_t batchSize,
                           int32_t dimType, int32_t stage, int32_t which, float *out, size_t cap, size_t *count)
{
    return guarded([&]() -> int32_t {
        if (ndim < 1 || ndim > 3 || dims == nullptr || batchSize <= 0) return E::ACL_ERROR_INVALID_PARAM;
        if (!valid_type(fftType) || !valid_dir(direction)) return E::ACL_ERROR_INVALID_PARAM;
        Plan p;
        p.fftType = fftType;
        p.direction = direction;
        p.batch = batchSize;
        p.dimType = (ndim == 1 && dimType == sip::kDimVertical) ? sip::kDimVertical : sip::kDimHorizontal;
        for (int i = 0; i < ndim; i++) {
            if (dims[i] <= 0 || dims[i] > sip::kMaxFftSize) return E::ACL_ERROR_INVALID_PARAM;
            p.sizes.push_back(dims[i]);
        }
        sip::plan_build(p);
        if (stage < 0 || stage >= (int32_t)p.stages.size()) return E::ACL_ERROR_INVALID_PARAM;
        const sip::Stage &st = p.stages[(size_t)stage];
        const sip::TableRef *t = nullptr;
        switch (which) {
            case 0: t = &st.tw; break;
            case 1: t = &st.rtw; break;
            case 2: t = &st.tw4lo; break;
            case 3: t = &st.tw4hi; break;
            case 4: t = &st.tab0; break;
            case 5: t = &st.tw2; break;
            default: return E::ACL_ERROR_INVALID_PARAM;
        }
        if (!t->used()) return E::ACL_ERROR_INVALID_PARAM;
        const std::vector<float2> v = t->gen();
        if (count) *count = v.size();
        if (out == nullptr || cap < 2 * v.size()) return E::ACL_ERROR_INVALID_PARAM;
        std::memcpy(out, v.data(), v.size() * sizeof(float2));
        return E::ACL_SUCCESS;
    });
}

const char *sipFftVersion(void) { return "sip_fft 0.2 (gfx950)"; }

int32_t sipFftSetPrecision(void *handle, int32_t mode)
{
    return guarded([&]() -> int32_t {
        Plan *p = lookup(handle);
        if (!p || mode < sip::PREC_FP32 || mode > sip::PREC_BF16_SPLIT3) return E::ACL_ERROR_INVALID_PARAM;
        std::lock_guard<std::mutex> lk(p->mu);
        p->precision = mode;
        if (p->initialized) return finish_plan(*p);  // re-plan with the new stage kernels
        return E::ACL_SUCCESS;
    });
}

}  // extern "C"
