/*
 * Plain-C / FFI view of the asdFft* entry points, plus this library's extensions.
 *
 * The asdFft* functions below are the same unmangled symbols as include/fft_api.h (the reference's
 * include/fft_api.h:48-96); where the C++ header takes a reference (asdFftHandle &, size_t &,
 * asdFftType &) this view takes a pointer, which is the same machine-level argument. Include one
 * of the two headers in a translation unit, not both: in C++ the extension block below is also
 * available after including fft_api.h (it is guarded for that case).
 *
 * Enum values (include/fft_api.h:26-44 in the reference):
 *   type      C2C = 0x10, C2R = 0x11, R2C = 0x12, STFT_* = 0x20-0x22, C2C_SEP = 0x30
 *   direction FORWARD = 0x10, INVERSE = 0x11
 *   dimType   HORIZONTAL = 0x10, VERTICAL = 0x11
 * Status codes: 0 success, 100000 invalid parameter, 200001 not supported, 500000 internal error.
 */
#ifndef SIP_FFT_H
#define SIP_FFT_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
#include "fft_api.h"
extern "C" {
#else
typedef struct aclTensor aclTensor;
typedef void *asdFftHandle;
typedef int32_t AspbStatus;

enum { ASCEND_FFT_C2C = 0x10, ASCEND_FFT_C2R = 0x11, ASCEND_FFT_R2C = 0x12, ASCEND_STFT_C2C = 0x20,
       ASCEND_STFT_C2R = 0x21, ASCEND_STFT_R2C = 0x22, ASCEND_FFT_C2C_SEP = 0x30 };
enum { ASCEND_FFT_FORWARD = 0x10, ASCEND_FFT_INVERSE = 0x11 };
enum { ASCEND_FFT_HORIZONTAL = 0x10, ASCEND_FFT_VERTICAL = 0x11 };

AspbStatus asdFftCreate(asdFftHandle *handle);
AspbStatus asdFftSetStream(asdFftHandle handle, void *stream);
AspbStatus asdFftMakePlan1D(asdFftHandle handle, int64_t fftSize, int32_t fftType, int32_t direction,
                            int64_t batchSize, int32_t dimType);
AspbStatus asdFftMakePlan2D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, int32_t fftType,
                            int32_t direction, int32_t batchSize);
AspbStatus asdFftMakePlan3D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, int64_t fftSizeZ,
                            int32_t fftType, int32_t direction, int32_t batchSize);
AspbStatus asdFftGetWorkspaceSize(asdFftHandle handle, size_t *workspaceSize);
AspbStatus asdFftSetWorkspace(asdFftHandle handle, void *workspace);
AspbStatus asdFftExecC2C(asdFftHandle handle, const aclTensor *input, const aclTensor *output);
AspbStatus asdFftExecC2R(asdFftHandle handle, const aclTensor *input, const aclTensor *output);
AspbStatus asdFftExecR2C(asdFftHandle handle, const aclTensor *input, const aclTensor *output);
AspbStatus asdFftExecC2CSeparated(asdFftHandle handle, const aclTensor *inputReal, const aclTensor *inputImag,
                                  const aclTensor *outputReal, const aclTensor *outputImag);
AspbStatus asdFftIstftMakePlan(asdFftHandle handle, const aclTensor *input, int64_t nFft, int64_t hopLengthOpt,
                               int64_t winLengthOpt, bool center, bool normalized, bool onesidedOpt,
                               int64_t lengthOpt, bool returnComplex);
AspbStatus asdFftExecIstft(asdFftHandle handle, const aclTensor *input, const aclTensor *windowOpt,
                           const aclTensor *output);
AspbStatus asdFftSynchronize(asdFftHandle handle);
AspbStatus asdFftGetType(asdFftHandle handle, int32_t *fftType);
AspbStatus asdFftDestroy(asdFftHandle handle);
#endif

/* ------------------------------------------------------------------ extensions (no counterpart
 * in the reference; used by the tests, the benchmark and the batching / serving layers) */

/* Writes a JSON description of a handle's plan (stages, kernels, radices, workspace) into buf. Returns 0, 100000 on a bad handle, or 100000 if
 * cap is too small (then *needed holds the size including the terminating NUL). */
int32_t sipFftPlanDescribe(void *handle, char *buf, size_t cap, size_t *needed);

/* Same description for a plan that is only laid out on the host (no device memory, no GPU
 * needed): ndim 1..3, dims[ndim], type/direction/dimType as for asdFftMakePlan*. */
int32_t sipFftPlanDescribeDry(int32_t ndim, const int64_t *dims, int32_t fftType, int32_t direction,
                              int64_t batchSize, int32_t dimType, char *buf, size_t cap, size_t *needed);

/* Host-generated content of one coefficient table of a plan laid out on the host (no GPU needed),
 * as interleaved (re, im) floats: stage = index into the plan's stages (as described by
 * sipFftPlanDescribeDry); which = 0 pass twiddles, 1 real-transform roots, 2 / 3 four-step
 * twiddles (low / high level), 4 auxiliary table (Bluestein chirp or kernel spectrum), 5 column pass
 * twiddles of a fused 2D tile. *count receives the number of complex entries; cap is the capacity
 * of out in floats (at least 2 * count). Returns 100000 for a table the stage does not have. */
int32_t sipFftPlanTableDry(int32_t ndim, const int64_t *dims, int32_t fftType, int32_t direction, int64_t batchSize,
                           int32_t dimType, int32_t stage, int32_t which, float *out, size_t cap, size_t *count);

/* Device-resident coefficient cache: map entries (live or released), live tables and their bytes.
 * A table is shared by every live plan that uses it and freed with the last of them. */
int32_t sipFftCoeffCacheStats(size_t *entries, size_t *live_tables, size_t *live_bytes);

/* asdFftExecC2C without a caller workspace. No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). Same checks and status codes as asdFftExecC2C; the plan's workspace comes from a
 * per-device scratch array of 64 MiB held by the library (looked up on the first call on a device,
 * kept until the process ends), so asdFftSetWorkspace is not needed, and a plan whose workspace
 * (asdFftGetWorkspaceSize) exceeds 64 MiB is refused with 100000. The call allocates nothing. All
 * calls on one GPU share that GPU's array: calls that overlap on one GPU race. */
int32_t sipFftExecC2CNoAlloc(void *handle, const aclTensor *input, const aclTensor *output);

/* Library version string. */
const char *sipFftVersion(void);

/* Precision form: run the radix-16 stages of C2C launches on the matrix cores with split operands
 * and FP32 accumulation. mode 0 = plain FP32 (default), 1 = FP16 hi+lo (per-column power-of-two
 * scaling), 2 = BF16 hi+lo, 3 = BF16 hi+mid+lo. Applies to leaf lengths 128, 256, 512, 1024, 2048
 * and 4096 (other launches stay plain). Call before or after MakePlan (the plan is rebuilt); the
 * workspace size does not change. */
int32_t sipFftSetPrecision(void *handle, int32_t mode);

/* Batching form: `count` C2C problems of uneven lengths run as one launch. Problem i is
 * batches[i] transforms of length lengths[i], contiguous complex64 at device address in_ptrs[i]
 * (8-byte aligned), written to out_ptrs[i]. The buffers are bound at plan time.
 * flags: bit 0 = capture the whole batch into one hipGraph (each exec is one graph launch);
 *        bit 1 = autotune tile shapes per length group (decisions cached in $SIP_FFT_TUNE_CACHE,
 *                default $HOME/.sip_fft_tune.jsonl);
 *        bit 2 = run every 2,3,5-smooth length up to 4096 in one persistent multi-length launch
 *                (with bit 1 the radix order of each such length is tuned).
 * Exec only enqueues on the bound stream. Workspace: query after MakePlan, caller-allocated. */
int32_t sipFftBatchCreate(void **handle);
int32_t sipFftBatchSetStream(void *handle, void *stream);
int32_t sipFftBatchMakePlan(void *handle, int32_t count, const int64_t *lengths, const int64_t *batches,
                            void *const *in_ptrs, void *const *out_ptrs, int32_t direction, int32_t flags);
int32_t sipFftBatchGetWorkspaceSize(void *handle, size_t *bytes);
int32_t sipFftBatchSetWorkspace(void *handle, void *workspace);
int32_t sipFftBatchExec(void *handle);
int32_t sipFftBatchDescribe(void *handle, char *buf, size_t cap, size_t *needed);
int32_t sipFftBatchDestroy(void *handle);

/* Serving form: a continuous-batching request loop (one scheduler thread per server).
 * Create: budget_bytes of device memory are reserved once as the admission arena (outputs owned by
 * the server, workspaces, row tables); nstreams streams carry up to two launches each; queued
 * requests of the same (type, n) are coalesced into one launch of at most max_coalesce_rows rows.
 * Submit: fftType C2C (0x10, complex64 in/out) or R2C (0x12, float in, n/2+1 bins out), forward;
 * input is a device address; output a device address, or NULL to let the server allocate it from the
 * arena (the result is then discarded after completion). SubmitMany hands over `count` requests
 * at once (arrays of the Submit arguments; outputs may be NULL, or hold NULL entries): all are
 * validated before any is queued (on a refusal nothing is queued and *refused, when given, is the
 * index of the refused request), then they enter the queue together, so those of one shape can
 * coalesce; tickets (may be NULL) receives one ticket per request. Wait blocks until the request finished and
 * returns its status and latency (arrival to observed completion). Stats returns JSON
 * (p50/p90/p99 latency, launches, coalescing, arena use, resident coefficient tables and their
 * evictions: the tables are kept on the device outside the arena, up to a sixteenth of the budget
 * or at least 64 MiB, least recently used dropped first); with reset = 1 the counters are cleared
 * only when the report fits `cap` (a size query with buf = NULL clears nothing). Benchmark runs an
 * open-loop Poisson stream (see serving.cpp) once per call and writes its JSON report into buf: pass
 * a buffer of 64 KiB (a short buffer loses the report and returns 100000 with *needed set). Calls
 * that overlap on one server are serialised internally. */
int32_t sipFftServeCreate(void **handle, size_t budget_bytes, int32_t nstreams, int64_t max_coalesce_rows);
int32_t sipFftServeSubmit(void *handle, int32_t fftType, int64_t n, int64_t batch, const void *input, void *output,
                          uint64_t *ticket);
int32_t sipFftServeSubmitMany(void *handle, int64_t count, const int32_t *fftTypes, const int64_t *n,
                              const int64_t *batch, const void *const *inputs, void *const *outputs,
                              uint64_t *tickets, int64_t *refused);
int32_t sipFftServeWait(void *handle, uint64_t ticket, double *latency_ms);
int32_t sipFftServeDrain(void *handle);
int32_t sipFftServeStats(void *handle, int32_t reset, char *buf, size_t cap, size_t *needed);
int32_t sipFftServeBenchmark(void *handle, double rate, double seconds, uint64_t seed, const void *pool,
                             size_t pool_bytes, double c2c_frac, int64_t n_lo, int64_t n_hi, int64_t b_max,
                             double drain_seconds, char *buf, size_t cap, size_t *needed);
int32_t sipFftServeDestroy(void *handle);

#ifdef __cplusplus
}
#endif

#endif
