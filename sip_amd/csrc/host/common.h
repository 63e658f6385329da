// Shared host-side helpers: status codes, logging, HIP error checks, exact fast division.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>

#include "utils/aspb_status.h"
#include "../kernels/stockham_args.h"

namespace sip {

using AsdSip::AspbStatus;
using AsdSip::ErrorType;

// ------------------------------------------------------------------ logging
// Levels and environment follow the reference's logger (core/utils/log/log_core.cpp:24-137):
// ASCEND_GLOBAL_LOG_LEVEL 0 debug, 1 info, 2 warning, 3 error (default), 4 none; messages go to
// stderr when ASCEND_SLOG_PRINT_TO_STDOUT=1 or when the level is error.
enum class LogLevel : int { DEBUG = 0, INFO = 1, WARN = 2, ERROR = 3, NONE = 4 };
bool log_enabled(LogLevel lvl);
void log_emit(LogLevel lvl, const char *file, int line, const std::string &msg);

#define SIP_LOG(LVL, MSG)                                                            \
    do {                                                                             \
        if (::sip::log_enabled(::sip::LogLevel::LVL)) {                              \
            std::ostringstream sip_log_os_;                                          \
            sip_log_os_ << MSG;                                                      \
            ::sip::log_emit(::sip::LogLevel::LVL, __FILE__, __LINE__, sip_log_os_.str()); \
        }                                                                            \
    } while (0)

// Thrown inside the library and mapped to a status code at the C boundary; never crosses it.
struct SipError : std::runtime_error {
    AspbStatus status;
    SipError(AspbStatus s, const std::string &m) : std::runtime_error(m), status(s) {}
};

// A failing HIP call becomes a SipError (mapped to a status at the C-ABI). When the runtime's
// last-error slot holds exactly the error this call just returned, it is read off, so a failure
// already reported through a status does not resurface in the caller's next HIP call; any other
// pending error is left for its owner.
#define SIP_HIP_CHECK(EXPR)                                                                        \
    do {                                                                                           \
        hipError_t sip_e_ = (EXPR);                                                                \
        if (sip_e_ != hipSuccess) {                                                                \
            if (hipPeekAtLastError() == sip_e_) (void)hipGetLastError();                           \
            throw ::sip::SipError(::AsdSip::ACL_ERROR_INTERNAL_ERROR,                              \
                                  std::string(#EXPR) + " failed: " + hipGetErrorString(sip_e_)); \
        }                                                                                          \
    } while (0)

// ------------------------------------------------------------------ exact division by a constant
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1; exact for
// every n < 2^31 (Granlund & Montgomery, "Division by invariant integers using multiplication").
inline FastDiv make_fastdiv(uint32_t d)
{
    if (d == 0) d = 1;
    uint32_t s = 0;
    while ((uint64_t(1) << s) < d) s++;
    const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
    FastDiv f;
    f.d = d;
    f.m = (uint32_t)m;
    f.s = s;
    f.pad_ = 0;
    return f;
}

inline uint32_t fastdiv_host(uint32_t n, const FastDiv &f)
{
    const uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
    return (uint32_t)(((uint64_t)hi + n) >> f.s);
}

}  // namespace sip
