// Minimal leveled logger with the reference's environment switches
// (core/utils/log/log_core.cpp:24-137: ASCEND_GLOBAL_LOG_LEVEL, ASCEND_SLOG_PRINT_TO_STDOUT).
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "common.h"

namespace sip {

static int g_level = -1;
static bool g_to_stdout = false;
static std::once_flag g_once;

static void init_log()
{
    g_level = (int)LogLevel::ERROR;
    if (const char *v = std::getenv("ASCEND_GLOBAL_LOG_LEVEL")) {
        int l = std::atoi(v);
        if (l >= 0 && l <= 4) g_level = l;
    }
    if (const char *v = std::getenv("ASCEND_SLOG_PRINT_TO_STDOUT")) g_to_stdout = std::atoi(v) == 1;
}

bool log_enabled(LogLevel lvl)
{
    std::call_once(g_once, init_log);
    return (int)lvl >= g_level && lvl != LogLevel::NONE;
}

void log_emit(LogLevel lvl, const char *file, int line, const std::string &msg)
{
    static const char *names[] = {"DEBUG", "INFO", "WARN", "ERROR", "NONE"};
    static std::mutex mu;
    if (!g_to_stdout && lvl != LogLevel::ERROR) return;
    std::lock_guard<std::mutex> lk(mu);
    std::fprintf(g_to_stdout ? stdout : stderr, "[sip_fft %s] %s:%d %s\n", names[(int)lvl], file, line, msg.c_str());
}

}  // namespace sip
