// FEFF file handoffs of the FMS and MKGTR module stages (work-directory seam).
#include <cstring>
#include <string>

#include "fms_common.h"

namespace {
void put_message(char* dst, size_t len, const std::string& msg) {
    if (!dst || len == 0) return;
    const size_t n = std::min(len - 1, msg.size());
    std::memcpy(dst, msg.data(), n);
    dst[n] = '\0';
}
}  // namespace

extern "C" {

int refeff_fms_run_in_dir(const char* work_dir, int32_t solver, int32_t* files_written, char* message,
                          size_t message_len) {
    (void)work_dir;
    (void)solver;
    if (files_written) *files_written = 0;
    put_message(message, message_len, "module stage not available in this build");
    return REFEFF_FMS_UNSUPPORTED;
}

int refeff_mkgtr_run_in_dir(const char* work_dir, int32_t* files_written, char* message, size_t message_len) {
    (void)work_dir;
    if (files_written) *files_written = 0;
    put_message(message, message_len, "module stage not available in this build");
    return REFEFF_FMS_UNSUPPORTED;
}

}  // extern "C"
