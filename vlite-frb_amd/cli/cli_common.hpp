// Shared helpers of the three command-line programs.
#pragma once

#include <sys/time.h>
#include <unistd.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vfrb.h"

namespace cli {

inline bool eq(const char* a, const char* b) { return std::strcmp(a, b) == 0; }

// Threads per block the reference's GPUs launch at most: its block-per-channel
// subtractbpCAL<<<NCHANOUT,NHALFCYC>>> (PIPELINE:323-324,658;
// STATS:248-249,414) fails, and leaves the filterbank as it was, when
// NHALFCYC exceeds it.
constexpr int kRefMaxBlockThreads = 1024;

// Exits with the reference's code 2 (I/O, init) or 1 (usage/plan) after
// printing the library's message.
[[noreturn]] inline void die(int code, const char* what, vfrb_status st) {
    std::fprintf(stderr, "%s: %s (%s)\n", what, vfrb_strerror(st), vfrb_last_error());
    std::exit(code);
}

#define CLI_CHECK(call, code)                      \
    do {                                           \
        vfrb_status st_ = (call);                  \
        if (st_ != VFRB_OK) cli::die(code, #call, st_); \
    } while (0)

inline double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Reads one float per line with "%f" (PIPELINE:360-433): the line count is
// the number of DMs, each line parsed with sscanf.
inline bool read_dm_list(const char* path, std::vector<float>& dms) {
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[256];
    dms.clear();
    while (std::fgets(line, sizeof line, f)) {
        float v = 0.0f;
        std::sscanf(line, "%f", &v);
        dms.push_back(v);
    }
    std::fclose(f);
    return true;
}

// Log sink for the frame sequencer: the reference writes these lines to the
// run's .log file.
inline void log_to_file(const char* line, void* user) {
    FILE* f = static_cast<FILE*>(user);
    if (f) std::fputs(line, f);
}

// Empties an output file opened for writing and rewinds it.
inline void truncate(FILE* f) {
    std::fflush(f);
    if (ftruncate(fileno(f), 0) != 0) std::perror("ftruncate");
    std::rewind(f);
}

inline void remove_if_exists(const char* path) {
    if (access(path, F_OK) != -1) std::remove(path);
}

}  // namespace cli
