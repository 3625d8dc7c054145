// libthom CLI: ./libthom <test|bench|tune> <routine> <implementation> ...
// Same dispatch as the reference's main.cpp:41-72: parse mode and routine,
// hand the rest to the routine's handler.  Exits non-zero when a result check
// failed.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "enums.h"
#include "routines.h"

static const std::unordered_map<std::string, RunMode> runModeMap = {
    {"test", RunMode::TEST}, {"bench", RunMode::BENCH}, {"tune", RunMode::TUNE}};

using RoutineHandler = void (*)(RunMode, int, char**);

static const std::unordered_map<std::string, RoutineHandler> routinesMap = {
#ifdef LIB_HAS_ROUTINE_REDUCTION
    {"reduction", reduction_handler},
#endif
#ifdef LIB_HAS_ROUTINE_SDOT
    {"sdot", sdot_handler},
#endif
#ifdef LIB_HAS_ROUTINE_SGEMV
    {"sgemv", sgemv_handler},
#endif
#ifdef LIB_HAS_ROUTINE_DGEMM
    {"dgemm", dgemm_handler},
#endif
#ifdef LIB_HAS_ROUTINE_HGEMM
    {"hgemm", hgemm_handler},
    {"hgemm_batched", hgemm_batched_handler},
#endif
#ifdef LIB_HAS_ROUTINE_W4A16
    {"w4a16", w4a16_handler},
    {"w4a16_grouped", w4a16_grouped_handler},
#endif
};

int main(int argc, char* argv[]) {
    if (argc < 3) {
        fprintf(stderr, "Not enough parameters, please provide at least ./libthom <test|bench|tune> <ROUTINE>\n");
        return 1;
    }
    const char* mode = argv[1];
    const char* routine = argv[2];
    auto m = runModeMap.find(mode);
    auto r = routinesMap.find(routine);
    if (m == runModeMap.end() || r == routinesMap.end()) {
        fprintf(stderr, "Invalid mode or routine: %s %s\nRoutines:", mode, routine);
        for (const auto& e : routinesMap) fprintf(stderr, " %s", e.first.c_str());
        fprintf(stderr, "\n");
        return EXIT_FAILURE;
    }
    r->second(m->second, argc, argv);
    if (cli::g_check_failed) {
        fprintf(stderr, "FAILED: results differ from the CPU check\n");
        return 3;
    }
    return EXIT_SUCCESS;
}
