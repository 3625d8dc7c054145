#pragma once

// Run modes of the libthom CLI (the reference's enums.h:4-8).
enum RunMode { TEST, BENCH, TUNE };
