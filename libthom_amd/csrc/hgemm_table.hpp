// Tile shapes (BM, BN, BK) compiled for hgemm and the grouped hgemm; the
// autotuner searches these times split-K.
#pragma once
#define LT_HGEMM_TILES(X) \
    X(64, 64, 64)         \
    X(64, 128, 64)        \
    X(128, 64, 64)        \
    X(128, 128, 32)       \
    X(128, 128, 64)       \
    X(128, 256, 64)       \
    X(256, 128, 64)       \
    X(256, 256, 64)
