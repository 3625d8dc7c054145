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

// Pipelined tiles (BM, BN, BK, STAGES, wave grid WGM x WGN): both operands
// reach LDS by DMA through a ring of STAGES slabs.
#define LT_HGEMM2_TILES(X)    \
    X(256, 256, 32, 4, 2, 4)  \
    X(256, 256, 32, 3, 2, 4)  \
    X(256, 256, 32, 5, 2, 4)  \
    X(256, 128, 64, 3, 2, 2)  \
    X(128, 256, 64, 3, 2, 2)  \
    X(256, 128, 32, 4, 2, 2)  \
    X(128, 256, 32, 4, 2, 2)  \
    X(128, 128, 64, 3, 2, 2)  \
    X(128, 128, 32, 4, 2, 2)
