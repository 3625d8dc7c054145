// Tuning database of the CLI: CSV files written by `libthom tune` and by
// libthom_amd/tuner.py (same columns), read back by the "tuned" implementations.
// Location: $LTHIP_TUNING_DIR, else libthom_amd/tuning next to the binary.
//   hgemm.csv : M,N,K,bm,bn,bk,split_k,stages,ms
//   w4a16.csv : M,N,K,group,kind,tile_n,tile_m,split_k,ms
//   batched.csv : count,bm,bn,bk,stages,policy,ms (written by libthom_amd/tuner.py batched)
#pragma once

#include <string>

#include "lthip.h"

namespace tuning {

std::string dir();
bool lookup_hgemm(int M, int N, int K, lthip_hgemm_config* c);
void append_hgemm(int M, int N, int K, const lthip_hgemm_config& c, float ms);
bool lookup_w4a16(int M, int N, int K, int group, lthip_w4a16_config* c);
void append_w4a16(int M, int N, int K, int group, const lthip_w4a16_config& c, float ms);
bool lookup_batched(int count, lthip_hgemm_config* c, int* policy);

}  // namespace tuning
