#include "tuning.h"

#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>
#include <sys/stat.h>
#include <vector>

namespace tuning {

std::string dir() {
    if (const char* d = getenv("LTHIP_TUNING_DIR")) return d;
    char buf[4096];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (n <= 0) return "tuning";
    buf[n] = 0;
    std::string exe(buf);
    const size_t slash = exe.rfind('/');
    const std::string bindir = slash == std::string::npos ? "." : exe.substr(0, slash);
    return bindir + "/../tuning";
}

static std::vector<std::vector<double>> read_rows(const std::string& file) {
    std::vector<std::vector<double>> rows;
    std::ifstream f(dir() + "/" + file);
    std::string line;
    bool header = true;
    while (std::getline(f, line)) {
        if (header) {
            header = false;
            continue;
        }
        std::vector<double> r;
        std::stringstream ss(line);
        std::string cell;
        while (std::getline(ss, cell, ',')) r.push_back(atof(cell.c_str()));
        rows.push_back(r);
    }
    return rows;
}

static void append_row(const std::string& file, const std::string& header, const std::string& row) {
    mkdir(dir().c_str(), 0755);
    const std::string path = dir() + "/" + file;
    std::ifstream probe(path);
    const bool exists = probe.good();
    probe.close();
    std::ofstream f(path, std::ios::app);
    if (!exists) f << header << "\n";
    f << row << "\n";
}

bool lookup_hgemm(int M, int N, int K, lthip_hgemm_config* c) {
    double best = std::numeric_limits<double>::infinity();
    for (const auto& r : read_rows("hgemm.csv")) {
        if (r.size() < 9 || (int)r[0] != M || (int)r[1] != N || (int)r[2] != K) continue;
        if (r[8] < best) {
            best = r[8];
            *c = lthip_hgemm_config{(int)r[3], (int)r[4], (int)r[5], (int)r[6], (int)r[7]};
        }
    }
    return best < std::numeric_limits<double>::infinity();
}

void append_hgemm(int M, int N, int K, const lthip_hgemm_config& c, float ms) {
    char row[256];
    snprintf(row, sizeof(row), "%d,%d,%d,%d,%d,%d,%d,%d,%.6f", M, N, K, c.bm, c.bn, c.bk, c.split_k, c.stages, ms);
    append_row("hgemm.csv", "M,N,K,bm,bn,bk,split_k,stages,ms", row);
}

bool lookup_w4a16(int M, int N, int K, int group, lthip_w4a16_config* c) {
    double best = std::numeric_limits<double>::infinity();
    for (const auto& r : read_rows("w4a16.csv")) {
        if (r.size() < 9 || (int)r[0] != M || (int)r[1] != N || (int)r[2] != K || (int)r[3] != group) continue;
        if (r[8] < best) {
            best = r[8];
            *c = lthip_w4a16_config{(int)r[4], (int)r[5], (int)r[6], (int)r[7]};
        }
    }
    return best < std::numeric_limits<double>::infinity();
}

void append_w4a16(int M, int N, int K, int group, const lthip_w4a16_config& c, float ms) {
    char row[256];
    snprintf(row, sizeof(row), "%d,%d,%d,%d,%d,%d,%d,%d,%.6f", M, N, K, group, c.kind, c.tile_n, c.tile_m, c.split_k,
             ms);
    append_row("w4a16.csv", "M,N,K,group,kind,tile_n,tile_m,split_k,ms", row);
}

bool lookup_batched(int count, lthip_hgemm_config* c, int* policy) {
    double best = std::numeric_limits<double>::infinity();
    for (const auto& r : read_rows("batched.csv")) {
        if (r.size() < 7 || (int)r[0] != count) continue;
        if (r[6] < best) {
            best = r[6];
            *c = lthip_hgemm_config{(int)r[1], (int)r[2], (int)r[3], 1, (int)r[4]};
            *policy = (int)r[5];
        }
    }
    return best < std::numeric_limits<double>::infinity();
}

}  // namespace tuning
