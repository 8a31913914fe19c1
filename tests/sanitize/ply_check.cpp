// Host sanitizer driver for the PLY loader (test infrastructure): built with -fsanitize=address,undefined together
// with gps-pbvr-web_amd/csrc/gps_ply.cpp by tests/test_sanitizers.py.  Loads each file named on the command line
// with gps_ply_load and, when it loads, reads every array and activates it; prints one line per file:
// "<status> <count> <sh degree> <checksum>" (status: 0 or the GPS_PLY_* code).
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/gps_hip.h"

// sum over the bytes b[i] of the concatenated arrays of b[i] (i * 2654435761 + 1), modulo 2^64
static uint64_t checksum(const void *p, size_t n, uint64_t &pos, uint64_t h) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; i++, pos++) h += (uint64_t)b[i] * (pos * 2654435761ull + 1ull);
    return h;
}

int main(int argc, char **argv) {
    for (int a = 1; a < argc; a++) {
        FILE *f = std::fopen(argv[a], "rb");
        if (!f) return 2;
        std::vector<uint8_t> bytes;
        uint8_t buf[65536];
        size_t k;
        while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + k);
        std::fclose(f);
        GpsPly *g = nullptr;
        const int rc = gps_ply_load(bytes.data(), bytes.size(), &g);
        if (rc != 0) {
            std::printf("%d 0 0 0\n", rc);
            continue;
        }
        uint64_t count = 0;
        uint32_t deg = 0;
        int binary = 0;
        gps_ply_info(g, &count, &deg, &binary);
        const size_t rest = 3u * ((deg + 1u) * (deg + 1u) - 1u);
        std::vector<float> pos(3 * count), ls(3 * count), rot(4 * count), op(count), dc(3 * count), sh(rest * count);
        gps_ply_read(g, pos.data(), ls.data(), rot.data(), op.data(), dc.data(), sh.data());
        std::vector<GpsGaussian> gpu(count);
        gps_ply_to_gpu(g, gpu.data());
        uint64_t h = 0, at = 0;
        h = checksum(pos.data(), pos.size() * 4, at, h);
        h = checksum(sh.data(), sh.size() * 4, at, h);
        h = checksum(gpu.data(), gpu.size() * sizeof(GpsGaussian), at, h);
        std::printf("0 %llu %u %llu\n", (unsigned long long)count, deg, (unsigned long long)h);
        gps_ply_free(g);
    }
    return 0;
}
