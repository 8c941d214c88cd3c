// Robustness of the DICOM reader (ddfan_dicom.cpp) under AddressSanitizer and
// UBSan: every truncation and a fixed-seed stream of byte mutations of each
// seed file (argv) is handed to the three readers; each call must return a
// status without touching memory outside its buffers.
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "ddfan.h"

namespace ddf {
int fail(int code, const char*, ...) { return code; }
}  // namespace ddf

static std::vector<uint8_t> slurp(const char* p) {
    std::vector<uint8_t> b;
    FILE* f = fopen(p, "rb");
    if (!f) return b;
    int c;
    while ((c = fgetc(f)) != EOF) b.push_back((uint8_t)c);
    fclose(f);
    return b;
}

static void spill(const std::string& p, const std::vector<uint8_t>& b, size_t n) {
    FILE* f = fopen(p.c_str(), "wb");
    fwrite(b.data(), 1, n, f);
    fclose(f);
}

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}

static int exercise(const std::string& p, int rows, int cols, std::vector<uint16_t>& px) {
    ddfan_ctpd_header h;
    ddfan_ct_image_header img;
    float stats[64];
    int ok = 0;
    ok += ddfan_ctpd_read_header(p.c_str(), &h, stats, 64) == DDFAN_OK;
    ok += ddfan_ct_image_read_header(p.c_str(), &img) == DDFAN_OK;
    ok += ddfan_dicom_read_u16(p.c_str(), rows, cols, px.data()) == DDFAN_OK;
    const char* two[2] = {p.c_str(), p.c_str()};
    std::vector<uint16_t> both(2 * px.size());
    ok += ddfan_ctpd_read_frames(two, 2, rows, cols, both.data(), 2) == DDFAN_OK;
    return ok;
}

int main(int argc, char** argv) {
    // argv: tmp file, then (seed path, rows, cols) triples
    if (argc < 5 || (argc - 2) % 3) return 2;
    const std::string tmp = argv[1];
    long calls = 0, accepted = 0;
    for (int a = 2; a + 2 < argc; a += 3) {
        const std::vector<uint8_t> seed = slurp(argv[a]);
        const int rows = atoi(argv[a + 1]), cols = atoi(argv[a + 2]);
        std::vector<uint16_t> px((size_t)rows * cols);
        if (seed.empty()) return 3;
        spill(tmp, seed, seed.size());
        if (exercise(tmp, rows, cols, px) == 0) return 4;  // the seed itself must read
        for (size_t n = 0; n < seed.size(); n += 1 + seed.size() / 400) {
            spill(tmp, seed, n);
            accepted += exercise(tmp, rows, cols, px);
            ++calls;
        }
        for (int m = 0; m < 1500; ++m) {
            std::vector<uint8_t> b = seed;
            const int k = 1 + (int)(rnd() % 4);
            for (int j = 0; j < k; ++j) {
                const size_t at = 132 + rnd() % (b.size() - 132);
                const uint64_t r = rnd();
                b[at] = (r & 3) == 0 ? 0xFF : (r & 3) == 1 ? 0x00 : (uint8_t)(r >> 8);
            }
            spill(tmp, b, b.size());
            accepted += exercise(tmp, rows, cols, px);
            ++calls;
        }
    }
    printf("dicom fuzz ok: %ld inputs, %ld reads accepted\n", calls, accepted);
    return 0;
}
