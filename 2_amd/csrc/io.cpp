// io.cpp -- host-side helpers of the output formats.
//
// CRC-64 with the ISO polynomial as Go's hash/crc64 computes it (the
// #dump002 format's checksum, dump/writer.go:12,30-37,60-67): reflected
// polynomial 0xD800000000000000, initial value and final xor all ones.
// Table-driven, 8 bytes per step (slicing-by-8).
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "demag_plan.hpp"

namespace {

struct Crc64Tables {
    uint64_t t[8][256];
    Crc64Tables()
    {
        const uint64_t poly = 0xD800000000000000ULL;
        for (int i = 0; i < 256; i++) {
            uint64_t c = (uint64_t)i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ poly : (c >> 1);
            t[0][i] = c;
        }
        for (int i = 0; i < 256; i++)
            for (int s = 1; s < 8; s++) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xff];
    }
};

const Crc64Tables &tables()
{
    static const Crc64Tables T;
    return T;
}

} // namespace

extern "C" unsigned long long mmx_crc64_iso(const void *data, unsigned long long n,
                                            unsigned long long crc)
{
    const Crc64Tables &T = tables();
    const unsigned char *p = static_cast<const unsigned char *>(data);
    uint64_t c = ~(uint64_t)crc;
    while (n >= 8) {
        uint64_t w;
        std::memcpy(&w, p, 8); // little-endian host
        c ^= w;
        c = T.t[7][c & 0xff] ^ T.t[6][(c >> 8) & 0xff] ^ T.t[5][(c >> 16) & 0xff] ^
            T.t[4][(c >> 24) & 0xff] ^ T.t[3][(c >> 32) & 0xff] ^ T.t[2][(c >> 40) & 0xff] ^
            T.t[1][(c >> 48) & 0xff] ^ T.t[0][c >> 56];
        p += 8;
        n -= 8;
    }
    while (n--) c = T.t[0][(c ^ *p++) & 0xff] ^ (c >> 8);
    return ~c;
}
