// pk_lz4.hip -- LZ4 block codec (host code, no device work) for Zarr chunks.
//
// The reference's Zarr stores default to blosc with the lz4 codec, byte
// shuffle and clevel 5 (io/cpp-zarr/src/zarr.cpp:15-18); its reader and writer
// hand each chunk to c-blosc (parallelreadzarr.cpp:148-157,
// parallelwritezarr.cpp).  Blosc splits a chunk into blocks and each block
// into streams compressed as raw LZ4 blocks.  This file is that LZ4 block
// layer, written from the published LZ4 block format: a sequence is a token
// (literal length in the high nibble, match length - 4 in the low one, 15
// meaning "more bytes follow", each 255 adding 255), the literals, a 2-byte
// little-endian offset and the extra match-length bytes; the last sequence
// carries literals only; a match starts at least 12 bytes and ends at least 5
// bytes before the end of the block.  The blosc framing is in zarr_io.py.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/pkdecon.h"

extern "C" pk_status pk_lz4_decode(const uint8_t* src, size_t n, uint8_t* dst, size_t cap, size_t* out_len) {
    if ((!src && n) || (!dst && cap) || !out_len) return PK_ERR_INVALID_ARG;
    size_t ip = 0, op = 0;
    while (ip < n) {
        const uint8_t token = src[ip++];
        size_t lit = token >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= n) return PK_ERR_INVALID_ARG;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > n - ip) return PK_ERR_INVALID_ARG;
        if (lit > cap - op) return PK_ERR_OUT_SIZE;
        std::memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip == n) break;  // the last sequence: literals only
        if (n - ip < 2) return PK_ERR_INVALID_ARG;
        const size_t off = (size_t)src[ip] | ((size_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return PK_ERR_INVALID_ARG;
        size_t ml = token & 15;
        if (ml == 15) {
            uint8_t b;
            do {
                if (ip >= n) return PK_ERR_INVALID_ARG;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += 4;
        if (ml > cap - op) return PK_ERR_OUT_SIZE;
        const uint8_t* from = dst + op - off;
        if (off >= ml) {
            std::memcpy(dst + op, from, ml);
        } else {
            for (size_t k = 0; k < ml; ++k) dst[op + k] = from[k];  // overlapping copy repeats the pattern
        }
        op += ml;
    }
    *out_len = op;
    return PK_OK;
}

extern "C" size_t pk_lz4_encode_bound(size_t n) { return n + n / 255 + 16; }

namespace {
inline uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
void put_len(uint8_t* dst, size_t& op, size_t v) {  // the bytes after a nibble of 15
    while (v >= 255) {
        dst[op++] = 255;
        v -= 255;
    }
    dst[op++] = (uint8_t)v;
}
}  // namespace

extern "C" pk_status pk_lz4_encode(const uint8_t* src, size_t n, uint8_t* dst, size_t cap, size_t* out_len) {
    if ((!src && n) || (!dst && cap) || !out_len) return PK_ERR_INVALID_ARG;
    if (cap < pk_lz4_encode_bound(n)) return PK_ERR_OUT_SIZE;
    constexpr int kHashBits = 16;
    std::vector<int64_t> table((size_t)1 << kHashBits, -1);
    size_t op = 0, anchor = 0, i = 0;
    auto emit = [&](size_t lit_end, size_t off, size_t ml) {  // literals [anchor, lit_end) then a match
        const size_t lit = lit_end - anchor;
        const size_t mcode = ml ? ml - 4 : 0;
        dst[op++] = (uint8_t)(((lit < 15 ? lit : 15) << 4) | (ml ? (mcode < 15 ? mcode : 15) : 0));
        if (lit >= 15) put_len(dst, op, lit - 15);
        std::memcpy(dst + op, src + anchor, lit);
        op += lit;
        if (ml) {
            dst[op++] = (uint8_t)(off & 0xff);
            dst[op++] = (uint8_t)(off >> 8);
            if (mcode >= 15) put_len(dst, op, mcode - 15);
        }
    };
    if (n >= 13) {
        const size_t match_end_max = n - 5;
        while (i + 12 <= n) {
            const uint32_t seq = rd32(src + i);
            const uint32_t h = (seq * 2654435761u) >> (32 - kHashBits);
            const int64_t ref = table[h];
            table[h] = (int64_t)i;
            if (ref >= 0 && i - (size_t)ref <= 65535 && rd32(src + ref) == seq) {
                size_t ml = 4;
                while (i + ml < match_end_max && src[ref + ml] == src[i + ml]) ++ml;
                emit(i, i - (size_t)ref, ml);
                i += ml;
                anchor = i;
                continue;
            }
            ++i;
        }
    }
    emit(n, 0, 0);  // trailing literals
    *out_len = op;
    return PK_OK;
}
