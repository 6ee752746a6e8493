// pk_tiff.hip -- TIFF LZW strip codec (host code, no device work).
//
// The reference reads LZW strips through libtiff (io/cpp-tiff/src/
// parallelreadtiff.cpp:283-300, one TIFFReadEncodedStrip per strip across
// OpenMP workers) and writes them with its own encoder
// (io/cpp-tiff/src/lzwencode.cpp, called per strip from parallelwritetiff.cpp).
// Both follow TIFF 6.0 section 13: codes 0-255 are bytes, 256 Clear, 257 End
// of Information, table entries from 258; codes are 9 to 12 bits wide, packed
// most significant bit first; the width grows one code early (the decoder reads
// wider codes once its next free entry reaches 2^w - 1, the encoder, whose table
// runs one entry ahead, once its next free entry passes 2^w - 1); the encoder
// emits Clear before its table reaches 4094 entries.  Each strip is independent,
// so callers decode or encode strips on as many host threads as they like.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/pkdecon.h"

namespace {

constexpr int kClear = 256, kEoi = 257, kFirst = 258, kMaxBits = 12, kTableMax = 4096;

struct BitWriter {
    uint8_t* dst;
    size_t cap, pos = 0;
    uint32_t acc = 0;  // pending bits, right-aligned
    int nacc = 0;
    bool ok = true;
    void put(int code, int width) {
        acc = (acc << width) | (uint32_t)code;
        nacc += width;
        while (nacc >= 8) {
            nacc -= 8;
            if (pos < cap) dst[pos++] = (uint8_t)(acc >> nacc);
            else ok = false;
        }
        acc &= (1u << nacc) - 1u;
    }
    void flush() {
        if (nacc > 0) {
            if (pos < cap) dst[pos++] = (uint8_t)(acc << (8 - nacc));
            else ok = false;
            nacc = 0;
            acc = 0;
        }
    }
};

}  // namespace

extern "C" pk_status pk_lzw_decode(const uint8_t* src, size_t n, uint8_t* dst, size_t cap, size_t* out_len) {
    if ((!src && n) || (!dst && cap) || !out_len) return PK_ERR_INVALID_ARG;
    // entry k: prefix code, last byte, first byte and string length
    std::vector<int16_t> prefix(kTableMax);
    std::vector<uint8_t> last(kTableMax), first(kTableMax);
    std::vector<uint32_t> len(kTableMax);
    for (int k = 0; k < 256; ++k) {
        prefix[k] = -1;
        last[k] = first[k] = (uint8_t)k;
        len[k] = 1;
    }
    size_t inpos = 0, out = 0;
    uint64_t acc = 0;  // unread bits, right-aligned
    int nacc = 0;
    int width = 9, next = kFirst, prev = -1;
    auto emit = [&](int code) -> bool {  // the string of code, written back to front
        const uint32_t L = len[code];
        if (out + L > cap) return false;
        size_t q = out + L;
        for (int c = code; c >= 0; c = prefix[c]) dst[--q] = last[c];
        out += L;
        return true;
    };
    for (;;) {
        while (nacc < width && inpos < n) {
            acc = (acc << 8) | src[inpos++];
            nacc += 8;
        }
        if (nacc < width) break;  // strip ends without End of Information
        nacc -= width;
        const int code = (int)((acc >> nacc) & ((1u << width) - 1u));
        if (code == kEoi) break;
        if (code == kClear) {
            width = 9;
            next = kFirst;
            prev = -1;
            continue;
        }
        if (prev < 0) {  // first code after Clear: a single byte
            if (code > 255) return PK_ERR_INVALID_ARG;
            if (!emit(code)) return PK_ERR_OUT_SIZE;
            prev = code;
            continue;
        }
        int add_first;
        if (code < next) {
            if (!emit(code)) return PK_ERR_OUT_SIZE;
            add_first = first[code];
        } else if (code == next) {  // the string being defined: prev + its own first byte
            add_first = first[prev];
        } else {
            return PK_ERR_INVALID_ARG;
        }
        if (next < kTableMax) {
            prefix[next] = (int16_t)prev;
            last[next] = (uint8_t)add_first;
            first[next] = first[prev];
            len[next] = len[prev] + 1;
            if (code == next && !emit(next)) return PK_ERR_OUT_SIZE;
            ++next;
        } else if (code == next) {
            return PK_ERR_INVALID_ARG;
        }
        if (next >= (1 << width) - 1 && width < kMaxBits) ++width;
        prev = code;
    }
    *out_len = out;
    return PK_OK;
}

extern "C" size_t pk_lzw_encode_bound(size_t n) {
    // one code of at most 12 bits per input byte, a Clear every 3836 codes,
    // the leading Clear, the last code and End of Information
    return n + n / 2 + n / 2000 + 16;
}

extern "C" pk_status pk_lzw_encode(const uint8_t* src, size_t n, uint8_t* dst, size_t cap, size_t* out_len) {
    if ((!src && n) || (!dst && cap) || !out_len) return PK_ERR_INVALID_ARG;
    // open-addressed (prefix code, byte) -> code table; a generation stamp
    // empties it at each Clear without touching every slot
    constexpr int kSlots = 1 << 14;
    std::vector<uint32_t> key(kSlots), stamp(kSlots, 0);
    std::vector<uint16_t> val(kSlots);
    uint32_t gen = 1;
    BitWriter w{dst, cap};
    int width = 9, next = kFirst;
    w.put(kClear, width);
    auto grow = [&]() {  // after an entry is (or would be) added
        if (next == kTableMax - 2) {
            w.put(kClear, width);
            width = 9;
            next = kFirst;
            ++gen;
        } else if (next > (1 << width) - 1) {
            ++width;
        }
    };
    if (n > 0) {
        int omega = src[0];
        for (size_t i = 1; i < n; ++i) {
            const uint8_t k = src[i];
            const uint32_t kk = ((uint32_t)omega << 8) | k;
            uint32_t h = (kk * 2654435761u) >> 18;  // 14-bit slot
            bool found = false;
            while (stamp[h] == gen) {
                if (key[h] == kk) {
                    found = true;
                    break;
                }
                h = (h + 1) & (kSlots - 1);
            }
            if (found) {
                omega = val[h];
                continue;
            }
            w.put(omega, width);
            stamp[h] = gen;
            key[h] = kk;
            val[h] = (uint16_t)next;
            ++next;
            grow();
            omega = k;
        }
        w.put(omega, width);
        ++next;  // the entry the decoder adds on this code
        grow();
    }
    w.put(kEoi, width);
    w.flush();
    if (!w.ok) return PK_ERR_OUT_SIZE;
    *out_len = w.pos;
    return PK_OK;
}
