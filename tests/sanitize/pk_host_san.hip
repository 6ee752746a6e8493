// Host-code driver for the AddressSanitizer / UndefinedBehaviorSanitizer build
// of libpkdecon's sources (host pass only; no kernel is launched and no device
// is needed).  It drives every entry point that runs on the host:
//   - the TIFF LZW and Zarr LZ4 strip codecs: round trips over structured and
//     random data, every output capacity below the encoded size, truncated and
//     bit-flipped streams and random garbage (the decoders must reject or stop
//     inside their buffers, never read or write past them);
//   - BackProjector's five back-projector types at odd, even and prime PSF
//     sizes and the three resolution rules, plus its argument errors;
//   - findGoodFactorNumber, the deskew and deskew-rotate size rules on
//     extreme geometry, strel('disk', r) for every radius up to 255;
//   - context creation without a device and the NULL-context error paths.
// Prints one "ok <section>" line per section and "all ok" at the end; any
// sanitizer finding aborts the process with a report on stderr.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pkdecon.h"
#include "../../petakit5d_amd/csrc/pk_misc.hpp"

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64() {  // splitmix64
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void require(bool c, const char* what) {
    if (!c) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        std::exit(1);
    }
}

std::vector<uint8_t> sample(int kind, size_t n) {
    std::vector<uint8_t> v(n);
    for (size_t i = 0; i < n; ++i) {
        switch (kind) {
            case 0: v[i] = 0; break;                                              // one long run
            case 1: v[i] = (uint8_t)next_u64(); break;                            // incompressible
            case 2: v[i] = (uint8_t)((i >> 1) * 7 + (i & 1 ? (i >> 9) : 0)); break;  // uint16 ramp bytes
            case 3: v[i] = (uint8_t)("abcabcabd"[i % 9]); break;                  // short repeats
            default: v[i] = (next_u64() % 17 == 0) ? (uint8_t)next_u64() : (uint8_t)(i / 300); break;
        }
    }
    return v;
}

using enc_fn = pk_status (*)(const uint8_t*, size_t, uint8_t*, size_t, size_t*);
using bound_fn = size_t (*)(size_t);

// Decoding into a heap buffer of exactly `cap` bytes: ASan flags any byte the
// decoder touches past it.
void decode_any(enc_fn dec, const uint8_t* src, size_t n, size_t cap) {
    std::vector<uint8_t> out(cap ? cap : 1);
    std::vector<uint8_t> in(src, src + n);  // exact-size copy: reads past n are caught too
    size_t len = 0;
    const pk_status st = dec(n ? in.data() : nullptr, n, cap ? out.data() : nullptr, cap, &len);
    if (st == PK_OK) require(len <= cap, "decoded length within capacity");
}

void codec(const char* name, enc_fn enc, enc_fn dec, bound_fn bound) {
    const size_t sizes[] = {1, 2, 7, 64, 255, 4097, 65535, 65536, 65537, 300000};
    for (int kind = 0; kind < 5; ++kind)
        for (size_t n : sizes) {
            const std::vector<uint8_t> src = sample(kind, n);
            std::vector<uint8_t> e(bound(n));
            size_t el = 0;
            require(enc(src.data(), n, e.data(), e.size(), &el) == PK_OK, "encode at the bound");
            require(el <= e.size(), "encoded length within the bound");
            e.resize(el);
            std::vector<uint8_t> d(n);
            size_t dl = 0;
            require(dec(e.data(), el, d.data(), n, &dl) == PK_OK && dl == n && d == src, "round trip");
            if (n <= 4097) {
                // every smaller output capacity: OUT_SIZE, never a write past it
                for (size_t cap = 0; cap < n; cap += (n > 300 ? 37 : 1)) decode_any(dec, e.data(), el, cap);
                // every truncation of the stream
                for (size_t m = 0; m < el; m += (el > 300 ? 13 : 1)) decode_any(dec, e.data(), m, n);
                // encoding into too small an output fails cleanly
                std::vector<uint8_t> small(el > 1 ? el - 1 : 1);
                size_t sl = 0;
                const pk_status st = enc(src.data(), n, small.data(), el > 1 ? el - 1 : 0, &sl);
                require(st != PK_OK || sl < el, "encode into a short buffer");
            }
            // bit flips
            for (int t = 0; t < 64 && el; ++t) {
                std::vector<uint8_t> f = e;
                f[next_u64() % el] ^= (uint8_t)(1u << (next_u64() % 8));
                decode_any(dec, f.data(), el, n);
            }
        }
    // garbage streams
    for (int t = 0; t < 2000; ++t) {
        const size_t n = 1 + next_u64() % 600;
        std::vector<uint8_t> g(n);
        for (auto& b : g) b = (uint8_t)next_u64();
        decode_any(dec, g.data(), n, next_u64() % 4096);
    }
    decode_any(dec, nullptr, 0, 16);
    size_t l = 0;
    require(enc(nullptr, 0, nullptr, 0, &l) != PK_ERR_NO_MEMORY, "empty encode");
    std::printf("ok %s\n", name);
}

std::vector<float> gauss_psf(const int64_t d[3], double sxy, double sz, double shear) {
    std::vector<float> p((size_t)(d[0] * d[1] * d[2]));
    const double c[3] = {(d[0] - 1) / 2.0, (d[1] - 1) / 2.0, (d[2] - 1) / 2.0};
    for (int64_t k = 0; k < d[2]; ++k)
        for (int64_t j = 0; j < d[1]; ++j)
            for (int64_t i = 0; i < d[0]; ++i) {
                const double z = k - c[2], y = i - c[0], x = j - c[1] - shear * z;
                p[(size_t)(i + d[0] * (j + d[1] * k))] =
                    (float)std::exp(-(x * x + y * y) / (2 * sxy * sxy) - z * z / (2 * sz * sz));
            }
    return p;
}

void backprojector() {
    const int64_t shapes[][3] = {{9, 9, 9}, {12, 10, 8}, {13, 11, 7}, {17, 17, 31}, {1, 9, 9}};
    for (const auto& d : shapes) {
        const std::vector<float> psf = gauss_psf(d, 1.3, 2.2, 0.4);
        const size_t P = psf.size();
        for (int type = PK_BP_TRADITIONAL; type <= PK_BP_WIENER_BUTTERWORTH; ++type)
            for (int rf = 0; rf < 3; ++rf) {
                std::vector<double> bp(P), otf(2 * P), info(10);
                const double ires[3] = {2.0, 2.0, 3.0};
                const pk_status st = pk_wb_backprojector(nullptr, psf.data(), d, type, 0.05, 1.0, 10, rf, ires,
                                                         bp.data(), otf.data(), info.data());
                // a 1-pixel dimension has no FWHM crossing: the reference raises an error there
                require(st == PK_OK || d[0] == 1, "BackProjector");
            }
    }
    std::vector<float> one(27, 1.f);
    std::vector<double> bp(27);
    const int64_t d3[3] = {3, 3, 3}, d0[3] = {0, 3, 3};
    require(pk_wb_backprojector(nullptr, one.data(), d0, 4, 1, 1, 10, 0, nullptr, bp.data(), nullptr, nullptr) != PK_OK,
            "empty PSF rejected");
    require(pk_wb_backprojector(nullptr, one.data(), d3, 4, 1, 1, 10, 2, nullptr, bp.data(), nullptr, nullptr) != PK_OK,
            "resFlag 2 without iRes rejected");
    require(pk_wb_backprojector(nullptr, nullptr, d3, 4, 1, 1, 10, 0, nullptr, bp.data(), nullptr, nullptr) != PK_OK,
            "NULL PSF rejected");
    (void)pk_wb_backprojector(nullptr, one.data(), d3, 9, 1, 1, 10, 0, nullptr, bp.data(), nullptr, nullptr);
    std::printf("ok backprojector\n");
}

void sizes() {
    for (int64_t n = 1; n < 20000; n += (n < 3000 ? 1 : 97))
        for (int dir = -1; dir <= 1; dir += 2)
            for (int mp : {5, 7}) {
                const int64_t g = pk_find_good_factor_number(n, dir, 0, mp);
                require(g >= 1, "good factor positive");
            }
    const double angles[] = {0.0, 32.45, -32.45, 89.999, 90.0, 135.0};
    const double dzs[] = {1e-6, 0.2, 0.5, 5.0, 1e3};
    const int64_t dims[][3] = {{1, 1, 1}, {64, 96, 40}, {2048, 2048, 512}, {1, 1 << 20, 3}};
    for (const auto& d : dims)
        for (double a : angles)
            for (double dz : dzs)
                for (int rv = 0; rv < 2; ++rv) {
                    int64_t o[3];
                    (void)pk_deskew_out_dims(d, a, dz, 0.108, rv, 1, o);
                    (void)pk_deskew_out_dims(d, a, dz, 0.108, rv, 0, o);
                    for (int obj = 0; obj < 2; ++obj) (void)pk_deskew_rotate_out_dims(d, a, dz, 0.108, rv, obj, 2.0, o);
                }
    int64_t o[3];
    const int64_t bad[3] = {0, 4, 4};
    require(pk_deskew_out_dims(bad, 32.45, 0.3, 0.108, 0, 1, o) != PK_OK, "empty deskew rejected");
    require(pk_deskew_rotate_out_dims(bad, 32.45, 0.3, 0.108, 0, 0, 2.0, o) != PK_OK, "empty DSR rejected");
    require(pk_deskew_out_dims(dims[1], 32.45, -0.1, 0.108, 0, 1, o) != PK_OK, "negative dz rejected");
    for (int r = 0; r <= 255; ++r) {
        const pk::DiskSE se = pk::disk_se(r);
        require(se.R >= r - 2 && se.R <= r, "disk extent");
    }
    std::printf("ok sizes\n");
}

void contexts() {
    pk_ctx* c = nullptr;
    const pk_status st = pk_ctx_create(0, 0, &c);
    require(st == PK_ERR_NO_DEVICE && c == nullptr, "no device: PK_ERR_NO_DEVICE");
    require(pk_ctx_create(0, 0, nullptr) == PK_ERR_INVALID_ARG, "NULL out rejected");
    pk_ctx_destroy(nullptr);
    require(pk_last_error(nullptr) != nullptr, "last error of NULL context");
    pk_opts o;
    pk_opts_default(&o);
    pk_bp_opts bo;
    pk_bp_opts_default(&bo);
    const int64_t d[3] = {8, 8, 4};
    float f = 1.f;
    require(pk_decon_lucy(nullptr, nullptr, &f, d, 1, &o, nullptr, nullptr) != PK_OK, "NULL context rejected");
    std::printf("ok contexts\n");
}

}  // namespace

int main(int argc, char** argv) {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    const char* only = argc > 1 ? argv[1] : "";
    if (!*only || !std::strcmp(only, "lzw")) codec("lzw", pk_lzw_encode, pk_lzw_decode, pk_lzw_encode_bound);
    if (!*only || !std::strcmp(only, "lz4")) codec("lz4", pk_lz4_encode, pk_lz4_decode, pk_lz4_encode_bound);
    if (!*only || !std::strcmp(only, "bp")) backprojector();
    if (!*only || !std::strcmp(only, "sizes")) sizes();
    if (!*only || !std::strcmp(only, "ctx")) contexts();
    std::printf("all ok\n");
    return 0;
}
