// A C++ caller of include/fk20.h, compiled against the header and linked with -lfk20 the way a
// reference caller (fk20benchmark.cu:405-407, fk20test_poly.cu:332-370) would be after the
// switch: no HIP or ctypes in sight, plain pointers, sizes and status codes.
//
// usage: abi_caller <dir> <rows> <N>
//   reads  <dir>/poly.bin  (rows x N fr_t) and <dir>/xext.bin (16 x N/8 g1p_t)
//   writes <dir>/h_plain.bin    fk20_poly2h_fft (drop-in, library scratch)
//          <dir>/h_noalloc.bin  fk20_poly2h_fft_noalloc (fixed workspace)
//          <dir>/h_ctx.bin      fk20_ctx_create / fk20_setup_load / fk20_ctx_reserve /
//                               fk20_prove_rows on a stream (the batched path)
//          <dir>/h_affine.bin   fk20_g1p_to_affine of h_plain (rows x N/8 g1a_t)
// Exits 0 when every call returned FK20_OK, 1 on the first failing call (named on stderr),
// 2 on a usage or file error.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fk20.h"

namespace {

bool read_file(const std::string &path, std::vector<unsigned char> &buf) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const size_t got = std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return got == buf.size();
}

bool write_file(const std::string &path, const std::vector<unsigned char> &buf) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t put = std::fwrite(buf.data(), 1, buf.size(), f);
    return std::fclose(f) == 0 && put == buf.size();
}

int fail(const char *call, int rc) {
    std::fprintf(stderr, "abi_caller: %s returned %d (%s), hip error %d\n", call, rc, fk20_status_string(rc),
                 fk20_last_hip_error());
    return 1;
}

}  // namespace

#define CHECK(call)                           \
    do {                                      \
        const int rc_ = (call);               \
        if (rc_ != FK20_OK) return fail(#call, rc_); \
    } while (0)

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: abi_caller <dir> <rows> <N>\n");
        return 2;
    }
    const std::string dir = argv[1];
    const unsigned rows = (unsigned)std::strtoul(argv[2], nullptr, 10);
    const unsigned N = (unsigned)std::strtoul(argv[3], nullptr, 10);
    const unsigned n2 = N / 8;
    std::vector<unsigned char> poly((size_t)rows * N * sizeof(fk20_fr_t));
    std::vector<unsigned char> xext((size_t)16 * n2 * sizeof(fk20_g1p_t));
    std::vector<unsigned char> h((size_t)rows * n2 * sizeof(fk20_g1p_t));
    std::vector<unsigned char> ha((size_t)rows * n2 * sizeof(fk20_g1a_t));
    if (!read_file(dir + "/poly.bin", poly) || !read_file(dir + "/xext.bin", xext)) {
        std::fprintf(stderr, "abi_caller: cannot read the inputs in %s\n", dir.c_str());
        return 2;
    }

    void *d_poly, *d_xext, *d_h, *d_ha;
    CHECK(fk20_malloc(&d_poly, poly.size()));
    CHECK(fk20_malloc(&d_xext, xext.size()));
    CHECK(fk20_malloc(&d_h, h.size()));
    CHECK(fk20_malloc(&d_ha, ha.size()));
    CHECK(fk20_memcpy_h2d(d_poly, poly.data(), poly.size()));
    CHECK(fk20_memcpy_h2d(d_xext, xext.data(), xext.size()));
    const fk20_fr_t *P = static_cast<const fk20_fr_t *>(d_poly);
    const fk20_g1p_t *X = static_cast<const fk20_g1p_t *>(d_xext);
    fk20_g1p_t *H = static_cast<fk20_g1p_t *>(d_h);

    // 1. the reference's host call, fk20_poly2h_fft (fk20.cuh:22), on the default stream
    CHECK(fk20_poly2h_fft(H, P, X, rows, N, nullptr));
    CHECK(fk20_synchronize());
    CHECK(fk20_memcpy_d2h(h.data(), d_h, h.size()));
    if (!write_file(dir + "/h_plain.bin", h)) return 2;
    CHECK(fk20_g1p_to_affine(static_cast<fk20_g1a_t *>(d_ha), H, (size_t)rows * n2, nullptr));
    CHECK(fk20_synchronize());
    CHECK(fk20_memcpy_d2h(ha.data(), d_ha, ha.size()));
    if (!write_file(dir + "/h_affine.bin", ha)) return 2;

    // 2. the allocation-free form
    CHECK(fk20_memset(d_h, 0, h.size()));
    CHECK(fk20_poly2h_fft_noalloc(H, P, X, rows, N, nullptr));
    CHECK(fk20_synchronize());
    CHECK(fk20_memcpy_d2h(h.data(), d_h, h.size()));
    if (!write_file(dir + "/h_noalloc.bin", h)) return 2;

    // 3. the batched path with a resident setup, on a stream
    fk20_stream_t stream;
    CHECK(fk20_stream_create(&stream));
    fk20_ctx_t *ctx;
    int setup;
    CHECK(fk20_ctx_create(&ctx));
    CHECK(fk20_setup_load(ctx, X, N, 8, &setup, stream));
    CHECK(fk20_ctx_reserve(ctx, rows, N));
    CHECK(fk20_memset(d_h, 0, h.size()));
    CHECK(fk20_prove_rows(ctx, setup, H, P, rows, stream));
    CHECK(fk20_stream_synchronize(stream));
    CHECK(fk20_memcpy_d2h(h.data(), d_h, h.size()));
    if (!write_file(dir + "/h_ctx.bin", h)) return 2;
    CHECK(fk20_ctx_destroy(ctx));
    CHECK(fk20_stream_destroy(stream));

    CHECK(fk20_free(d_poly));
    CHECK(fk20_free(d_xext));
    CHECK(fk20_free(d_h));
    CHECK(fk20_free(d_ha));
    std::printf("abi_caller: rows=%u N=%u ok\n", rows, N);
    return 0;
}
