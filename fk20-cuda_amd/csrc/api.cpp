// C ABI of libfk20.so: argument checks, per-device tables and scratch, and the host-side
// launch sequences (include/fk20.h documents every entry point and the reference
// interface it replaces).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/fk20.h"
#include "common.hip.h"
#include "fast.hip.h"

using namespace fk20;

// ---------------------------------------------------------------- static device memory
//
// The constant tables and twiddle programs of the drop-in entry points live in static device
// memory (one copy per device, filled on the first call on that device), so no entry point
// allocates them.

static __device__ DevTables g_tables_dev;
static constexpr size_t PROG_OPS_CAP = (size_t)1 << 16;  // words; the programs take about 48 K
static __device__ uint32_t g_prog_ops[PROG_OPS_CAP];
static __device__ uint32_t g_prog_index[TW_INDEX_SIZE];

// Workspace of fk20_poly2h_fft_noalloc, the entry point for callers that cannot allocate:
// a fixed array per device, held until the process ends.  Host code uses only the address
// hipGetSymbolAddress returns for the current device (DeviceState::noalloc), never the
// array's name.
static constexpr size_t NOALLOC_SCRATCH_BYTES = FK20_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

thread_local int g_last_hip_error = 0;

int hip_status(hipError_t e) {
    if (e == hipSuccess) return FK20_OK;
    g_last_hip_error = (int)e;
    return e == hipErrorOutOfMemory ? FK20_ENOMEM : FK20_EHIP;
}

#define FK20_TRY(expr)                                   \
    do {                                                 \
        hipError_t _e = (expr);                          \
        if (_e != hipSuccess) return hip_status(_e);     \
    } while (0)

hipStream_t S(fk20_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

bool valid_N(unsigned N) { return N == 1024 || N == 2048 || N == 4096; }
bool valid_n(unsigned n) { return n == 128 || n == 256 || n == 512; }
unsigned log2u(unsigned n) { return 31u - (unsigned)__builtin_clz(n); }

// ---------------------------------------------------------------- per-device state

struct DeviceState {
    DevTables *tables = nullptr;            // device copy (g_tables_dev)
    Programs prog;                          // twiddle programs of the drop-in transforms
    void *scratch = nullptr;                // workspace of the drop-in entry points
    size_t scratch_bytes = 0;
    unsigned char *noalloc = nullptr;       // this device's address of `scratch`
};

std::mutex g_mu;
// one heap-allocated entry per device, so pointers handed out stay valid when the table grows
std::vector<std::unique_ptr<DeviceState>> g_dev;

}  // namespace

void fk20::build_host_tables(DevTables &T) {
    // w = 5^((r-1)/512) in Montgomery form, by square-and-multiply on the host copy of
    // the device Fr arithmetic
    fs five, w, acc;
    fs_zero(five);
    five.l[0] = 5;
    fs_to_mont(five, five);
    // exponent (r-1)/512 as 8 words
    uint32_t e[8];
    for (int i = 0; i < 8; i++) e[i] = FR_R[i];
    e[0] -= 1;
    for (int s = 0; s < 9; s++)
        for (int i = 0; i < 8; i++) e[i] = (e[i] >> 1) | (i < 7 ? (e[i + 1] << 31) : 0);
    fs_set(acc, FR_ONE);
    for (int i = 255; i >= 0; i--) {
        fs_mul(acc, acc, acc);
        if ((e[i / 32] >> (i % 32)) & 1) fs_mul(acc, acc, five);
    }
    w = acc;
    fs_set(acc, FR_ONE);
    for (int i = 0; i <= 512; i++) {
        fs std;
        fs_from_mont(std, acc);
        memcpy(T.tw_mont[i], acc.l, 32);
        memcpy(T.tw_std[i], std.l, 32);
        fs_mul(acc, acc, w);
    }
    // 2^-j: (r+1)/2 is 2^-1
    fs half, two;
    fs_zero(two);
    two.l[0] = 2;
    fs_to_mont(two, two);
    // inverse of 2 by Fermat
    fs_set(acc, FR_ONE);
    for (int i = 255; i >= 0; i--) {
        const uint32_t ew = FR_RM2[i / 32];
        fs_mul(acc, acc, acc);
        if ((ew >> (i % 32)) & 1) fs_mul(acc, acc, two);
    }
    half = acc;
    fs_set(acc, FR_ONE);
    for (int j = 0; j < 10; j++) {
        fs std;
        fs_from_mont(std, acc);
        memcpy(T.inv2k_mont[j], acc.l, 32);
        memcpy(T.inv2k_std[j], std.l, 32);
        fs t;
        fs_set(t, T.tw_mont[384]);
        fs_mul(t, t, acc);
        fs_from_mont(std, t);
        memcpy(T.inv2k_w384_std[j], std.l, 32);
        fs_mul(acc, acc, half);
    }
}

namespace {

// Returns the state of the current device, creating its tables on first use.
int device_state(DeviceState **out) {
    int dev = 0;
    FK20_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
    if (!g_dev[dev]) g_dev[dev].reset(new DeviceState());
    DeviceState &st = *g_dev[dev];
    // set-up of this device, once, with the device current: symbol addresses of the static
    // arrays, then the tables and programs copied into them (no allocation)
    if (!st.noalloc) {
        void *d = nullptr;
        FK20_TRY(hipGetSymbolAddress(&d, HIP_SYMBOL(scratch)));
        st.noalloc = reinterpret_cast<unsigned char *>(d);
    }
    if (!st.tables) {
        DevTables h;
        build_host_tables(h);
        void *d = nullptr;
        FK20_TRY(hipGetSymbolAddress(&d, HIP_SYMBOL(g_tables_dev)));
        FK20_TRY(hipMemcpy(d, &h, sizeof(DevTables), hipMemcpyHostToDevice));
        st.tables = reinterpret_cast<DevTables *>(d);
    }
    if (!st.prog.ops) {
        void *ops = nullptr, *index = nullptr;
        FK20_TRY(hipGetSymbolAddress(&ops, HIP_SYMBOL(g_prog_ops)));
        FK20_TRY(hipGetSymbolAddress(&index, HIP_SYMBOL(g_prog_index)));
        FK20_TRY(load_programs(st.prog, reinterpret_cast<uint32_t *>(ops), PROG_OPS_CAP,
                               reinterpret_cast<uint32_t *>(index)));
    }
    *out = &st;
    return FK20_OK;
}

// Grows the library-owned scratch of the current device to at least `bytes`.  The old
// buffer is released after the device is idle, since queued work may still use it.
int device_scratch(DeviceState *st, size_t bytes, void **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (st->scratch_bytes < bytes) {
        if (st->scratch) {
            FK20_TRY(hipDeviceSynchronize());
            FK20_TRY(hipFree(st->scratch));
            st->scratch = nullptr;
            st->scratch_bytes = 0;
        }
        FK20_TRY(hipMalloc(&st->scratch, bytes));
        st->scratch_bytes = bytes;
    }
    *out = st->scratch;
    return FK20_OK;
}

uint32_t *W(const void *p) { return reinterpret_cast<uint32_t *>(const_cast<void *>(p)); }

}  // namespace

namespace fk20 {
int status_from_hip(hipError_t e) { return hip_status(e); }

const DevTables *current_device_tables(hipError_t *err) {
    DeviceState *st;
    *err = hipSuccess;
    if (device_state(&st) != FK20_OK) {
        *err = (hipError_t)(g_last_hip_error ? g_last_hip_error : (int)hipErrorUnknown);
        return nullptr;
    }
    return st->tables;
}
}  // namespace fk20

extern "C" {

int fk20_last_hip_error(void) { return g_last_hip_error; }

const char *fk20_status_string(int status) {
    switch (status) {
    case FK20_OK: return "ok";
    case FK20_EINVAL: return "invalid argument";
    case FK20_ENOMEM: return "out of device memory";
    case FK20_EHIP: return "HIP runtime error";
    case FK20_ENODEV: return "no gfx950 device";
    case FK20_ETIMEDOUT: return "a bounded device-side wait expired";
    case FK20_ERANGE: return "request larger than the fixed workspace of an allocation-free entry point";
    default: return "unknown status";
    }
}

// ---------------------------------------------------------------- drop-in entry points

int fk20_poly2toeplitz_coefficients(fk20_fr_t *tc, const fk20_fr_t *poly, unsigned rows, unsigned N,
                                    fk20_stream_t stream) {
    if (!valid_N(N) || (rows && (!tc || !poly))) return FK20_EINVAL;
    return hip_status(launch_toeplitz(W(tc), W(poly), rows, N, S(stream)));
}

int fk20_poly2toeplitz_coefficients_fft(fk20_fr_t *tc_fft, const fk20_fr_t *poly, unsigned rows, unsigned N,
                                        fk20_stream_t stream) {
    if (!valid_N(N) || (rows && (!tc_fft || !poly))) return FK20_EINVAL;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    return hip_status(launch_toeplitz_fft(W(tc_fft), W(poly), rows, N, st->tables, S(stream)));
}

static int fr_transform(fk20_fr_t *out, const fk20_fr_t *in, unsigned count, unsigned n, bool inv,
                        fk20_stream_t stream) {
    if (!valid_n(n) || (count && (!out || !in))) return FK20_EINVAL;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    return hip_status(launch_fr_ntt(W(out), W(in), count, n, inv, st->tables, S(stream)));
}

int fr_fft(fk20_fr_t *output, const fk20_fr_t *input, unsigned count, unsigned n, fk20_stream_t stream) {
    return fr_transform(output, input, count, n, false, stream);
}

int fr_ift(fk20_fr_t *output, const fk20_fr_t *input, unsigned count, unsigned n, fk20_stream_t stream) {
    return fr_transform(output, input, count, n, true, stream);
}

// The G1 transforms run column-batched (lanes = transforms) through D in scratch, so the
// output may alias the input as in the reference (g1p_fft.cu:17).
static int g1_transform(fk20_g1p_t *out, const fk20_g1p_t *in, unsigned count, unsigned n, bool inv,
                        fk20_stream_t stream) {
    if (!valid_n(n) || (count && (!out || !in))) return FK20_EINVAL;
    if (!count) return FK20_OK;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    void *D;
    const size_t d_bytes = (size_t)count * n * 144;
    rc = device_scratch(st, d_bytes + (size_t)count * 4, &D);
    if (rc) return rc;
    uint32_t *flags = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(D) + d_bytes);
    return hip_status(launch_plain(inv ? PLAIN_G1_IFT : PLAIN_G1_FFT, W(out), W(in), nullptr, count, log2u(n),
                                   reinterpret_cast<uint32_t *>(D), nullptr, nullptr, flags, st->tables, st->prog,
                                   S(stream)));
}

int g1p_fft(fk20_g1p_t *output, const fk20_g1p_t *input, unsigned count, unsigned n, fk20_stream_t stream) {
    return g1_transform(output, input, count, n, false, stream);
}

int g1p_ift(fk20_g1p_t *output, const fk20_g1p_t *input, unsigned count, unsigned n, fk20_stream_t stream) {
    return g1_transform(output, input, count, n, true, stream);
}

// MSM with the bases converted once per call into packed Montgomery form in scratch,
// followed there by the scalars' Booth digits.
int fk20_msm(fk20_g1p_t *hext_fft, const fk20_fr_t *tc_fft, const fk20_g1p_t *xext_fft, unsigned rows,
             unsigned n2, fk20_stream_t stream) {
    if (!valid_n(n2) || (rows && (!hext_fft || !tc_fft || !xext_fft))) return FK20_EINVAL;
    if (!rows) return FK20_OK;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    const size_t x_bytes = (size_t)16 * n2 * 144;
    void *scratch;
    rc = device_scratch(st, x_bytes + msm_plain_digit_bytes(rows, n2), &scratch);
    if (rc) return rc;
    uint32_t *xpk = reinterpret_cast<uint32_t *>(scratch);
    int8_t *digits = reinterpret_cast<int8_t *>(scratch) + x_bytes;
    hipStream_t s = S(stream);
    FK20_TRY(launch_std_to_packed(xpk, W(xext_fft), (size_t)16 * n2, s));
    return hip_status(launch_msm_plain(W(hext_fft), W(tc_fft), xpk, digits, rows, n2, s));
}

int fk20_hext2h(fk20_g1p_t *h, unsigned rows, unsigned n2, fk20_stream_t stream) {
    if (!valid_n(n2) || (rows && !h)) return FK20_EINVAL;
    return hip_status(launch_hext2h(W(h), rows, n2, S(stream)));
}

int fk20_hext_fft2h_fft(fk20_g1p_t *h_fft, const fk20_g1p_t *hext_fft, unsigned rows, unsigned n2,
                        fk20_stream_t stream) {
    if (!valid_n(n2) || (rows && (!h_fft || !hext_fft))) return FK20_EINVAL;
    if (!rows) return FK20_OK;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    void *D;
    const size_t d_bytes = (size_t)rows * n2 * 144;
    rc = device_scratch(st, d_bytes + (size_t)rows * 4, &D);
    if (rc) return rc;
    uint32_t *flags = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(D) + d_bytes);
    return hip_status(launch_plain(PLAIN_HEXT_FFT2H_FFT, W(h_fft), W(hext_fft), nullptr, rows, log2u(n2),
                                   reinterpret_cast<uint32_t *>(D), nullptr, nullptr, flags, st->tables, st->prog,
                                   S(stream)));
}

// fk20_poly2h_fft.cu:47-80: toeplitz -> NTT (16 per row) -> MSM -> IFFT -> zero -> FFT, with
// the 1/n2 of the IFFT moved onto the NTT output (a field multiplication instead of a
// scalar multiplication per point) and the transforms column-batched over the rows.
// Workspace: tcT, D, the packed bases, the subgroup flag of the bases (one 256-byte slot)
// and the MSM digits, in that order.
static size_t poly2h_scratch_bytes(unsigned rows, unsigned n2) {
    return (size_t)rows * 16 * n2 * 32 + (size_t)rows * n2 * 144 + (size_t)16 * n2 * 144 + 256 +
           msm_plain_digit_bytes(rows, n2);
}

static int poly2h_fft_in(void *scratch, DeviceState *st, fk20_g1p_t *h_fft, const fk20_fr_t *poly,
                         const fk20_g1p_t *xext_fft, unsigned rows, unsigned n2, fk20_stream_t stream) {
    const size_t tc_bytes = (size_t)rows * 16 * n2 * 32;
    const size_t d_bytes = (size_t)rows * n2 * 144;
    const size_t x_bytes = (size_t)16 * n2 * 144;
    char *base = reinterpret_cast<char *>(scratch);
    uint32_t *tcT = reinterpret_cast<uint32_t *>(base);
    uint32_t *D = reinterpret_cast<uint32_t *>(base + tc_bytes);
    uint32_t *xpk = reinterpret_cast<uint32_t *>(base + tc_bytes + d_bytes);
    uint32_t *flag = reinterpret_cast<uint32_t *>(base + tc_bytes + d_bytes + x_bytes);
    int8_t *digits = reinterpret_cast<int8_t *>(base + tc_bytes + d_bytes + x_bytes + 256);
    hipStream_t s = S(stream);
    // one flag for all rows: a base outside G1 can put every row's MSM output outside it
    FK20_TRY(launch_subgroup_groups(flag, W(xext_fft), (size_t)16 * n2, (size_t)16 * n2, s));
    FK20_TRY(launch_std_to_packed(xpk, W(xext_fft), (size_t)16 * n2, s));
    return hip_status(launch_plain(PLAIN_POLY2H_FFT, W(h_fft), W(poly), xpk, rows, log2u(n2), D, tcT, digits,
                                   flag, st->tables, st->prog, s));
}

int fk20_poly2h_fft(fk20_g1p_t *h_fft, const fk20_fr_t *poly, const fk20_g1p_t *xext_fft, unsigned rows,
                    unsigned N, fk20_stream_t stream) {
    if (!valid_N(N) || (rows && (!h_fft || !poly || !xext_fft))) return FK20_EINVAL;
    if (!rows) return FK20_OK;
    const unsigned n2 = N / 8;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    void *scr;
    rc = device_scratch(st, poly2h_scratch_bytes(rows, n2), &scr);
    if (rc) return rc;
    return poly2h_fft_in(scr, st, h_fft, poly, xext_fft, rows, n2, stream);
}

// The same launch sequence with its workspace in this device's fixed `scratch` array: no
// allocation (the first call on a device copies the constant tables into static device
// memory, which allocates nothing either).  Calls that overlap on one device race on the
// array; nothing in it needs clearing (every byte is written before it is read).
int fk20_poly2h_fft_noalloc(fk20_g1p_t *h_fft, const fk20_fr_t *poly, const fk20_g1p_t *xext_fft, unsigned rows,
                            unsigned N, fk20_stream_t stream) {
    if (!valid_N(N) || (rows && (!h_fft || !poly || !xext_fft))) return FK20_EINVAL;
    if (!rows) return FK20_OK;
    const unsigned n2 = N / 8;
    if (poly2h_scratch_bytes(rows, n2) > NOALLOC_SCRATCH_BYTES) return FK20_ERANGE;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    return poly2h_fft_in(st->noalloc, st, h_fft, poly, xext_fft, rows, n2, stream);
}

int fk20_setup2xext_fft(fk20_g1p_t *xext_fft, const fk20_g1p_t *setup, unsigned N, fk20_stream_t stream) {
    if (!valid_N(N) || !xext_fft || !setup) return FK20_EINVAL;
    DeviceState *st;
    int rc = device_state(&st);
    if (rc) return rc;
    void *D;
    const size_t d_bytes = (size_t)16 * (N / 8) * 144;
    rc = device_scratch(st, d_bytes + 16 * 4, &D);
    if (rc) return rc;
    uint32_t *flags = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(D) + d_bytes);
    hipStream_t s = S(stream);
    FK20_TRY(launch_setup_gather(W(xext_fft), W(setup), N, s));
    return hip_status(launch_plain(PLAIN_G1_FFT, W(xext_fft), W(xext_fft), nullptr, 16, log2u(N / 8),
                                   reinterpret_cast<uint32_t *>(D), nullptr, nullptr, flags, st->tables, st->prog,
                                   s));
}

// ---------------------------------------------------------------- helpers

int fk20_mad_probe(uint64_t *out, unsigned blocks, unsigned iters, fk20_stream_t stream) {
    if (!out || !blocks) return FK20_EINVAL;
    return hip_status(launch_mad_probe(out, blocks, iters, S(stream)));
}

int fk20_fp_mul_probe(int limbs, uint32_t *out, const uint32_t *xs, const uint32_t *y, unsigned blocks,
                      unsigned iters, fk20_stream_t stream) {
    if ((limbs != 29 && limbs != 32) || !out || !xs || !y || !blocks) return FK20_EINVAL;
    return hip_status(launch_fp_mul_probe(limbs, out, xs, y, blocks, iters, S(stream)));
}

int fk20_g1p_to_affine(fk20_g1a_t *out, const fk20_g1p_t *in, size_t count, fk20_stream_t stream) {
    if (count && (!out || !in)) return FK20_EINVAL;
    return hip_status(launch_to_affine(W(out), W(in), count, S(stream)));
}

int fk20_g1p_compress(uint8_t *out, const fk20_g1p_t *in, size_t count, fk20_stream_t stream) {
    if (count && (!out || !in)) return FK20_EINVAL;
    return hip_status(launch_compress(W(out), W(in), count, S(stream)));
}

int fk20_fp_op(int op, uint64_t *out, const uint64_t *a, const uint64_t *b, size_t count, fk20_stream_t stream) {
    if (op < 0 || op > 6 || (count && (!out || !a || (op <= 2 && !b)))) return FK20_EINVAL;
    return hip_status(launch_fp_op(op, W(out), W(a), W(b), count, S(stream)));
}

int fk20_fr_op(int op, fk20_fr_t *out, const fk20_fr_t *a, const fk20_fr_t *b, size_t count,
               fk20_stream_t stream) {
    if (op < 0 || op > 6 || (count && (!out || !a || (op <= 2 && !b)))) return FK20_EINVAL;
    return hip_status(launch_fr_op(op, W(out), W(a), W(b), count, S(stream)));
}

int fk20_g1p_op(int op, fk20_g1p_t *out, const fk20_g1p_t *p, const void *q, const fk20_fr_t *k, size_t count,
                fk20_stream_t stream) {
    if (op < 0 || op > 7 || (count && (!out || !p))) return FK20_EINVAL;
    const bool needs_q = op == 0 || op == 1 || op == 5 || op == 7;
    if (count && ((needs_q && !q) || (op == 3 && !k))) return FK20_EINVAL;
    return hip_status(launch_g1_op(op, W(out), W(p), W(q), W(k), count, S(stream)));
}

int fk20_device_count(int *count) {
    if (!count) return FK20_EINVAL;
    return hip_status(hipGetDeviceCount(count));
}

int fk20_set_device(int device) { return hip_status(hipSetDevice(device)); }

int fk20_malloc(void **ptr, size_t bytes) {
    if (!ptr) return FK20_EINVAL;
    return hip_status(hipMalloc(ptr, bytes));
}

int fk20_free(void *ptr) { return hip_status(hipFree(ptr)); }

int fk20_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int fk20_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int fk20_memset(void *dst, int value, size_t bytes) { return hip_status(hipMemset(dst, value, bytes)); }

int fk20_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
}

int fk20_synchronize(void) { return hip_status(hipDeviceSynchronize()); }

int fk20_stream_create(fk20_stream_t *stream) {
    if (!stream) return FK20_EINVAL;
    hipStream_t s;
    int rc = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (rc == FK20_OK) *stream = s;
    return rc;
}

int fk20_stream_destroy(fk20_stream_t stream) { return hip_status(hipStreamDestroy(S(stream))); }

int fk20_stream_synchronize(fk20_stream_t stream) { return hip_status(hipStreamSynchronize(S(stream))); }

int fk20_event_create(fk20_event_t *event) {
    if (!event) return FK20_EINVAL;
    hipEvent_t e;
    int rc = hip_status(hipEventCreate(&e));
    if (rc == FK20_OK) *event = e;
    return rc;
}

int fk20_event_destroy(fk20_event_t event) { return hip_status(hipEventDestroy(reinterpret_cast<hipEvent_t>(event))); }

int fk20_event_record(fk20_event_t event, fk20_stream_t stream) {
    return hip_status(hipEventRecord(reinterpret_cast<hipEvent_t>(event), S(stream)));
}

int fk20_event_elapsed(float *ms, fk20_event_t start, fk20_event_t stop) {
    if (!ms) return FK20_EINVAL;
    FK20_TRY(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
    return hip_status(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
}

}  // extern "C"
