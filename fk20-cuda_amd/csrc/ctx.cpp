// Context API of the batched FK20 path: resident setups (window tables), twiddle programs,
// workspace and optional hipGraph replay.  See include/fk20.h for the interface.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/fk20.h"
#include "fast.hip.h"

using namespace fk20;

namespace {

typedef unsigned __int128 u128;

// ---------------------------------------------------------------- twiddle programs

// k = q * lambda + rem for a 256-bit k (8 words); q, rem < 2^128 because k < 2^255 and
// lambda > 2^127.
void divmod_lambda(const uint32_t *k, u128 &q, u128 &rem) {
    const u128 lam = ((u128)GLV_LAMBDA[1] << 64) | GLV_LAMBDA[0];
    u128 r = 0;
    q = 0;
    for (int i = 255; i >= 0; i--) {
        const bool top = (r >> 127) != 0;
        r = (r << 1) | ((k[i / 32] >> (i % 32)) & 1u);
        q <<= 1;
        if (top || r >= lam) {
            r -= lam;
            q |= 1;
        }
    }
    rem = r;
}

// width-W non-adjacent form, least significant digit first; digits odd in (-2^(W-1), 2^(W-1))
std::vector<int> wnaf(u128 k, int W) {
    std::vector<int> d;
    const int full = 1 << W, half = 1 << (W - 1);
    while (k != 0) {
        int digit = 0;
        if (k & 1) {
            digit = (int)(k & (u128)(full - 1));
            if (digit >= half) digit -= full;
            if (digit > 0) k -= (u128)digit;
            else k += (u128)(-digit);
        }
        d.push_back(digit);
        k >>= 1;
    }
    return d;
}

// Ladder cost of a program over digit strings d1, d2, in 32 x 32-bit multiply-adds of the
// 14 x 29-bit field arithmetic (a product 392, a square 301, a sum of two products sharing
// one reduction 588): one Jacobian doubling (3 squares, 2 products, 1 pair: 2275) per position
// below the top, one mixed addition (4 squares, 5 products, 1 pair: 3752) per non-zero digit
// after the first, and one product (x times beta or beta^2) per digit on phi(P) or phi^2(P).
int program_cost(const std::vector<int> &d1, const std::vector<int> &d2, int t) {
    const int L = (int)std::max(d1.size(), d2.size());
    int nz = 0, nphi = 0;
    for (int x : d1) nz += x != 0;
    for (int x : d2) nz += x != 0;
    for (int part = 0; part < 2; part++)
        if ((t + part) % 3 != 0)
            for (int x : part ? d2 : d1) nphi += x != 0;
    return std::max(L - 1, 0) * 2275 + std::max(nz - 1, 0) * 3752 + nphi * 392;
}

// lambda (the eigenvalue of phi on G1) in Montgomery form
fs lambda_mont() {
    fs lam;
    fs_zero(lam);
    lam.l[0] = (uint32_t)GLV_LAMBDA[0];
    lam.l[1] = (uint32_t)(GLV_LAMBDA[0] >> 32);
    lam.l[2] = (uint32_t)GLV_LAMBDA[1];
    lam.l[3] = (uint32_t)(GLV_LAMBDA[1] >> 32);
    fs_to_mont(lam, lam);
    return lam;
}

bool fs_equal(const fs &a, const fs &b) { return memcmp(a.l, b.l, 32) == 0; }

// Evaluates a program on the scalar 1 with phi = lambda (Montgomery Fr): it must give k, and
// the ladder's incomplete formulas (ladder.hip.h) must meet no exceptional case for an input
// of order r: before each mixed addition the accumulator a P and the entry b P satisfy
// a != 0 and a != +-b (mod r).  The table's co-Z additions 2P + (2i - 1)P are never
// exceptional (|2 -+ (2i - 1)| < 2^W is not a multiple of r).
bool check_program(const uint32_t *ops, const fs &expect_mont) {
    fs lam, lam2, acc, one, zero, neg;
    lam = lambda_mont();
    fs_mul(lam2, lam, lam);
    fs_set(one, FR_ONE);
    fs_zero(acc);
    fs_zero(zero);
    bool started = false;
    for (int i = 0;; i++) {
        const uint32_t op = ops[i];
        for (uint32_t d = 0; d < (op & TWOP_DBL_MASK); d++) fs_add(acc, acc, acc);
        if (op & TWOP_END) break;
        fs term, m;
        fs_zero(m);
        m.l[0] = 2 * ((op >> TWOP_ENTRY_SHIFT) & 7) + 1;
        fs_to_mont(term, m);
        if (op & TWOP_PHI2) fs_mul(term, term, lam2);
        else if (op & TWOP_PHI) fs_mul(term, term, lam);
        if (op & TWOP_NEG) fs_sub(term, zero, term);
        if (started) {
            fs_sub(neg, zero, term);
            if (fs_equal(acc, zero) || fs_equal(acc, term) || fs_equal(acc, neg)) return false;
        }
        fs_add(acc, acc, term);
        started = true;
    }
    return started && fs_equal(acc, expect_mont);
}

// k * P as a program over the wNAF digits of a split of k; appends it to `ops` and returns
// true, or returns false if no split passes check_program.  Twelve splits are candidates:
// k lambda^-t for t = 0, 1, 2 (then k1's digits act on phi^t(P) and k2's on phi^(t+1)(P)),
// each as itself or as r minus it (then every digit negated, since (r - k) P = -k P), each
// with 0 <= k1 < lambda or with k1 - lambda and k2 + 1; the cheapest that passes is kept.
// The choice matters most for w^-128 = -w^128 = -(z lambda + z), whose halves are 64-bit.
bool append_program(std::vector<uint32_t> &ops, const uint32_t *k, int W, const fs &expect_mont) {
    const u128 lam = ((u128)GLV_LAMBDA[1] << 64) | GLV_LAMBDA[0];
    const fs lam_m = lambda_mont();
    fs lam_inv;  // lambda^-1 = lambda^2
    fs_mul(lam_inv, lam_m, lam_m);
    fs km, kt;
    fs_zero(km);
    for (int i = 0; i < 8; i++) km.l[i] = k[i];
    fs_to_mont(km, km);
    struct Cand {
        int cost, t;
        std::vector<int> d1, d2;
    };
    std::vector<Cand> cands;
    kt = km;
    for (int t = 0; t < 3; t++) {
        if (t) fs_mul(kt, kt, lam_inv);
        fs ks;
        fs_from_mont(ks, kt);
        for (int neg = 0; neg < 2; neg++) {
            uint32_t kk[8];
            if (neg) {  // r - k (k = 0 has no negation to try)
                bool zero = true;
                for (int i = 0; i < 8; i++) zero = zero && ks.l[i] == 0;
                if (zero) continue;
                uint64_t borrow = 0;
                for (int i = 0; i < 8; i++) {
                    const uint64_t d = (uint64_t)FR_R[i] - ks.l[i] - borrow;
                    kk[i] = (uint32_t)d;
                    borrow = (d >> 63) & 1;
                }
            } else {
                memcpy(kk, ks.l, 32);
            }
            u128 k2, k1;
            divmod_lambda(kk, k2, k1);
            for (int shifted = 0; shifted < 2; shifted++) {
                // shifted: k1 - lambda < 0 and k2 + 1
                if (shifted && k1 == 0) continue;
                Cand c;
                c.t = t;
                c.d1 = wnaf(shifted ? lam - k1 : k1, W);
                c.d2 = wnaf(shifted ? k2 + 1 : k2, W);
                if (shifted) for (int &x : c.d1) x = -x;
                if (neg) {
                    for (int &x : c.d1) x = -x;
                    for (int &x : c.d2) x = -x;
                }
                c.cost = program_cost(c.d1, c.d2, t);
                cands.push_back(c);
            }
        }
    }
    std::stable_sort(cands.begin(), cands.end(), [](const Cand &a, const Cand &b) { return a.cost < b.cost; });
    for (const Cand &c : cands) {
        std::vector<uint32_t> prog;
        const int L = (int)std::max(c.d1.size(), c.d2.size());
        bool started = false;
        uint32_t pending = 0;
        for (int t = L - 1; t >= 0; t--) {
            if (started) pending++;
            for (int part = 0; part < 2; part++) {
                const std::vector<int> &d = part ? c.d2 : c.d1;
                const int digit = t < (int)d.size() ? d[t] : 0;
                if (!digit) continue;
                const uint32_t e = (uint32_t)((digit < 0 ? -digit : digit) - 1) / 2;
                const int power = (c.t + part) % 3;  // phi^power of P
                const uint32_t phi = power == 1 ? TWOP_PHI : power == 2 ? TWOP_PHI2 : 0u;
                prog.push_back(pending | phi | (e << TWOP_ENTRY_SHIFT) | (digit < 0 ? TWOP_NEG : 0u));
                pending = 0;
                started = true;
            }
        }
        prog.push_back(TWOP_END | pending);
        if (check_program(prog.data(), expect_mont)) {
            ops.insert(ops.end(), prog.begin(), prog.end());
            return true;
        }
    }
    return false;
}

// Builds the programs of w^idx and w^-idx (w of order 512), each checked on the host.
bool build_programs_host(std::vector<uint32_t> &ops, std::vector<uint32_t> &index, int W = TW_WNAF) {
    DevTables T;
    build_host_tables(T);
    ops.clear();
    index.assign(2 * 512, 0);
    for (int dir = 0; dir < 2; dir++) {
        for (int idx = 0; idx < 512; idx++) {
            const int e = dir == TW_FWD ? idx : (512 - idx) % 512;
            index[dir * 512 + idx] = (uint32_t)ops.size();
            fs expect;
            fs_set(expect, T.tw_mont[e]);
            if (!append_program(ops, T.tw_std[e], W, expect)) return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------- context

struct Setup {
    bool live = false;
    bool commit = false;  // an SRS for KZG commitments (fk20_commit_setup_load)
    unsigned N = 0, n2 = 0, win = 0, nwin = 0;
    uint32_t *table = nullptr;
    uint32_t *inf = nullptr;
};

bool valid_N(unsigned N) { return N == 1024 || N == 2048 || N == 4096; }
bool valid_win(unsigned w) { return w >= 8 && w <= 13; }
unsigned nwin_of(unsigned w) { return (256 + w - 1) / w; }

int status_of(hipError_t e) { return status_from_hip(e); }

}  // namespace

namespace {
// Every program (w^+-i, then 2^-j) on the host, each checked, with its index table.
// Ladder cost of an encoded program (the units of program_cost): its doublings, its mixed
// additions after the first entry and its endomorphism products.
int ops_cost(const uint32_t *ops) {
    int dbl = 0, add = 0, phi = 0;
    for (unsigned k = 0;; k++) {
        dbl += (int)(ops[k] & TWOP_DBL_MASK);
        if (ops[k] & TWOP_END) break;
        add++;
        phi += (ops[k] & (TWOP_PHI | TWOP_PHI2)) != 0;
    }
    return dbl * 2275 + std::max(add - 1, 0) * 3752 + phi * 392;
}

// The twiddle layers' item tables (index[TW_LAYER ..]): every (q, e) of layer_twiddle, per
// size, stage and direction, sorted by decreasing cost of e's program (stable: ties keep the
// closed-form order).
bool layer_tables_host(const std::vector<uint32_t> &ops, std::vector<uint32_t> &index) {
    for (unsigned L = 7; L <= 9; L++)
        for (unsigned sd = 1; sd < L; sd += 2)
            for (unsigned dir = 0; dir < 2; dir++) {
                const unsigned cnt = layer_count(sd, L), off = layer_table_offset(L, sd, dir);
                if (off + cnt > (unsigned)TW_LAYER_SIZE) return false;
                std::vector<std::pair<int, uint32_t>> items;
                for (unsigned k = 0; k < cnt; k++) {
                    const LayerTw t = layer_twiddle(sd, L, k);
                    items.push_back({-ops_cost(ops.data() + index[dir * 512 + t.e]), t.q | (t.e << 16)});
                }
                std::stable_sort(items.begin(), items.end(),
                                 [](const std::pair<int, uint32_t> &a, const std::pair<int, uint32_t> &b) {
                                     return a.first < b.first;
                                 });
                for (unsigned k = 0; k < cnt; k++) index[TW_LAYER + off + k] = items[k].second;
            }
    return true;
}

bool all_programs_host(std::vector<uint32_t> &ops, std::vector<uint32_t> &index, uint32_t *scale_off,
                       int W = TW_WNAF) {
    if (!build_programs_host(ops, index, W)) return false;
    DevTables T;
    build_host_tables(T);
    index.resize(TW_INDEX_SIZE);
    for (int j = 0; j < 10; j++) {
        index[TW_SCALE + j] = scale_off[j] = (uint32_t)ops.size();
        fs expect;
        fs_set(expect, T.inv2k_mont[j]);
        if (!append_program(ops, T.inv2k_std[j], W, expect)) return false;
    }
    return layer_tables_host(ops, index);
}
}  // namespace

hipError_t fk20::build_programs(Programs &P, unsigned wnaf) {
    std::vector<uint32_t> ops, index;
    if (wnaf < 3 || wnaf > 5 || !all_programs_host(ops, index, P.scale_off, (int)wnaf)) return hipErrorInvalidValue;
    P.owned = true;
    P.wnaf = wnaf;
    hipError_t err = hipMalloc(&P.ops, ops.size() * 4);
    if (err == hipSuccess) err = hipMalloc(&P.index, index.size() * 4);
    if (err == hipSuccess) err = hipMemcpy(P.ops, ops.data(), ops.size() * 4, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(P.index, index.data(), index.size() * 4, hipMemcpyHostToDevice);
    if (err != hipSuccess) free_programs(P);
    return err;
}

hipError_t fk20::load_programs(Programs &P, uint32_t *ops_dev, size_t ops_cap, uint32_t *index_dev) {
    std::vector<uint32_t> ops, index;
    if (!all_programs_host(ops, index, P.scale_off) || ops.size() > ops_cap) return hipErrorInvalidValue;
    hipError_t err = hipMemcpy(ops_dev, ops.data(), ops.size() * 4, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(index_dev, index.data(), index.size() * 4, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        P.ops = ops_dev;
        P.index = index_dev;
        P.owned = false;
        P.wnaf = TW_WNAF;
    }
    return err;
}

void fk20::free_programs(Programs &P) {
    if (P.owned) {
        if (P.ops) (void)hipFree(P.ops);
        if (P.index) (void)hipFree(P.index);
    }
    P.ops = P.index = nullptr;
}

struct fk20_ctx {
    int device = 0;
    const DevTables *T = nullptr;
    Programs prog;
    std::vector<Setup> setups;
    void *work = nullptr;
    size_t work_bytes = 0;
    bool use_graph = false;
    bool use_chain = false;  // transform stages as one persistent launch (k_chain)
    uint64_t chain_wait = CHAIN_WAIT_TICKS;
    // control blocks of the persistent launches of the last enqueued (or captured) pass: word 1
    // is set by a wave whose bounded wait expired
    std::vector<uint32_t *> chain_ctl;
    // device-side order of the calls that use `work`: `done` is recorded on the stream of the
    // last such call, and a call on another stream waits for it first
    hipEvent_t done = nullptr;
    hipStream_t done_stream = nullptr;
    bool done_valid = false;
    int out_format = FK20_OUT_G1P;
    // schedule of the transform kernels (fk20_ctx_set_schedule); prog is built for sch_wnaf
    unsigned sch_wnaf = TW_WNAF, sch_occ = DEFAULT_OCCUPANCY;
    bool sch_r22 = true;
    // fk20_prove_batch splits its rows into `nsub` halves at 64-row boundaries and runs each
    // half's launch sequence on its own stream, forked from and joined back to the caller's
    // stream, so one half's stage launches fill the SIMDs the other half's idle at a boundary
    unsigned nsub = 2;
    hipStream_t sub[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
    std::vector<fk20_batch_item_t> gkey;  // items of the captured graph
    hipStream_t gstream = nullptr;
    hipGraphExec_t gexec = nullptr;
    // per-launch timing: events[0] before the first launch, events[k] after launch k
    bool timing = false;
    std::vector<hipEvent_t> events;
    std::vector<int> kinds;
    unsigned nmarks = 0;
    hipStream_t tstream = nullptr;
    std::mutex mu;
};

namespace {
void mark_launch(void *arg, int kind) {
    fk20_ctx *c = static_cast<fk20_ctx *>(arg);
    if (c->nmarks + 1 >= c->events.size()) return;
    c->kinds[c->nmarks] = kind;
    (void)hipEventRecord(c->events[++c->nmarks], c->tstream);
}
}  // namespace

extern "C" {

size_t fk20_setup_table_bytes(unsigned N, unsigned window) {
    if (!valid_N(N) || !valid_win(window)) return 0;
    return (size_t)16 * (N / 8) * nwin_of(window) * ((size_t)1 << (window - 1)) * TABLE_ENTRY_WORDS * 4;
}

// Per item: tcT (16 n2 Fr per row), D (n2 points per row), the chain flags (n2 words per
// 64-row group) and 256-byte control blocks (the chain's work counter; item 0's is used):
// two of them, so that the two halves of a split batch fit in the whole's size.
size_t fk20_rows_workspace_bytes(unsigned rows, unsigned N) {
    if (!valid_N(N)) return 0;
    const size_t n2 = N / 8;
    return (size_t)rows * n2 * (16 * 32 + 144) + (size_t)((rows + 63) / 64) * n2 * 4 + 512;
}

int fk20_ctx_create(fk20_ctx_t **out) {
    if (!out) return FK20_EINVAL;
    fk20_ctx *c = new fk20_ctx();
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) c->T = current_device_tables(&e);
    if (e == hipSuccess) e = build_programs(c->prog, c->sch_wnaf);
    if (e != hipSuccess) {
        fk20_ctx_destroy(c);
        return status_of(e);
    }
    *out = c;
    return FK20_OK;
}

int fk20_ctx_destroy(fk20_ctx_t *c) {
    if (!c) return FK20_EINVAL;
    (void)hipDeviceSynchronize();
    for (Setup &s : c->setups) {
        if (s.table) (void)hipFree(s.table);
        if (s.inf) (void)hipFree(s.inf);
    }
    if (c->work) (void)hipFree(c->work);
    free_programs(c->prog);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; k++) {
        if (c->sub[k]) (void)hipStreamDestroy(c->sub[k]);
        if (c->join[k]) (void)hipEventDestroy(c->join[k]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->done) (void)hipEventDestroy(c->done);
    delete c;
    return FK20_OK;
}

size_t fk20_output_point_bytes(int format) {
    return format == FK20_OUT_G1P || format == FK20_OUT_G1P_DA ? 144
         : format == FK20_OUT_COMPRESSED || format == FK20_OUT_COMPRESSED_DA ? 48 : 0;
}

int fk20_ctx_set_output(fk20_ctx_t *c, int format) {
    if (!c || !fk20_output_point_bytes(format)) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->out_format != format && c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);  // the format is captured in the launch arguments
        c->gexec = nullptr;
        c->gkey.clear();
    }
    c->out_format = format;
    return FK20_OK;
}

int fk20_ctx_set_streams(fk20_ctx_t *c, unsigned n) {
    if (!c || n < 1 || n > 2) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->nsub != n && c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);  // captured with the other stream layout
        c->gexec = nullptr;
        c->gkey.clear();
    }
    c->nsub = n;
    return FK20_OK;
}

int fk20_ctx_set_schedule(fk20_ctx_t *c, unsigned wnaf, unsigned radix, unsigned waves_per_simd) {
    if (!c || !stage_kernels(wnaf, radix == 4, waves_per_simd) || (radix != 2 && radix != 4)) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (wnaf != c->sch_wnaf) {
        // queued passes may still read the old programs
        Programs np;
        hipError_t e = build_programs(np, wnaf);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            free_programs(np);
            return status_of(e);
        }
        free_programs(c->prog);
        c->prog = np;
    }
    if (c->gexec) {  // captured with the other kernels
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
        c->gkey.clear();
    }
    c->sch_wnaf = wnaf;
    c->sch_r22 = radix == 4;
    c->sch_occ = waves_per_simd;
    return FK20_OK;
}

int fk20_ctx_get_schedule(fk20_ctx_t *c, unsigned *wnaf, unsigned *radix, unsigned *waves_per_simd) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (wnaf) *wnaf = c->sch_wnaf;
    if (radix) *radix = c->sch_r22 ? 4 : 2;
    if (waves_per_simd) *waves_per_simd = c->sch_occ;
    return FK20_OK;
}

int fk20_ctx_get_output(fk20_ctx_t *c, int *format) {
    if (!c || !format) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    *format = c->out_format;
    return FK20_OK;
}

int fk20_ctx_set_chain(fk20_ctx_t *c, int enable) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->use_chain != (enable != 0) && c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);  // captured with the other launch sequence
        c->gexec = nullptr;
        c->gkey.clear();
    }
    c->use_chain = enable != 0;
    return FK20_OK;
}

int fk20_ctx_set_chain_wait(fk20_ctx_t *c, uint64_t ticks) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->chain_wait != ticks && c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);  // the bound is a captured launch argument
        c->gexec = nullptr;
        c->gkey.clear();
    }
    c->chain_wait = ticks;
    return FK20_OK;
}

int fk20_ctx_set_graph(fk20_ctx_t *c, int enable) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    c->use_graph = enable != 0;
    return FK20_OK;
}

static hipError_t grow_work(fk20_ctx *c, size_t bytes);

// Window tables of nbases points (standard-form g1p_t, device), each first multiplied by the
// standard-form scalar `scale` (device, 8 words).  Fills st.table / st.inf.  *outside: some
// base lies outside the order-r subgroup G1; *small_order: a base of small order outside G1
// gave an infinity entry, which the affine tables cannot hold (nothing is kept then).
static hipError_t build_tables(Setup &st, const fk20_g1p_t *points, unsigned nbases, const uint32_t *scale,
                               hipStream_t s, bool *outside, bool *small_order) {
    *outside = *small_order = false;
    const unsigned H = 1u << (st.win - 1);
    uint32_t *xs = nullptr, *zt = nullptr;
    // bases per build launch: one thread per (base, window), batch-inversion scratch bounded
    // by 4 GiB (every base of an N = 4096 setup in one launch up to w = 13)
    const size_t per_base = (size_t)st.nwin * H * TABLE_ENTRY_WORDS * 4;
    const size_t zt_per_base = table_build_scratch_bytes(1, st.win, st.nwin);
    unsigned chunk = (unsigned)std::max<size_t>(1, ((size_t)4 << 30) / zt_per_base);
    if (chunk > nbases) chunk = nbases;
    hipError_t e = hipMalloc(&st.table, (size_t)nbases * per_base);
    // inf[0 .. nbases), then two words: outside G1, table-build error
    if (e == hipSuccess) e = hipMalloc(&st.inf, (size_t)nbases * 4 + 8);
    if (e == hipSuccess) e = hipMalloc(&xs, (size_t)nbases * 144);
    if (e == hipSuccess) e = hipMalloc(&zt, table_build_scratch_bytes(chunk, st.win, st.nwin));
    uint32_t *words = st.inf ? st.inf + nbases : nullptr;
    if (e == hipSuccess) e = hipMemsetAsync(words, 0, 8, s);
    if (e == hipSuccess)
        e = launch_subgroup_groups(words, reinterpret_cast<const uint32_t *>(points), nbases, nbases, s);
    if (e == hipSuccess)
        e = launch_scale_bases(xs, st.inf, reinterpret_cast<const uint32_t *>(points), nbases, scale, s);
    for (unsigned b0 = 0; e == hipSuccess && b0 < nbases; b0 += chunk)
        e = launch_table_build(st.table, zt, xs, st.inf, words + 1, b0, std::min(chunk, nbases - b0), st.win,
                               st.nwin, s);
    uint32_t host_words[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(host_words, words, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) {
        *outside = host_words[0] != 0;
        *small_order = host_words[1] != 0;
    }
    if (xs) (void)hipFree(xs);
    if (zt) (void)hipFree(zt);
    if (e != hipSuccess || *small_order) {
        if (st.table) (void)hipFree(st.table);
        if (st.inf) (void)hipFree(st.inf);
        st.table = st.inf = nullptr;
    }
    return e;
}

int fk20_setup_load(fk20_ctx_t *c, const fk20_g1p_t *xext_fft, unsigned N, unsigned window, int *setup_id,
                    fk20_stream_t stream) {
    if (!c || !xext_fft || !setup_id || !valid_N(N) || !valid_win(window)) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    Setup st;
    st.N = N;
    st.n2 = N / 8;
    st.win = window;
    st.nwin = nwin_of(window);
    const unsigned logn2 = st.n2 == 128 ? 7 : st.n2 == 256 ? 8 : 9;
    // the bases carry the IFFT's 1/n2.  Every base must lie in G1: outside it the result
    // depends on the exact integer multipliers (fk20_poly2h_fft follows the reference's).
    bool outside, small_order;
    hipError_t e = build_tables(st, xext_fft, 16 * st.n2, c->T->inv2k_std[logn2],
                                reinterpret_cast<hipStream_t>(stream), &outside, &small_order);
    if (e != hipSuccess) return status_of(e);
    if (outside) {
        (void)hipFree(st.table);
        (void)hipFree(st.inf);
    }
    if (outside || small_order) return FK20_EINVAL;
    st.live = true;
    c->setups.push_back(st);
    *setup_id = (int)c->setups.size() - 1;
    return FK20_OK;
}

int fk20_commit_setup_load(fk20_ctx_t *c, const fk20_g1p_t *srs, unsigned N, unsigned window, int *setup_id,
                           fk20_stream_t stream) {
    if (!c || !srs || !setup_id || !valid_N(N) || !valid_win(window)) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    Setup st;
    st.commit = true;
    st.N = N;
    st.n2 = N / 8;
    st.win = window;
    st.nwin = nwin_of(window);
    // tw_std[0] = w^0 = 1 (standard form): the SRS points are taken as they are
    bool outside, small_order;  // commitments multiply each point by its coefficient only: any point will do
    hipError_t e = build_tables(st, srs, N, c->T->tw_std[0], reinterpret_cast<hipStream_t>(stream), &outside,
                                &small_order);
    if (e != hipSuccess) return status_of(e);
    if (small_order) return FK20_EINVAL;
    st.live = true;
    c->setups.push_back(st);
    *setup_id = (int)c->setups.size() - 1;
    return FK20_OK;
}

// Orders a call that uses the context workspace on stream s after the previous one.
static hipError_t order_before(fk20_ctx *c, hipStream_t s) {
    if (!c->done_valid || c->done_stream == s) return hipSuccess;
    return hipStreamWaitEvent(s, c->done, 0);
}

static hipError_t order_after(fk20_ctx *c, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (!c->done) e = hipEventCreateWithFlags(&c->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(c->done, s);
    c->done_valid = e == hipSuccess;
    c->done_stream = s;
    return e;
}

size_t fk20_commit_setup_table_bytes(unsigned N, unsigned window) {
    if (!valid_N(N) || !valid_win(window)) return 0;
    return (size_t)N * nwin_of(window) * ((size_t)1 << (window - 1)) * TABLE_ENTRY_WORDS * 4;
}

size_t fk20_commit_workspace_bytes(unsigned rows, unsigned N) {
    if (!valid_N(N)) return 0;
    return commit_workspace_bytes(rows, N);
}

int fk20_commit_rows(fk20_ctx_t *c, int setup_id, fk20_g1p_t *commitments, const fk20_fr_t *polynomial,
                     unsigned rows, fk20_stream_t stream) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (setup_id < 0 || setup_id >= (int)c->setups.size() || !c->setups[setup_id].live ||
        !c->setups[setup_id].commit)
        return FK20_EINVAL;
    if (!rows) return FK20_OK;
    if (!commitments || !polynomial) return FK20_EINVAL;
    const Setup &st = c->setups[setup_id];
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = grow_work(c, commit_workspace_bytes(rows, st.N));
    if (e == hipSuccess) e = order_before(c, s);
    if (e == hipSuccess)
        e = launch_commit(reinterpret_cast<uint32_t *>(commitments), reinterpret_cast<const uint32_t *>(polynomial),
                          st.table, st.inf, reinterpret_cast<uint32_t *>(c->work), rows, st.N, st.win, s);
    if (e == hipSuccess) e = order_after(c, s);
    return status_of(e);
}

int fk20_setup_unload(fk20_ctx_t *c, int id) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (id < 0 || id >= (int)c->setups.size() || !c->setups[id].live) return FK20_EINVAL;
    (void)hipDeviceSynchronize();
    Setup &s = c->setups[id];
    (void)hipFree(s.table);
    (void)hipFree(s.inf);
    s = Setup();
    if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
        c->gkey.clear();
    }
    return FK20_OK;
}

static hipError_t grow_work(fk20_ctx *c, size_t bytes) {
    if (c->work_bytes >= bytes) return hipSuccess;
    if (c->work) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return e;
        (void)hipFree(c->work);
        c->work = nullptr;
        c->work_bytes = 0;
        if (c->gexec) {
            (void)hipGraphExecDestroy(c->gexec);
            c->gexec = nullptr;
        }
    }
    hipError_t e = hipMalloc(&c->work, bytes);
    if (e == hipSuccess) c->work_bytes = bytes;
    return e;
}

int fk20_ctx_reserve(fk20_ctx_t *c, unsigned rows, unsigned N) {
    if (!c || !valid_N(N)) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    return status_of(grow_work(c, fk20_rows_workspace_bytes(rows, N)));
}

static size_t items_workspace_bytes(const fk20_ctx *c, const fk20_batch_item_t *items, unsigned n) {
    size_t total = 0;
    for (unsigned k = 0; k < n; k++) total += fk20_rows_workspace_bytes(items[k].rows, c->setups[items[k].setup_id].N);
    return total;
}

static bool same_items(const std::vector<fk20_batch_item_t> &a, const fk20_batch_item_t *b, unsigned n) {
    if (a.size() != n) return false;
    for (unsigned k = 0; k < n; k++)
        if (a[k].setup_id != b[k].setup_id || a[k].rows != b[k].rows || a[k].polynomial != b[k].polynomial ||
            a[k].h_fft != b[k].h_fft)
            return false;
    return true;
}

// Carves each item's workspace (tcT, D, chain flags, control block) from w and launches the
// batch on s.
static hipError_t enqueue_launch(fk20_ctx *c, const fk20_batch_item_t *items, unsigned n, hipStream_t s,
                                 LaunchHook hook, char *w, bool chain) {
    Launch L;
    memset(&L, 0, sizeof L);
    L.nitems = n;
    uint32_t *ctl = nullptr;
    for (unsigned k = 0; k < n; k++) {
        const Setup &st = c->setups[items[k].setup_id];
        Item &it = L.item[k];
        const size_t rows = items[k].rows;
        it.poly = reinterpret_cast<const uint32_t *>(items[k].polynomial);
        it.h_fft = reinterpret_cast<uint32_t *>(items[k].h_fft);
        it.tcT = reinterpret_cast<uint32_t *>(w);
        w += rows * 16 * st.n2 * 32;
        it.D = reinterpret_cast<uint32_t *>(w);
        w += rows * st.n2 * 144;
        it.flags = reinterpret_cast<uint32_t *>(w);
        w += (rows + 63) / 64 * st.n2 * 4;
        if (!k) ctl = reinterpret_cast<uint32_t *>(w);
        w += 256;  // one control block carved per item: fk20_rows_workspace_bytes reserves two
        it.table = st.table;
        it.inf = st.inf;
        it.rows = items[k].rows;
        it.N = st.N;
        it.logn2 = st.n2 == 128 ? 7 : st.n2 == 256 ? 8 : 9;
        it.win = st.win;
        it.out_format = (uint32_t)c->out_format;
        it.stage = -1;
    }
    if (chain) c->chain_ctl.push_back(ctl);
    // the persistent form runs radix-2 stages: the schedule's width and occupancy, radix 2
    const StageKernels *K = stage_kernels(c->sch_wnaf, chain ? false : c->sch_r22, c->sch_occ);
    if (!K) return hipErrorInvalidValue;
    return launch_batch(L, c->T, c->prog.ops, c->prog.index, *K, s, hook, chain ? ctl : nullptr, c->chain_wait);
}

static hipError_t enqueue_items(fk20_ctx *c, const fk20_batch_item_t *items, unsigned n, hipStream_t s,
                                LaunchHook hook, void *workspace = nullptr) {
    // the persistent form only on the context's own workspace (fk20_prove_batch reads its
    // error words back; the serving loop's fk20_prove_batch_ws does not wait for its batches)
    const bool chain = c->use_chain && !workspace;
    char *w = reinterpret_cast<char *>(workspace ? workspace : c->work);
    unsigned groups = 0;
    for (unsigned k = 0; k < n; k++) groups += (items[k].rows + 63) / 64;
    // the context path with at least two 64-row groups, launches not timed one by one
    if (workspace || c->nsub < 2 || groups < 2 || hook.fn) return enqueue_launch(c, items, n, s, hook, w, chain);
    // halves: each item's rows split at a 64-row boundary (items of one group go to the first)
    fk20_batch_item_t half[2][MAX_ITEMS];
    unsigned nh[2] = {0, 0};
    for (unsigned k = 0; k < n; k++) {
        const unsigned rg = (items[k].rows + 63) / 64, r0 = std::min(items[k].rows, (rg + 1) / 2 * 64);
        const size_t pbytes = (size_t)c->setups[items[k].setup_id].N * 32;
        const size_t obytes = (size_t)(c->setups[items[k].setup_id].N / 8) * fk20_output_point_bytes(c->out_format);
        fk20_batch_item_t a = items[k], b = items[k];
        a.rows = r0;
        b.rows = items[k].rows - r0;
        b.polynomial = reinterpret_cast<const fk20_fr_t *>(reinterpret_cast<const char *>(items[k].polynomial) + r0 * pbytes);
        b.h_fft = reinterpret_cast<fk20_g1p_t *>(reinterpret_cast<char *>(items[k].h_fft) + r0 * obytes);
        half[0][nh[0]++] = a;
        if (b.rows) half[1][nh[1]++] = b;
    }
    if (!nh[1]) return enqueue_launch(c, items, n, s, hook, w, chain);
    // the halves' workspaces: a 64-row split keeps the flag words, and each half carves one
    // 256-byte control block of the two the whole item reserves, so both fit in the whole's size
    size_t off = 0;  // bytes enqueue_launch carves for the first half (256 below the reserved size per item)
    for (unsigned k = 0; k < nh[0]; k++)
        off += fk20_rows_workspace_bytes(half[0][k].rows, c->setups[half[0][k].setup_id].N) - 256;
    // The internal streams are created on the first split, not with the context: HIP spreads
    // streams over a few hardware queues (4 per process here), and two idle streams created
    // up front took queues that the server's two slot streams then had to share -- serving
    // throughput at saturation measured 8.8 K rows/s that way against 11.2 K created lazily.
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        if (!c->sub[k]) e = hipStreamCreateWithFlags(&c->sub[k], hipStreamNonBlocking);
        if (e == hipSuccess && !c->join[k]) e = hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming);
    }
    if (e == hipSuccess && !c->fork) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(c->fork, s);
    for (int h = 0; h < 2 && e == hipSuccess; h++) e = hipStreamWaitEvent(c->sub[h], c->fork, 0);
    if (e == hipSuccess) e = enqueue_launch(c, half[0], nh[0], c->sub[0], hook, w, chain);
    if (e == hipSuccess) e = enqueue_launch(c, half[1], nh[1], c->sub[1], hook, w + off, chain);
    for (int h = 0; h < 2 && e == hipSuccess; h++) {
        e = hipEventRecord(c->join[h], c->sub[h]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, c->join[h], 0);
    }
    return e;
}

int fk20_prove_batch(fk20_ctx_t *c, const fk20_batch_item_t *items, unsigned nitems, fk20_stream_t stream) {
    if (!c || (nitems && !items) || nitems > MAX_ITEMS) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    // drop empty items; validate the rest
    std::vector<fk20_batch_item_t> v;
    for (unsigned k = 0; k < nitems; k++) {
        const fk20_batch_item_t &it = items[k];
        if (it.setup_id < 0 || it.setup_id >= (int)c->setups.size() || !c->setups[it.setup_id].live ||
            c->setups[it.setup_id].commit)
            return FK20_EINVAL;
        if (!it.rows) continue;
        if (!it.polynomial || !it.h_fft) return FK20_EINVAL;
        v.push_back(it);
    }
    if (v.empty()) return FK20_OK;
    const unsigned n = (unsigned)v.size();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipError_t e = grow_work(c, items_workspace_bytes(c, v.data(), n));
    if (e == hipSuccess) e = order_before(c, s);
    if (e != hipSuccess) return status_of(e);
    if (c->timing) {
        // direct launches with an event after each one (read back by fk20_ctx_launch_times)
        c->tstream = s;
        c->nmarks = 0;
        c->chain_ctl.clear();
        e = hipEventRecord(c->events[0], s);
        if (e == hipSuccess) e = enqueue_items(c, v.data(), n, s, LaunchHook{mark_launch, c});
    } else if (!c->use_graph || s == nullptr) {
        c->chain_ctl.clear();
        e = enqueue_items(c, v.data(), n, s, LaunchHook{nullptr, nullptr});
    } else {
        // hipGraph: capture the launch sequence once per (items, stream), replay after
        if (!c->gexec || c->gstream != s || !same_items(c->gkey, v.data(), n)) {
            if (c->gexec) {
                (void)hipGraphExecDestroy(c->gexec);
                c->gexec = nullptr;
            }
            hipGraph_t g = nullptr;
            c->chain_ctl.clear();
            e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return status_of(e);
            hipError_t le = enqueue_items(c, v.data(), n, s, LaunchHook{nullptr, nullptr});
            e = hipStreamEndCapture(s, &g);
            if (le != hipSuccess) e = le;
            if (e == hipSuccess) e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) return status_of(e);
            c->gkey = v;
            c->gstream = s;
        }
        e = hipGraphLaunch(c->gexec, s);  // chain_ctl still holds the captured control blocks
    }
    if (e == hipSuccess) e = order_after(c, s);
    if (e != hipSuccess || !c->use_chain) return status_of(e);
    // persistent form: wait for the pass and report an expired wait
    e = hipStreamSynchronize(s);
    uint32_t expired = 0;
    for (uint32_t *ctl : c->chain_ctl) {
        uint32_t word = 0;
        if (e == hipSuccess) e = hipMemcpy(&word, ctl + 1, 4, hipMemcpyDeviceToHost);
        expired |= word;
    }
    if (e != hipSuccess) return status_of(e);
    return expired ? FK20_ETIMEDOUT : FK20_OK;
}

int fk20_prove_batch_ws(fk20_ctx_t *c, const fk20_batch_item_t *items, unsigned nitems, void *workspace,
                        size_t workspace_bytes, fk20_stream_t stream) {
    if (!c || (nitems && !items) || nitems > MAX_ITEMS || !workspace) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    std::vector<fk20_batch_item_t> v;
    for (unsigned k = 0; k < nitems; k++) {
        const fk20_batch_item_t &it = items[k];
        if (it.setup_id < 0 || it.setup_id >= (int)c->setups.size() || !c->setups[it.setup_id].live ||
            c->setups[it.setup_id].commit)
            return FK20_EINVAL;
        if (!it.rows) continue;
        if (!it.polynomial || !it.h_fft) return FK20_EINVAL;
        v.push_back(it);
    }
    if (v.empty()) return FK20_OK;
    if (items_workspace_bytes(c, v.data(), (unsigned)v.size()) > workspace_bytes) return FK20_ENOMEM;
    return status_of(enqueue_items(c, v.data(), (unsigned)v.size(), reinterpret_cast<hipStream_t>(stream),
                                   LaunchHook{nullptr, nullptr}, workspace));
}

int fk20_setup_info(fk20_ctx_t *c, int id, unsigned *N, unsigned *window) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (id < 0 || id >= (int)c->setups.size() || !c->setups[id].live) return FK20_EINVAL;
    if (N) *N = c->setups[id].N;
    if (window) *window = c->setups[id].win;
    return FK20_OK;
}

int fk20_prove_rows(fk20_ctx_t *c, int id, fk20_g1p_t *h_fft, const fk20_fr_t *poly, unsigned rows,
                    fk20_stream_t stream) {
    if (!c) return FK20_EINVAL;
    fk20_batch_item_t it;
    it.setup_id = id;
    it.rows = rows;
    it.polynomial = poly;
    it.h_fft = h_fft;
    return fk20_prove_batch(c, &it, 1, stream);
}

int fk20_ctx_set_launch_timing(fk20_ctx_t *c, int enable) {
    if (!c) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    if (enable && c->events.empty()) {
        c->events.resize(64, nullptr);
        c->kinds.assign(64, -1);
        for (hipEvent_t &e : c->events) {
            hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess) return status_of(err);
        }
    }
    c->timing = enable != 0;
    return FK20_OK;
}

int fk20_ctx_launch_times(fk20_ctx_t *c, float *ms, int *kinds, int max, int *count) {
    if (!c || !ms || !kinds || !count) return FK20_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    const int n = std::min<int>((int)c->nmarks, max);
    for (int k = 0; k < n; k++) {
        hipError_t e = hipEventSynchronize(c->events[k + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms[k], c->events[k], c->events[k + 1]);
        if (e != hipSuccess) return status_of(e);
        kinds[k] = c->kinds[k];
    }
    *count = n;
    return FK20_OK;
}

int fk20_twiddle_selfcheck(unsigned *total_ops, unsigned *wnaf_width) {
    std::vector<uint32_t> ops, index;
    const bool ok = build_programs_host(ops, index);
    if (total_ops) *total_ops = (unsigned)ops.size();
    if (wnaf_width) *wnaf_width = (unsigned)TW_WNAF;
    return ok ? FK20_OK : FK20_EINVAL;
}

int fk20_twiddle_program(int which, unsigned idx, unsigned wnaf, uint32_t *ops, unsigned max_ops, unsigned *n_ops) {
    if (!n_ops || which < 0 || which > 2 || (which < 2 && idx >= 512) || (which == 2 && idx >= 10) || wnaf < 3 ||
        wnaf > 5)
        return FK20_EINVAL;
    DevTables T;
    build_host_tables(T);
    const uint32_t *k;
    const uint32_t *km;
    if (which == 2) {
        k = T.inv2k_std[idx];
        km = T.inv2k_mont[idx];
    } else {
        const unsigned e = which == TW_FWD ? idx : (512 - idx) % 512;
        k = T.tw_std[e];
        km = T.tw_mont[e];
    }
    fs expect;
    fs_set(expect, km);
    std::vector<uint32_t> prog;
    if (!append_program(prog, k, (int)wnaf, expect)) return FK20_EINVAL;
    *n_ops = (unsigned)prog.size();
    if (ops && max_ops) memcpy(ops, prog.data(), 4 * std::min<size_t>(max_ops, prog.size()));
    return FK20_OK;
}

}  // extern "C"
