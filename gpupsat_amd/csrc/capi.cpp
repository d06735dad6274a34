// capi.cpp -- the extern "C" boundary of libgpsat.so: batches, clause evaluation,
// BCP fixpoints and the device-memory / timing helpers.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpsat.h"
#include "gpsat_internal.hpp"

namespace gpsat {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

Tuning& tuning()
{
    static Tuning t = [] {
        Tuning x;
        if (const char* e = getenv("GPSAT_TUNE_EVAL_STAGE")) x.eval_stage_arena = atoi(e);
        if (const char* e = getenv("GPSAT_TUNE_EVAL_FAST3")) x.eval_fast3 = atoi(e);
        if (const char* e = getenv("GPSAT_TUNE_EVAL_WPB")) x.eval_waves_per_block = atoi(e);
        if (const char* e = getenv("GPSAT_TUNE_EVAL_PKK")) x.eval_pkk = atoi(e);
        if (const char* e = getenv("GPSAT_TUNE_EVAL_PKK_WAVES")) x.eval_pkk_waves = atoi(e);
        if (const char* e = getenv("GPSAT_TUNE_EVAL_PKK_COMPACT")) x.eval_pkk_compact = atoi(e);
        if (const char* e = getenv("GPSAT_TUNE_SOLVE_REF_PER_WAVE")) x.solve_ref_per_wave = atoi(e);
        if (const char* e = getenv("GPSAT_TUNE_EVAL_FAST3_SMALL64")) x.eval_fast3_small64 = atoi(e);
        return x;
    }();
    return t;
}
}  // namespace gpsat

using namespace gpsat;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) {                                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                      \
            return GPSAT_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

#define ARG_CHECK(cond, msg)                 \
    do {                                     \
        if (!(cond)) {                       \
            set_error(msg);                  \
            return GPSAT_ERR_ARG;            \
        }                                    \
    } while (0)

extern "C" {

const char* gpsat_last_error(void) { return g_last_error.c_str(); }
const char* gpsat_version(void) { return "v0.0.1"; }

static int* tune_slot(const char* key)
{
    if (!key) return nullptr;
    if (!strcmp(key, "eval_fast3")) return &tuning().eval_fast3;
    if (!strcmp(key, "eval_stage_arena")) return &tuning().eval_stage_arena;
    if (!strcmp(key, "eval_waves_per_block")) return &tuning().eval_waves_per_block;
    if (!strcmp(key, "eval_pkk")) return &tuning().eval_pkk;
    if (!strcmp(key, "eval_pkk_waves")) return &tuning().eval_pkk_waves;
    if (!strcmp(key, "eval_pkk_compact")) return &tuning().eval_pkk_compact;
    if (!strcmp(key, "solve_ref_per_wave")) return &tuning().solve_ref_per_wave;
    if (!strcmp(key, "eval_fast3_small64")) return &tuning().eval_fast3_small64;
    return nullptr;
}

int gpsat_tune_set(const char* key, int32_t value)
{
    int* p = tune_slot(key);
    ARG_CHECK(p, "unknown tuning key");
    *p = value;
    return GPSAT_OK;
}

int gpsat_tune_get(const char* key, int32_t* value)
{
    int* p = tune_slot(key);
    ARG_CHECK(p && value, "unknown tuning key or null output");
    *value = *p;
    return GPSAT_OK;
}

int gpsat_device_count(int32_t* count)
{
    ARG_CHECK(count, "count is null");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return GPSAT_OK;
}

int gpsat_device_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return GPSAT_OK;
}

int gpsat_dev_alloc(int64_t bytes, void** ptr)
{
    ARG_CHECK(ptr && bytes >= 0, "bad allocation request");
    *ptr = nullptr;
    if (bytes == 0) return GPSAT_OK;
    hipError_t e = hipMalloc(ptr, (size_t)bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? GPSAT_ERR_NOMEM : GPSAT_ERR_HIP;
    }
    return GPSAT_OK;
}

int gpsat_dev_free(void* ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr));
    return GPSAT_OK;
}

int gpsat_memcpy_h2d(void* dst, const void* src, int64_t bytes)
{
    ARG_CHECK(bytes >= 0, "negative size");
    if (bytes) HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return GPSAT_OK;
}

int gpsat_memcpy_d2h(void* dst, const void* src, int64_t bytes)
{
    ARG_CHECK(bytes >= 0, "negative size");
    if (bytes) HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return GPSAT_OK;
}

int gpsat_memset_dev(void* dst, int32_t value, int64_t bytes)
{
    ARG_CHECK(bytes >= 0, "negative size");
    if (bytes) HIP_TRY(hipMemset(dst, value, (size_t)bytes));
    return GPSAT_OK;
}

int gpsat_event_create(void** ev)
{
    ARG_CHECK(ev, "ev is null");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *ev = (void*)e;
    return GPSAT_OK;
}

int gpsat_event_destroy(void* ev)
{
    if (ev) HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return GPSAT_OK;
}

int gpsat_event_record(void* ev, gpsat_batch* b)
{
    ARG_CHECK(ev, "ev is null");
    HIP_TRY(hipEventRecord((hipEvent_t)ev, b ? b->d.stream : nullptr));
    return GPSAT_OK;
}

int gpsat_event_elapsed_ms(void* start, void* stop, float* ms)
{
    ARG_CHECK(start && stop && ms, "null argument");
    HIP_TRY(hipEventSynchronize((hipEvent_t)stop));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return GPSAT_OK;
}

int gpsat_copy_dev(gpsat_batch* b, void* dst, const void* src, int64_t bytes)
{
    ARG_CHECK(bytes >= 0 && bytes % 16 == 0, "copy size must be a multiple of 16");
    ARG_CHECK(bytes == 0 || (dst && src), "null device buffer");
    HIP_TRY(launch_copy(src, dst, bytes, b ? b->d.stream : nullptr));
    return GPSAT_OK;
}

// Packed clause words of one 3-literal instance (three 10-bit literal codes per word).  The
// order of an instance's clauses, and of the literals inside a clause, changes no result of
// the evaluation or BCP, so the words are laid out for the kernel's 16-byte LDS gathers: lane
// c % 64 of a wavefront reads clause c, and for each 16-lane quarter and each of the three
// positions a greedy pass prefers clauses (and literal orders) whose literal slots fall in
// distinct residues mod 16, i.e. in distinct LDS bank groups (a 16-byte slot spans 4 of the
// 64 banks).  Candidates come from the next 48 unplaced clauses.
static void pack3_instance(const int32_t* L, int32_t m, uint32_t* out)
{
    static const int P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    std::vector<char> used((size_t)std::max(m, 1), 0);
    int32_t head = 0;
    for (int32_t b = 0; b < m; b += 16) {
        uint32_t ur[3] = {0, 0, 0};
        for (int32_t ln = 0; ln < 16 && b + ln < m; ln++) {
            while (head < m && used[head]) head++;
            int32_t best = head;
            int bp = 0, bc = 4, seen = 0;
            for (int32_t c = head; c < m && seen < 48 && bc > 0; c++) {
                if (used[c]) continue;
                seen++;
                for (int q = 0; q < 6 && bc > 0; q++) {
                    int cost = 0;
                    for (int k = 0; k < 3; k++) cost += (int)((ur[k] >> (L[3 * c + P[q][k]] & 15)) & 1u);
                    if (cost < bc) {
                        bc = cost;
                        best = c;
                        bp = q;
                    }
                }
            }
            used[best] = 1;
            uint32_t w = 0;
            for (int k = 0; k < 3; k++) {
                const uint32_t l = (uint32_t)L[3 * best + P[bp][k]];
                ur[k] |= 1u << (l & 15u);
                w |= l << (10 * k);
            }
            out[b + ln] = w;
        }
    }
}

// Packed words of one instance for the general path: clause c -> out[2c], out[2c + 1] with
// slot k's literal code at bit 10 * (k % 3) of word k / 3; slots not holding a literal hold
// the always-false slot code 32W + 1.  The kernel body reads K = max(3, longest clause) slots
// per clause, one 16-byte LDS read per slot position across the wavefront.  As for the
// 3-literal words, the order of the clauses and the slot of each literal change no result,
// so a greedy pass chooses them: for each 16-lane quarter of a 64-clause group and each slot
// position it prefers literal slots in distinct residues mod 16 (distinct LDS bank groups),
// taking candidates from the next 32 unplaced clauses and placing each literal in the free
// position where its residue is still unused.
static void packk_instance(const int32_t* lits, const int64_t* coff, int32_t m, int32_t W, int K, uint32_t* out)
{
    const uint32_t fl = 32u * (uint32_t)W + 1u;
    std::vector<char> used((size_t)std::max(m, 1), 0);
    int32_t head = 0;
    int pos_best[kPkkMaxLen], pos_try[kPkkMaxLen];
    for (int32_t b = 0; b < m; b += 16) {
        uint32_t ur[kPkkMaxLen] = {0, 0, 0, 0, 0, 0};
        for (int32_t ln = 0; ln < 16 && b + ln < m; ln++) {
            while (head < m && used[head]) head++;
            int32_t best = head;
            int bc = 1 << 30, seen = 0;
            for (int32_t c = head; c < m && seen < 32 && bc > 0; c++) {
                if (used[c]) continue;
                seen++;
                const int64_t cb = coff[c];
                const int len = (int)(coff[c + 1] - cb);
                uint32_t taken = 0;
                int cost = 0;
                for (int k = 0; k < len; k++) {
                    const uint32_t r = (uint32_t)lits[cb + k] & 15u;
                    int p = -1;
                    for (int q = 0; q < K && p < 0; q++)
                        if (!((taken >> q) & 1u) && !((ur[q] >> r) & 1u)) p = q;
                    if (p < 0) {
                        for (int q = 0; q < K && p < 0; q++)
                            if (!((taken >> q) & 1u)) p = q;
                        cost++;
                    }
                    taken |= 1u << p;
                    pos_try[k] = p;
                }
                if (cost < bc) {
                    bc = cost;
                    best = c;
                    for (int k = 0; k < len; k++) pos_best[k] = pos_try[k];
                }
            }
            used[best] = 1;
            uint32_t code[kPkkMaxLen];
            for (int k = 0; k < kPkkMaxLen; k++) code[k] = fl;
            const int64_t cb = coff[best];
            const int len = (int)(coff[best + 1] - cb);
            for (int k = 0; k < len; k++) {
                code[pos_best[k]] = (uint32_t)lits[cb + k];
                ur[pos_best[k]] |= 1u << ((uint32_t)lits[cb + k] & 15u);
            }
            uint32_t w[2] = {0, 0};
            for (int k = 0; k < kPkkMaxLen; k++) w[k / 3] |= code[k] << (10 * (k % 3));
            out[2 * (b + ln)] = w[0];
            out[2 * (b + ln) + 1] = w[1];
        }
    }
}

int gpsat_batch_create(const int32_t* lits, const int64_t* clause_off, const int64_t* inst_clause_off,
                       const int32_t* n_vars, int32_t n_inst, gpsat_batch** out)
{
    ARG_CHECK(out, "out is null");
    *out = nullptr;
    ARG_CHECK(n_inst >= 0, "negative instance count");
    ARG_CHECK(n_inst == 0 || (clause_off && inst_clause_off && n_vars), "null batch arrays");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return GPSAT_ERR_NODEVICE;
    }
    gpsat_batch* b = new gpsat_batch();
    DeviceBatch& d = b->d;
    d.n_inst = n_inst;
    const int64_t M = n_inst ? inst_clause_off[n_inst] - inst_clause_off[0] : 0;
    const int64_t c0 = n_inst ? inst_clause_off[0] : 0;
    const int64_t L = n_inst ? clause_off[c0 + M] - clause_off[c0] : 0;
    const int64_t l0 = n_inst ? clause_off[c0] : 0;
    d.n_clauses = M;
    d.n_lits = L;
    d.h_meta.resize(n_inst);
    d.h_coff.resize(M + n_inst);
    d.h_lits.assign(lits ? lits + l0 : nullptr, lits ? lits + l0 + L : nullptr);
    int64_t wbase = 0;
    bool fast3 = true, every3 = true, packk = true;
    std::vector<uint8_t> h_klen(n_inst, 0);
    for (int32_t i = 0; i < n_inst; i++) {
        const int64_t cb = inst_clause_off[i], ce = inst_clause_off[i + 1];
        if (ce < cb || n_vars[i] < 0) {
            delete b;
            set_error("malformed batch offsets");
            return GPSAT_ERR_ARG;
        }
        InstMeta& mi = d.h_meta[i];
        mi.lit_base = clause_off[cb] - l0;
        mi.coff_base = (cb - c0) + i;
        mi.word_base = wbase;
        mi.m = (int32_t)(ce - cb);
        mi.n = n_vars[i];
        for (int64_t c = cb; c <= ce; c++) {
            const int64_t rel = clause_off[c] - clause_off[cb];
            if (rel < 0 || rel > INT32_MAX) {
                delete b;
                set_error("clause offsets out of range");
                return GPSAT_ERR_ARG;
            }
            d.h_coff[mi.coff_base + (c - cb)] = (int32_t)rel;
        }
        for (int64_t k = clause_off[cb]; k < clause_off[ce]; k++) {
            const int32_t l = lits[k];
            if (l < 0 || (l >> 1) >= n_vars[i]) {
                delete b;
                set_error("literal code out of range for its instance");
                return GPSAT_ERR_ARG;
            }
        }
        bool all3 = true;
        int64_t max_len = 0;
        for (int64_t c = cb; c < ce; c++) {
            const int64_t len = clause_off[c + 1] - clause_off[c];
            all3 = all3 && len == 3;
            max_len = std::max(max_len, len);
        }
        packk = packk && pkk_fits(mi.n, (int)std::min<int64_t>(max_len, INT32_MAX));
        h_klen[i] = (uint8_t)std::min<int64_t>(max_len, 255);
        fast3 = fast3 && all3 && fast3_fits(mi.n, mi.m);
        every3 = every3 && all3;
        wbase += words_per_row(mi.n);
        d.max_n = std::max(d.max_n, mi.n);
        d.max_m = std::max(d.max_m, mi.m);
    }
    d.total_words = wbase;
    d.fast3 = fast3 && n_inst > 0;
    d.all3 = every3 && n_inst > 0;
    // the register-resident kernel reads one 4-byte word per clause instead of three
    // (instances split over host threads)
    std::vector<uint32_t> h_pk3;
    if (d.fast3) {
        h_pk3.resize(M);
        const int32_t T = (int32_t)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const int32_t nt = std::min(T, 1 + n_inst / 64);
        const auto work = [&](int32_t t) {
            for (int32_t i = t; i < n_inst; i += nt) {
                const int64_t cb = inst_clause_off[i];
                pack3_instance(lits + clause_off[cb], d.h_meta[i].m, h_pk3.data() + (cb - c0));
            }
        };
        std::vector<std::thread> pool;
        for (int32_t t = 1; t < nt; t++) pool.emplace_back(work, t);
        work(0);
        for (std::thread& th : pool) th.join();
    }
    // the packed general kernel reads one 8-byte word pair per clause instead of the offsets
    // and the literals (batches that the 3-literal kernel does not take)
    std::vector<uint32_t> h_pkk;
    // (GPSAT_PKK_FOR_FAST3=1 builds them for 3-literal batches too, so both kernels can be
    // compared on one batch with eval_fast3 = 0)
    const char* pf3 = getenv("GPSAT_PKK_FOR_FAST3");
    const bool use_pkk = packk && (!d.fast3 || (pf3 && atoi(pf3) != 0)) && n_inst > 0 && M > 0;
    if (use_pkk) {
        h_pkk.resize(2 * M);
        const int32_t T = (int32_t)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const int32_t nt = std::min(T, 1 + n_inst / 16);
        const auto work = [&](int32_t t) {
            for (int32_t i = t; i < n_inst; i += nt) {
                const int64_t cb = inst_clause_off[i];
                packk_instance(lits, clause_off + cb, d.h_meta[i].m, words_per_row(d.h_meta[i].n),
                               std::max<int>(3, h_klen[i]), h_pkk.data() + 2 * (cb - c0));
            }
        };
        std::vector<std::thread> pool;
        for (int32_t t = 1; t < nt; t++) pool.emplace_back(work, t);
        work(0);
        for (std::thread& th : pool) th.join();
    }
    hipError_t e = hipGetDevice(&d.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking);
    if (e == hipSuccess && n_inst) e = hipMalloc(&d.meta, sizeof(InstMeta) * n_inst);
    if (e == hipSuccess && L) e = hipMalloc(&d.lits, sizeof(int32_t) * L);
    if (e == hipSuccess && n_inst) e = hipMalloc(&d.coff, sizeof(int32_t) * (M + n_inst));
    if (e == hipSuccess && n_inst) e = hipMemcpy(d.meta, d.h_meta.data(), sizeof(InstMeta) * n_inst, hipMemcpyHostToDevice);
    if (e == hipSuccess && L) e = hipMemcpy(d.lits, d.h_lits.data(), sizeof(int32_t) * L, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_inst)
        e = hipMemcpy(d.coff, d.h_coff.data(), sizeof(int32_t) * (M + n_inst), hipMemcpyHostToDevice);
    if (e == hipSuccess && d.fast3) e = hipMalloc(&d.pk3, sizeof(uint32_t) * M);
    if (e == hipSuccess && d.fast3) e = hipMemcpy(d.pk3, h_pk3.data(), sizeof(uint32_t) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess && use_pkk) e = hipMalloc(&d.pkk, sizeof(uint32_t) * 2 * M);
    if (e == hipSuccess && use_pkk) e = hipMemcpy(d.pkk, h_pkk.data(), sizeof(uint32_t) * 2 * M, hipMemcpyHostToDevice);
    if (e == hipSuccess && use_pkk) e = hipMalloc(&d.klen, n_inst);
    if (e == hipSuccess && use_pkk) e = hipMemcpy(d.klen, h_klen.data(), n_inst, hipMemcpyHostToDevice);
    if (e == hipSuccess && use_pkk) {
        std::vector<int32_t> order(n_inst);
        for (int32_t i = 0; i < n_inst; i++) order[i] = i;
        const auto cost = [&](int32_t i) { return (int64_t)d.h_meta[i].m * std::max<int64_t>(3, h_klen[i]); };
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) { return cost(a) > cost(c); });
        e = hipMalloc(&d.korder, sizeof(int32_t) * n_inst);
        if (e == hipSuccess) e = hipMemcpy(d.korder, order.data(), sizeof(int32_t) * n_inst, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        set_error(std::string("batch upload: ") + hipGetErrorString(e));
        gpsat_batch_destroy(b);
        return e == hipErrorOutOfMemory ? GPSAT_ERR_NOMEM : GPSAT_ERR_HIP;
    }
    *out = b;
    return GPSAT_OK;
}

int gpsat_batch_destroy(gpsat_batch* b)
{
    if (!b) return GPSAT_OK;
    DeviceBatch& d = b->d;
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    (void)hipFree(d.meta);
    (void)hipFree(d.lits);
    (void)hipFree(d.coff);
    (void)hipFree(d.pk3);
    (void)hipFree(d.pkk);
    (void)hipFree(d.klen);
    (void)hipFree(d.korder);
    (void)hipFree(d.gtab);
    if (d.stream) (void)hipStreamDestroy(d.stream);
    delete b;
    return GPSAT_OK;
}

int gpsat_batch_eval_kernel(const gpsat_batch* b, int32_t* kind)
{
    ARG_CHECK(b && kind, "null argument");
    *kind = (int32_t)eval_kernel(b->d);
    return GPSAT_OK;
}

int gpsat_batch_info(const gpsat_batch* b, int32_t* n_inst, int64_t* n_clauses, int64_t* n_lits, int32_t* max_vars)
{
    ARG_CHECK(b, "null batch");
    if (n_inst) *n_inst = b->d.n_inst;
    if (n_clauses) *n_clauses = b->d.n_clauses;
    if (n_lits) *n_lits = b->d.n_lits;
    if (max_vars) *max_vars = b->d.max_n;
    return GPSAT_OK;
}

int gpsat_batch_row_words(const gpsat_batch* b, int32_t A, int64_t* total_words)
{
    ARG_CHECK(b && total_words, "null argument");
    ARG_CHECK(A >= 0, "negative A");
    *total_words = (int64_t)A * b->d.total_words;
    return GPSAT_OK;
}

int gpsat_batch_sync(gpsat_batch* b)
{
    ARG_CHECK(b, "null batch");
    HIP_TRY(hipStreamSynchronize(b->d.stream));
    return GPSAT_OK;
}

static int check_A(int32_t A)
{
    if (A < 1 || A > 64) {
        set_error("A must be in [1, 64]");
        return GPSAT_ERR_ARG;
    }
    return GPSAT_OK;
}

int gpsat_clause_eval_dev(gpsat_batch* b, const uint32_t* rows_dev, int32_t A, int32_t* counts_dev)
{
    ARG_CHECK(b, "null batch");
    if (int r = check_A(A)) return r;
    ARG_CHECK(b->d.n_inst == 0 || (rows_dev && counts_dev), "null device buffer");
    HIP_TRY(launch_eval_bcp(b->d, rows_dev, A, counts_dev, nullptr, nullptr, nullptr, nullptr, true, false,
                            b->d.stream));
    return GPSAT_OK;
}

int gpsat_bcp_fixpoint_dev(gpsat_batch* b, const uint32_t* rows_dev, int32_t A, uint32_t* out_rows_dev,
                           int32_t* conflict_dev, int32_t* props_dev)
{
    ARG_CHECK(b, "null batch");
    if (int r = check_A(A)) return r;
    ARG_CHECK(b->d.n_inst == 0 || (rows_dev && out_rows_dev && conflict_dev && props_dev), "null device buffer");
    HIP_TRY(launch_eval_bcp(b->d, rows_dev, A, nullptr, out_rows_dev, conflict_dev, props_dev, nullptr, false, true,
                            b->d.stream));
    return GPSAT_OK;
}

int gpsat_eval_bcp_dev(gpsat_batch* b, const uint32_t* rows_dev, int32_t A, int32_t* counts_dev,
                       uint32_t* out_rows_dev, int32_t* conflict_dev, int32_t* props_dev)
{
    ARG_CHECK(b, "null batch");
    if (int r = check_A(A)) return r;
    ARG_CHECK(b->d.n_inst == 0 || (rows_dev && counts_dev && out_rows_dev && conflict_dev && props_dev),
              "null device buffer");
    HIP_TRY(launch_eval_bcp(b->d, rows_dev, A, counts_dev, out_rows_dev, conflict_dev, props_dev, nullptr, true, true,
                            b->d.stream));
    return GPSAT_OK;
}

// Host-buffer forms: stage in HBM, run, copy back, free.
namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
};
}  // namespace

int gpsat_preprocess_batch(gpsat_batch* b, int32_t* status, uint32_t* solved_rows, int32_t* out_lits,
                           int64_t* out_clause_off, int64_t* out_inst_clause_off)
{
    ARG_CHECK(b, "null batch");
    ARG_CHECK(status && solved_rows && out_lits && out_clause_off && out_inst_clause_off, "null output buffer");
    DeviceBatch& d = b->d;
    const int32_t N = d.n_inst;
    out_inst_clause_off[0] = 0;
    out_clause_off[0] = 0;
    if (N == 0) return GPSAT_OK;
    const size_t L = (size_t)std::max<int64_t>(d.n_lits, 1), C = (size_t)d.n_clauses + N;
    DevBuf dl, dn, ol, oc, om, ds, dsol;
    HIP_TRY(hipMalloc(&dl.p, L * 4));
    HIP_TRY(hipMalloc(&dn.p, C * 4));
    HIP_TRY(hipMalloc(&ol.p, L * 4));
    HIP_TRY(hipMalloc(&oc.p, C * 4));
    HIP_TRY(hipMalloc(&om.p, (size_t)N * 4));
    HIP_TRY(hipMalloc(&ds.p, (size_t)N * 4));
    HIP_TRY(hipMalloc(&dsol.p, (size_t)std::max<int64_t>(d.total_words, 1) * 4));
    HIP_TRY(launch_preprocess(d, (int32_t*)dl.p, (int32_t*)dn.p, (int32_t*)ol.p, (int32_t*)oc.p, (int32_t*)om.p,
                              (int32_t*)ds.p, (uint32_t*)dsol.p, d.stream));
    std::vector<int32_t> h_ol(L), h_oc(C), h_om(N);
    HIP_TRY(hipMemcpyAsync(h_ol.data(), ol.p, L * 4, hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipMemcpyAsync(h_oc.data(), oc.p, C * 4, hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipMemcpyAsync(h_om.data(), om.p, (size_t)N * 4, hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipMemcpyAsync(status, ds.p, (size_t)N * 4, hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipMemcpyAsync(solved_rows, dsol.p, (size_t)d.total_words * 4, hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    // per-instance compacted regions -> one global CSR (gpsat_batch_create's layout)
    int64_t nc = 0, nl = 0;
    for (int32_t i = 0; i < N; i++) {
        const InstMeta& mi = d.h_meta[i];
        const int32_t* oci = h_oc.data() + mi.coff_base;
        for (int32_t c = 0; c < h_om[i]; c++) {
            for (int32_t k = oci[c]; k < oci[c + 1]; k++) out_lits[nl++] = h_ol[mi.lit_base + k];
            out_clause_off[++nc] = nl;
        }
        out_inst_clause_off[i + 1] = nc;
    }
    return GPSAT_OK;
}

int gpsat_clause_eval(gpsat_batch* b, const uint32_t* rows, int32_t A, int32_t* counts)
{
    ARG_CHECK(b, "null batch");
    if (int r = check_A(A)) return r;
    if (b->d.n_inst == 0) return GPSAT_OK;
    ARG_CHECK(rows && counts, "null host buffer");
    const size_t rw = (size_t)A * b->d.total_words * 4, cb = (size_t)b->d.n_inst * A * 3 * 4;
    DevBuf dr, dc;
    HIP_TRY(hipMalloc(&dr.p, std::max<size_t>(rw, 4)));
    HIP_TRY(hipMalloc(&dc.p, cb));
    HIP_TRY(hipMemcpyAsync(dr.p, rows, rw, hipMemcpyHostToDevice, b->d.stream));
    HIP_TRY(launch_eval_bcp(b->d, (const uint32_t*)dr.p, A, (int32_t*)dc.p, nullptr, nullptr, nullptr, nullptr, true,
                            false, b->d.stream));
    HIP_TRY(hipMemcpyAsync(counts, dc.p, cb, hipMemcpyDeviceToHost, b->d.stream));
    HIP_TRY(hipStreamSynchronize(b->d.stream));
    return GPSAT_OK;
}

int gpsat_bcp_fixpoint(gpsat_batch* b, const uint32_t* rows, int32_t A, uint32_t* out_rows, int32_t* conflict,
                       int32_t* props)
{
    ARG_CHECK(b, "null batch");
    if (int r = check_A(A)) return r;
    if (b->d.n_inst == 0) return GPSAT_OK;
    ARG_CHECK(rows && out_rows && conflict && props, "null host buffer");
    const size_t rw = (size_t)A * b->d.total_words * 4, fb = (size_t)b->d.n_inst * A * 4;
    DevBuf dr, dout, dcf, dpr;
    HIP_TRY(hipMalloc(&dr.p, std::max<size_t>(rw, 4)));
    HIP_TRY(hipMalloc(&dout.p, std::max<size_t>(rw, 4)));
    HIP_TRY(hipMalloc(&dcf.p, fb));
    HIP_TRY(hipMalloc(&dpr.p, fb));
    HIP_TRY(hipMemcpyAsync(dr.p, rows, rw, hipMemcpyHostToDevice, b->d.stream));
    HIP_TRY(launch_eval_bcp(b->d, (const uint32_t*)dr.p, A, nullptr, (uint32_t*)dout.p, (int32_t*)dcf.p,
                            (int32_t*)dpr.p, nullptr, false, true, b->d.stream));
    HIP_TRY(hipMemcpyAsync(out_rows, dout.p, rw, hipMemcpyDeviceToHost, b->d.stream));
    HIP_TRY(hipMemcpyAsync(conflict, dcf.p, fb, hipMemcpyDeviceToHost, b->d.stream));
    HIP_TRY(hipMemcpyAsync(props, dpr.p, fb, hipMemcpyDeviceToHost, b->d.stream));
    HIP_TRY(hipStreamSynchronize(b->d.stream));
    return GPSAT_OK;
}

}  // extern "C"
