// gpsat_internal.hpp -- host-side structures shared by the C-ABI and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "common.hpp"

namespace gpsat {

// A batch of instances resident in HBM.
struct DeviceBatch {
    int32_t n_inst = 0;
    int32_t max_n = 0;
    int32_t max_m = 0;
    int64_t n_clauses = 0;
    int64_t n_lits = 0;
    int64_t total_words = 0;      // sum of W_i (rows buffer holds A * total_words words)
    bool fast3 = false;           // every instance fits the register-resident 3-literal kernel
    bool all3 = false;            // every clause of every instance has exactly 3 literals
    InstMeta* meta = nullptr;     // device [n_inst]
    int32_t* lits = nullptr;      // device [n_lits]
    int32_t* coff = nullptr;      // device [n_clauses + n_inst], per-instance relative offsets
    // device [n_clauses] when fast3: clause c's three literal codes in bits 0-9, 10-19 and
    // 20-29 (codes are below 2n <= 512); instance i's clauses start at meta.coff_base - i
    uint32_t* pk3 = nullptr;
    // device [2 * n_clauses] when every instance fits the packed general path (and the batch
    // is not fast3): clause c's word pair, six 10-bit slots (see pkk_fits); klen [n_inst]
    // holds each instance's longest clause
    uint32_t* pkk = nullptr;
    uint8_t* klen = nullptr;
    int32_t* korder = nullptr;    // device [n_inst]: instances by decreasing m * max(3, klen)
    // HBM slices for the eval/BCP mask tables of instances too large for LDS (allocated by
    // the first launch that needs them, freed with the batch)
    mutable uint32_t* gtab = nullptr;
    mutable size_t gtab_words = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    // host mirrors used by the solver set-up and the job chooser
    std::vector<InstMeta> h_meta;
    std::vector<int32_t> h_lits;
    std::vector<int32_t> h_coff;
};

void set_error(const std::string& msg);

// Launch-shape choices, set by the autotuner (gpsat_tune_*) or the GPSAT_TUNE_* environment.
struct Tuning {
    int eval_stage_arena = 0;     // copy each instance's clause arena to LDS before evaluating
    int eval_fast3 = 1;           // use the register-resident kernel when the batch allows it
    int eval_waves_per_block = 4; // independent wavefronts (instances) per workgroup: 1, 2 or 4
    int eval_pkk = 1;             // use the packed general kernel when the batch allows it
    int eval_pkk_waves = 4;       // wavefronts sharing one instance in the packed kernel: 1-4
    int eval_pkk_compact = 32;    // packed kernel: once at most this many rows (0-32) are still
                                  // moving after a BCP pass, the passes go on over a table of
                                  // 32-bit masks of those rows (0: never)
    int solve_ref_per_wave = 1;   // trajectory mode: one instance per wavefront (0: per lane)
    int eval_fast3_small64 = 8;   // 3-literal kernel: at most this many rows open after the
                                  // evaluation run on the 64-bit table instead of a compacted one
                                  // (at ~1.4 passes the bit-by-bit compaction costs more than the
                                  // 32-bit passes save)
};
Tuning& tuning();

// Limits of the register-resident evaluation kernel: every clause has exactly three
// literals, W <= 16 words per row (n <= 256) and m <= 1024 (16 groups of 64 clauses).
constexpr int kFast3MaxWords = 16;
constexpr int kFast3MaxGroups = 16;
inline bool fast3_fits(int n, int m) { return words_per_row(n) <= kFast3MaxWords && m <= 64 * kFast3MaxGroups; }

// Limits of the packed general kernel: clauses of at most 6 literals, W <= 31 words per row
// (n <= 496, so every literal code and the two special slot codes 32W and 32W + 1 fit 10
// bits), any m.
constexpr int kPkkMaxWords = 31;
constexpr int kPkkMaxLen = 6;
inline bool pkk_fits(int n, int max_len) { return words_per_row(n) <= kPkkMaxWords && max_len <= kPkkMaxLen; }

size_t eval_bcp_lds_bytes(int max_n);
// gtab_ext: caller-provided HBM slices for the mask tables of instances too large for LDS
// (eval_bcp_gtab_words of them); null: the batch's own, allocated on first use.
hipError_t launch_eval_bcp(const DeviceBatch& b, const uint32_t* rows, int A, int32_t* counts, uint32_t* out_rows,
                           int32_t* conflict, int32_t* props, int32_t* passes, bool do_eval, bool do_bcp,
                           hipStream_t st, uint32_t* gtab_ext = nullptr);
// The kernel an eval/BCP launch on this batch runs under the current tuning (the kinds of
// gpsat_batch_eval_kernel).
enum EvalKernel { EK_FAST3 = 0, EK_PACKED = 1, EK_CSR = 2, EK_CSR_HBM = 3 };
EvalKernel eval_kernel(const DeviceBatch& b);
// Words of HBM mask-table slices a general eval/BCP launch on this batch needs (0: none).
size_t eval_bcp_gtab_words(const DeviceBatch& b);

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st);

hipError_t launch_preprocess(const DeviceBatch& b, int32_t* dlits, int32_t* dlen, int32_t* out_lits, int32_t* out_coff,
                             int32_t* out_m, int32_t* status, uint32_t* solved, hipStream_t st);

}  // namespace gpsat

// The opaque handle behind the C ABI's gpsat_batch.
struct gpsat_batch {
    gpsat::DeviceBatch d;
};
