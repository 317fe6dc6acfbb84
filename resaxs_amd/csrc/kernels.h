// Device-side data structures and kernel launchers of the Debye-SAXS engine.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace rsx {

// Class id stored in the w bits of a column/row record that contributes nothing
// (padding, atoms temporarily removed from a column set).
constexpr uint32_t CLS_SKIP = 0x7fffffffu;

// One unit of pair-sum work: up to 64 row atoms (one per lane) against a
// contiguous run of column atoms.  Records are float4 {x, y, z, class bits}.
// slot selects the row of the partial-sum table the job writes; scale
// multiplies its sum.  diag != 0 means cols[0..63] are the same atoms as
// rows[0..63] (the diagonal tile of a triangle), so only j > i counts fully,
// j == i counts half and j < i not at all (the job scale is then 2).
struct PairJob {
    const float4* rows;
    const float4* cols;
    int nrow;
    int ncol;
    int slot;
    float scale;
    int diag;
    int pad;
};

// Per-q-chunk constants of the uniform-grid path: a = first q of the chunk,
// *_rev = value / (2 pi) (v_sin/v_cos take revolutions).  The seed phases
// q r / (2 pi) are formed in FP64 and reduced there: an FP32 constant would
// carry one rounding error shared by every pair, i.e. a coherent phase shift.
struct ChunkConst {
    double a_rev;    // a / (2 pi)
    double mid_rev;  // (a + dq/2) / (2 pi)
    float a;
    float pad;
};

// highest power of the FP64 sin(dq r / 2) series is x^(2 SIN_MAX_TERMS + 1)
constexpr int SIN_MAX_TERMS = 8;

// series length for a layout whose pairs satisfy dq r / 2 <= x_max: the first
// omitted term (x^(2m+2) / (2m+3)! relative) stays below 1e-16 at 1.25 x_max
// (margin for device-resident positions that drift from the laid-out ones)
inline int sin_series_terms(double x_max)
{
    const double x = 1.25 * x_max;
    double term = x * x / 6.0;  // m = 0: x^2 / 3!
    int m = 0;
    while (m < SIN_MAX_TERMS && term > 1e-16) {
        ++m;
        term *= x * x / ((2.0 * m + 2.0) * (2.0 * m + 3.0));
    }
    return m < 2 ? 2 : m;
}

// pair-kernel paths (K lists differ) and recurrence variants
enum { PATH_DIRECT = 0, PATH_REC64 = 1, PATH_REC32 = 2, PATH_REC32R = 3 };
enum {
    VARIANT_WAVE = 0,    // FP64 recurrence, one wave per (job, q chunk)
    VARIANT_SHARED = 1,  // FP64 recurrence, the q chunks of a job share the pair geometry through LDS
    VARIANT_FP32 = 2,    // FP32 difference-form recurrence, one wave per (job, q chunk)
    VARIANT_FP32R = 3,   // the same recurrence, each class run reduced over the rows when it ends
    VARIANT_COUNT = 4
};

struct DebyeArgs {
    const PairJob* jobs;
    int n_chunks;         // q chunks per job (one wave each)
    const float* ftab;    // [n_class + 1][ldf] FP32 form factors per class, zero padded;
                          // row zero_class (= n_class) is all zeros
    int ldf;              // n_chunks * K
    int zero_class;
    const ChunkConst* chunks;
    double half_dq;       // dq / 2 (FP64)
    float dq;
    const double* invq;   // [ldf] 1/q, 0 for q == 0 and padding
    const double* qrev;   // [ldf] q / (2 pi)  (direct path)
    const float* qsplit;  // [2 ldf] q / (2 pi) as FP32 pairs (hi, lo), hi + lo = qrev to ~1e-14
                          // relative (direct path, FP32 phase)
    int phase32;          // direct path: phases in FP32 (every q r / (2 pi) <= DIRECT_PHASE32_MAX_REV)
    const float* qv;      // [ldf] q           (direct path)
    double* partials;     // [n_slots][ldf]
    int variant;          // recurrence kernel: VARIANT_WAVE, VARIANT_SHARED, VARIANT_FP32 or VARIANT_FP32R
    int* queue;           // work counter of the persistent FP32 launch (device int), or null
    int persistent_waves; // > 0: persistent launch of this many waves pulling items from queue
    int sin_terms;        // terms the sin(dq r / 2) series needs (sin_series_terms); the launcher
                          // picks the shortest instantiated length >= it
};

// Largest phase, in revolutions, of the direct kernel's FP32 phase form: v_sin_f32 takes
// revolutions and is exact in its reduction over its documented domain |x| <= 256 (measured
// on gfx950: 1.2e-7 absolute error, unbiased, to 32 revolutions); the bound leaves the
// rounding of the phase itself (half an ulp, 4e-6 revolutions at 64) random per pair.
constexpr double DIRECT_PHASE32_MAX_REV = 64.0;

// Debye pair sums.  uniform: q_k = a + k dq within each chunk (FP64
// three-term recurrence); otherwise direct FP32 sin per term.
hipError_t launch_debye(const DebyeArgs& a, int n_jobs, int K, bool uniform, hipStream_t s);
// Resident one-wave blocks per CU of an FP32 pair kernel (variant VARIANT_FP32 or
// VARIANT_FP32R, the given K and series length; the persistent work-queue form or the
// one-wave-per-item form); 0 if none.
int rec32_blocks_per_cu(int variant, int K, int sin_terms, bool persistent);
// K values the launcher accepts, ascending.
int debye_k_count(int path);
int debye_k_value(int path, int i);

// out[o*out_ld + k] (+)= sum_{s in [slot_begin[o], slot_begin[o+1])} partials[s*ldf + k]
// in a fixed order.  q0_index >= 0: that k is set to q0_val[o] instead (the
// exact (sum F)^2 rule for q == 0, added when accumulate != 0).  max_slots bounds the
// slots of any output; with scratch (reduce_scratch_doubles of them, null if 0) the
// slots are summed in chunks by more blocks, then the chunks in order.
hipError_t launch_reduce(const double* partials, int ldf, int n_q, const int* slot_begin, int n_out,
                         double* out, int out_ld, int accumulate, const double* q0_val, int q0_index,
                         int max_slots, double* scratch, hipStream_t s);
size_t reduce_scratch_doubles(int n_q, int n_out, int max_slots);

// out[i] = mean over frames (uniform weights) of in[f*n_q + i]
hipError_t launch_frame_mean(const double* in, int n_frames, int n_q, double* out, hipStream_t s);

// cols[slot[t]] = v for t < m
hipError_t launch_scatter_value(float4* cols, const int* slot, float4 v, int m, hipStream_t s);
// cols[slot[t]] = vals[t] and s2c[slot[t]] = callers[t] (slots whose atom changed)
hipError_t launch_scatter_slots(float4* cols, int* s2c, const int* slot, const float4* vals, const int* callers, int m,
                                hipStream_t s);

// cols[s] = {bodies[src[s]].xyz, class}, class = type(bodies[src[s]]) * d1 + (counts ? counts[src[s]] : 0).
// Checks of device-resident bodies, reported through *flag (bits or-ed in, never cleared
// here): a type not below n_types gets the skip class and sets GUARD_TYPE; a body farther
// than sqrt(r2_max) from centre sets GUARD_RADIUS.
constexpr int GUARD_TYPE = 1;
constexpr int GUARD_RADIUS = 2;
struct GatherGuard {
    int n_types;
    float cx, cy, cz;
    float r2_max;
    int* flag;
};
hipError_t launch_gather_cols(float4* cols, const float4* bodies, const int* src, const uint32_t* counts,
                              int d1, int n, const GatherGuard& g, hipStream_t s);

// Stable counting sort of device-resident bodies by class, class = type * d1 + counts[i]:
// slot s receives the s-th body in (class, caller index) order, the order the host layout
// gives them.  cols[s] = {x, y, z, class}, s2c[s] = caller index; a type not below n_types
// sorts last with the skip class.  The GatherGuard checks apply.  Scratch: table
// [ceil(n / 1024)][n_class + 1] ints, tot [n_class + 1] ints.  n_class + 1 <=
// CLASS_SORT_MAX_KEYS (the per-block histogram lives in LDS).
constexpr int CLASS_SORT_MAX_KEYS = 16384;
struct ClassSortArgs {
    const float4* bodies;
    const uint32_t* counts;
    int d1, n, n_types, n_class;
    int* table;
    int* tot;
    float4* cols;
    int* s2c;
    GatherGuard g;
};
hipError_t launch_class_sort(const ClassSortArgs& a, hipStream_t s);

// ---- SASA ----
// Inputs of a local recount on the device: the box [lo, hi], the moved range, the counter
// set this recount uses (parity) and, from LOCAL_STAGE_HEADER bytes on, the moved atoms.
struct LocalStage {
    float lo[3], hi[3];
    int start, len;
    int parity;
};
constexpr int LOCAL_STAGE_HEADER = 48;  // bytes before the moved atoms (float4-aligned)
// Counters of a local recount: two sets of LOCAL_HDR_INTS ints ([listed, changed, slab
// overflow, unused]); recount set p is cleared by the recount before it, which works on set
// 1 - p (or by the slab build).  The changed (atom, count) pairs follow the two sets.
constexpr int LOCAL_HDR_INTS = 4;
constexpr int LOCAL_HDR_TOTAL = 2 * LOCAL_HDR_INTS;
// Records per bucket of the slab cell list.
constexpr int SLAB_CAP = 64;

struct SasaArgs {
    const float4* atoms;     // x, y, z, type bits
    int n;
    int n_types;             // an atom whose type is not below n_types has no dots (count 0)
                             // and occludes nothing
    const float* r2;         // [n_types] (probe + r)^2
    const float* rext;       // [n_types] sqrt(r2) (prefilter)
    const float4* dots;      // [n_types][n_dots]
    int n_dots;
    float cell;              // cell edge
    float inv_cell;
    int n_buckets;           // power of two
    int* bucket_of;          // [n]
    int* bucket_count;       // [n_buckets + 1]
    int* bucket_start;       // [n_buckets + 1]
    int* bucket_fill;        // [n_buckets]
    int* tile_sum;           // [n_buckets / 1024] scan scratch
    int* sorted;             // [n] atom ids grouped by bucket
    float4* sorted_atoms;    // [n] the atoms in bucket order
    uint32_t* counts;        // [n] exposed dots (only the listed atoms when subset is set)
    const int* subset;       // null: count every atom; else count atoms subset[0 .. n_subset)
    int n_subset;
    // Local recount (local_hdr set; subset unused): the place kernel (or, with a slab, the
    // slab update) lists the atoms whose position lies in the box of local_stage in local_list
    // (in no fixed order); the dots kernel counts only those, over a fixed grid, and for each
    // count that differs from the one in counts appends (atom, new count) to local_changed
    // before overwriting it.  Counters in set local_stage->parity of local_hdr
    // (LOCAL_HDR_TOTAL ints): [0] atoms listed, [1] counts changed (entries from changed_cap
    // on are counted but not written), [2] slab overflow.
    int* local_hdr = nullptr;
    int* local_list = nullptr;   // [n]
    int* local_changed = nullptr;  // [changed_cap][2]
    int changed_cap = 0;
    // The local recount's inputs on the device (LocalStage).  With patch_cap > 0 the first
    // kernel also writes the stage's moved atoms into atoms[start, start + len) (len <=
    // patch_cap; atoms_w is the writable view of atoms).
    const LocalStage* local_stage = nullptr;
    float4* atoms_w = nullptr;
    int patch_cap = 0;
    // Slab cell list of a local recount (slab set): bucket b (of slab_buckets, a power of two)
    // holds its records at slab[b * SLAB_CAP, b * SLAB_CAP + min(slab_count[b], SLAB_CAP)), in
    // no fixed order; slab_pos[i] is the index of atom i's record (-1: none).  A local recount
    // keeps it current for the moved atoms instead of rebuilding the hashed cell list: a moved
    // atom's record is rewritten in place when its bucket is unchanged, otherwise replaced by a
    // tombstone (type bits all ones: no dots, occludes nothing) and appended to its new bucket.
    // A bucket that would pass SLAB_CAP raises the set's overflow counter; the counts of that
    // recount are then not to be used.
    float4* slab = nullptr;
    int* slab_count = nullptr;
    int* slab_pos = nullptr;
    int slab_buckets = 0;
};
hipError_t launch_sasa(const SasaArgs& a, hipStream_t s);
// The slab of a local recount built from atoms[0, n): slab_count and both counter sets
// cleared, every atom appended to its bucket.
hipError_t launch_slab_build(const SasaArgs& a, hipStream_t s);

// ---- adjoint ----
// Adjoint over the forward's triangle pair jobs (slot-ordered atoms).
struct AdjointArgs {
    const PairJob* jobs;     // triangle jobs, slot = execution index
    int n_chunks;
    const float* ftab;       // [n_class][ldf] FP32 class factors
    int ldf;                 // n_chunks * K
    const ChunkConst* chunks;
    double half_dq;
    const double* wq;        // [ldf] w_k = dchi2/dI_k (0 for q == 0 and padding)
    const double* invq;      // [ldf]
    const float* invq32;     // [ldf] the same in FP32 (FP32 kernel)
    const float* cb;         // [n_class][ldf] 2 w_k F_c,k (FP32 kernel; launch_adjoint_coef)
    const float* cg;         // [n_class][ldf] 2 w_k F_c,k / q_k
    double dq_rev;           // dq / (2 pi)
    int fp32;                // 1: FP32 difference-form kernel, 0: FP64 recurrence
    int sin_terms;           // series length (sin_series_terms)
    int col_range;           // columns per job (L)
    float* row_partial;      // [n_jobs * n_chunks][64][3] FP32 row-side partial gradients
    float* col_partial;      // [n_jobs * n_chunks][col_range][3] FP32 column-side partials
    // Batched VJP (FP32 kernel only): per-frame weights [n_frames][ldf], the job's frame in its
    // pad; cb and cg then hold 2 F_c,k and 2 F_c,k / q_k, and the kernel scales them by the
    // job's frame weights.  Null: one weight vector, folded into cb and cg.
    const float* wframe = nullptr;
    // Classes of the factor table: a row or column record whose class is not below n_class (a
    // type outside the table, flagged by the gather) adds nothing and reads no table entry.
    int n_class = 0;
    // Class-pair coefficient form (FP32 kernel, plain layouts): beta and gamma per (group, row
    // class, column class, q chunk) from launch_adjoint_pair_coef, read through scalar loads.
    // A group is a frame of the batch (coef_grouped: the job's frame in its pad) or the one
    // structure.  Null: the per-lane form above.
    const float* coef = nullptr;
    int coef_grouped = 0;
};
// one wave per (job, q chunk); n_jobs = number of jobs
hipError_t launch_adjoint(const AdjointArgs& a, int n_jobs, int K, hipStream_t s);
// cb, cg of the FP32 adjoint from the class factors and the weights
hipError_t launch_adjoint_coef(const float* ftab, const double* wq, const double* invq, int n_class, int ldf, float* cb,
                               float* cg, hipStream_t s);
int adjoint_k_count(bool fp32);
int adjoint_k_value(bool fp32, int i);
// q per thread of the class-pair coefficient form
int adjoint_pair_k_count();
int adjoint_pair_k_value(int i);
// Floats per (group, row class, column class, chunk) entry of the class-pair table, 16-float
// aligned: beta_k and gamma_k interleaved (2 k and 2 k + 1 for k < K), beta_K at 2 K -- one
// stream, consumed in order, so its scalar loads run ahead of their use.
constexpr int adjoint_pair_stride(int K)
{
    return (2 * K + 1 + 15) / 16 * 16;
}

// The class-pair table: for group g, row class ci, column class cl and chunk c (k = c K + s),
// with f_k = F_ci,k (times wframe[g][k] when wframe is set) and b_k = f_k cb[cl][k]:
// beta_s = b_s + b_{s-1} (b outside the chunk taken as 0) and gamma_s = f_k cg[cl][k], the
// per-lane form's products with the same rounding; layout as adjoint_pair_stride states.
struct AdjointPairCoefArgs {
    const float* ftab;    // [n_class][ldf]
    const float* cb;      // [n_class][ldf]
    const float* cg;      // [n_class][ldf]
    const float* wframe;  // [groups][ldf] or null
    int groups, n_class, n_chunks, K, ldf;
    float* coef;          // [groups][n_class][n_class][n_chunks][adjoint_pair_stride(K)]
};
hipError_t launch_adjoint_pair_coef(const AdjointPairCoefArgs& a, hipStream_t s);

struct AdjointReduceArgs {
    const float* row_partial;
    const float* col_partial;
    const int* tri_slot;     // [n_row_blocks][tri_stride]: slot of row block I's m-th column range
    int tri_stride;
    int n;                   // atoms
    int n_chunks;
    int col_range;
    const int* slot_to_caller;
    float* grad;             // [n][3] in caller order
};
// grad of every atom: its row and column partials over all chunks, fixed order
hipError_t launch_adjoint_reduce(const AdjointReduceArgs& a, hipStream_t s);

// The same over a batch of frames, each with its own triangle of jobs: atom s (batch column
// order) belongs to frame atom_frame[s], whose atoms are [frame_off[f], frame_off[f + 1])
// and whose (row block, column range) -> slot table starts at tri_slot + tri_off[f] with
// stride ceil(atoms / col_range).
struct AdjointBatchReduceArgs {
    const float* row_partial;
    const float* col_partial;
    const int* tri_slot;
    const int* tri_off;      // [n_frames]
    const int* atom_frame;   // [n]
    const int* frame_off;    // [n_frames + 1]
    int n;                   // atoms of all frames
    int n_chunks;
    int col_range;
    const int* slot_to_caller;
    float* grad;             // [n][3] in caller order
};
hipError_t launch_adjoint_batch_reduce(const AdjointBatchReduceArgs& a, hipStream_t s);

// Adjoint weights on the device.  wq[0 .. ldfa): wq[k] = w[k] for k < n_q except the q = 0
// entry (q0_index >= 0), 0 there and in the padding.
hipError_t launch_adjoint_weights(const double* w, int n_q, int q0_index, int ldfa, double* wq, hipStream_t s);
// chi^2 at the least-squares scale, one workgroup, fixed-order FP64 sums over k < n_q:
// c = sum I_exp I / sigma^2 / sum I^2 / sigma^2, chi2 = sum (I_exp - c I)^2 / sigma^2;
// out = {chi2, c}; wq as launch_adjoint_weights with w_k = -2 c (I_exp,k - c I_k) / sigma_k^2.
hipError_t launch_chi2_weights(const double* iq, const float* i_exp, const float* sigma, int n_q, int q0_index,
                               int ldfa, double* out, double* wq, hipStream_t s);

// out[0] = (sum_s f0col[class(cols[s])])^2, fixed-order FP64 sum (the q = 0 value)
hipError_t launch_q0(const float4* cols, int n, const double* f0col, double* out, hipStream_t s);

// dst[i] = src[i] for n float4 (grid-stride), the HBM calibration copy
hipError_t launch_copy(float4* dst, const float4* src, long long n, hipStream_t s);

// ---- distance-histogram Debye (approximate; large N) ----
// Pair distances binned by r_b = (b + 1/2) dr: h0[type pair][bin] counts unordered
// pairs, and with exposed-dot counts (hydration) h1[type][bin] sums the partner's count
// of pairs whose other atom has that type, h2[bin] the product of both counts.  Integer
// sums: the histogram is exact and order-free.
struct HistArgs {
    const float4* atoms;       // [n] x, y, z, type bits
    const uint32_t* counts;    // [n] exposed-dot counts, or null (plain)
    int n;
    int n_types;
    float inv_dr;
    int n_bins;
    const int* pair_index;     // [n_types][n_types] -> type pair (symmetric)
    unsigned int* h0;          // [n_pairs][n_bins]
    unsigned long long* h1;    // [n_types][n_bins]
    unsigned long long* h2;    // [n_bins]
};
hipError_t launch_hist_pairs(const HistArgs& a, hipStream_t s);

// Plain (unweighted) pair-distance histograms through workgroup-private LDS bins.  Atoms
// sorted by type; a tile = rows [r0, r0 + nr) of one type x columns [c0, c0 + nc) of
// one type (same type: only pairs with column index > row index).  Bins are counted in
// windows of HIST_WINDOW (grid y = window), each flushed with one global add per
// nonzero bin.
constexpr int HIST_WINDOW = 16384;
constexpr int HIST_ROWS = 256;
struct HistTile {
    int r0, nr, c0, nc;
    int pair;  // histogram row (type pair)
    int diag;  // rows and columns from the same type segment
};
hipError_t launch_hist_tiles(const float4* sorted_atoms, const HistTile* tiles, int n_tiles, float inv_dr,
                             int n_bins, unsigned int* h0, hipStream_t s);
// out[x][k] = sum_b hist[x][b] sin(q_k r_b) / (q_k r_b) (sum_b hist[x][b] at q = 0), FP64, from
// hw = hist / r_b (the widening below); partial: hist_transform_scratch doubles
size_t hist_transform_scratch(int n_hist, int n_bins, int n_q);
hipError_t launch_hist_transform(const double* hw, int n_hist, int n_bins, double dr, const double* q, int n_q,
                                 double* partial, double* out, hipStream_t s);
// widen u32 / u64 histograms to doubles divided by the bin centre: dst[i] = src[i] / ((i mod n_bins + 0.5) dr)
hipError_t launch_widen_u32(const unsigned int* src, double* dst, long long n, int n_bins, double dr, hipStream_t s);
hipError_t launch_widen_u64(const unsigned long long* src, double* dst, long long n, int n_bins, double dr,
                            hipStream_t s);

// workspace reset kernels (used instead of memset nodes so launches can be graph-captured)
hipError_t launch_fill_i32(int* p, int value, long long n, hipStream_t s);

}  // namespace rsx
