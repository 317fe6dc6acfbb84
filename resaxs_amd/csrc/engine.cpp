// Host engine of the MI355X Debye-SAXS library and its C ABI (include/rsx_capi.h).
//
// Device layout (HBM):
//   cols      float4 [n]      atoms in "slot" order: grouped by form-factor class,
//                              {x, y, z, class bits}; the pair kernels read it both
//                              as rows (lanes) and as columns (scalar loads)
//   ftab      float  [n_class][ldf]  F_class(q_k), zero padded to ldf = n_chunks*K
//   partials  double [n_slots][ldf]  one row per pair job
//   iq        double [n_q]    running I(q) (FP64)
// Classes: without hydration class = type; with hydration
// class = type * (n_dots + 1) + exposed-dot count, so F_class(q) =
// f_type(q) + (count / n_dots) * w * f_water(q) exactly as
// resaxslib/src/gpu_saxs_pt_wf.cl:130-138 composes it.

#include "kernels.h"
#include "../../include/rsx_capi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

namespace {

using rsx::ChunkConst;
using rsx::CLS_SKIP;
using rsx::PairJob;

constexpr double PI = 3.14159265358979323846;

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg)
{
    throw Fail{code, msg};
}

void hipchk(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(RSX_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    void ensure(size_t n)
    {
        if (n <= cap && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
        if (e != hipSuccess) fail(RSX_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        cap = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// pinned host staging buffer: copies from it are truly asynchronous, so its contents
// must stay unchanged until the stream has passed the copy (every user synchronizes its
// stream before the next use of the same buffer)
template <typename T>
struct HBuf {
    T* p = nullptr;
    size_t cap = 0;
    void ensure(size_t n)
    {
        if (n <= cap && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (n == 0) n = 1;
        hipError_t e = hipHostMalloc((void**)&p, n * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) fail(RSX_OUT_OF_MEMORY, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        cap = n;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};


inline float bitsf(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// host-side generators (dot spheres and the default atomic factor table)
// ---------------------------------------------------------------------------

// rsasa_algorithm.cpp:96-116, float arithmetic of the reference's float build
void fibonacci_dots(int nhalf, float radius, float probe, float* out)
{
    const float phi = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const float ext = radius + probe;
    int d = 0;
    for (int i = -(nhalf - 1); i <= nhalf; ++i, ++d) {
        const float theta = 2.0f * (float)PI * (float)i / phi;
        const float sphi = (float)i / (float)nhalf;
        const float cphi = (float)(std::sqrt((double)((nhalf + i) * (nhalf - i))) / nhalf);
        out[4 * d + 0] = cphi * std::sin(theta) * ext;
        out[4 * d + 1] = cphi * std::cos(theta) * ext;
        out[4 * d + 2] = sphi * ext;
        out[4 * d + 3] = radius;
    }
}

// rsasa_algorithm.cpp:58-80: rings of dots at polar angles pi i / (n_sect / 2), each
// ring with about sin(phi) n_sect dots; float arithmetic throughout, as the
// reference's float instantiation evaluates it.  Returns the dot count; writes at
// most cap dots.
int sphere_dots(float radius, float n_sect, float probe, float* out, int cap)
{
    const float n_vert = 0.5f * n_sect;
    const float ext = radius + probe;
    int d = 0;
    for (unsigned i = 0; (float)i < n_vert; ++i) {
        const float phi = ((float)PI * (float)i) / n_vert;
        const float z = std::cos(phi) * ext;
        const float xy = std::sin(phi);
        const float ring = xy * n_sect;
        for (unsigned j = 0; (float)j < ring - 1.0f; ++j, ++d) {
            const float theta = (2.0f * (float)PI * (float)j) / ring;
            if (d < cap) {
                out[4 * d + 0] = xy * std::cos(theta) * ext;
                out[4 * d + 1] = xy * std::sin(theta) * ext;
                out[4 * d + 2] = z;
                out[4 * d + 3] = radius;
            }
        }
    }
    return d;
}

// rsasa_algorithm.cpp:157-208: dodecahedron (20) + icosahedron (12) vertices
void dodeca_icosa_dots(float radius, float probe, float* out)
{
    const float ext = radius + probe;
    const float phi = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const float v1 = (1.0f / std::sqrt(3.0f)) * ext;
    const float v2 = (1.0f / phi) * v1, v3 = phi * v1;
    const double cc = (double)phi * std::cos(3.0 * PI / 10.0);
    const float ico = (float)((double)((std::sqrt(5.0f) + 5.0f) / 10.0f) / std::sqrt(3.0 - 1.0 / std::pow(cc, 2.0)) *
                              (double)ext);
    const float a = ico, b = phi * ico;
    const float sgn[2] = {1.0f, -1.0f};
    int d = 0;
    auto put = [&](float x, float y, float z) {
        out[4 * d + 0] = x;
        out[4 * d + 1] = y;
        out[4 * d + 2] = z;
        out[4 * d + 3] = radius;
        ++d;
    };
    for (float sx : sgn)
        for (float sy : sgn)
            for (float sz : sgn) put(sx * v1, sy * v1, sz * v1);
    for (float sy : sgn)
        for (float sz : sgn) put(0.0f, sy * v2, sz * v3);
    for (float sx : sgn)
        for (float sy : sgn) put(sx * v2, sy * v3, 0.0f);
    for (float sx : sgn)
        for (float sz : sgn) put(sx * v3, 0.0f, sz * v2);
    for (float sx : sgn)
        for (float sz : sgn) put(sx * a, 0.0f, sz * b);
    for (float sy : sgn)
        for (float sz : sgn) put(0.0f, sy * b, sz * a);
    for (float sx : sgn)
        for (float sy : sgn) put(sx * b, sy * a, 0.0f);
}

// ff_coef.hpp:39-48 Cromer-Mann a1..a4, c, b1..b4 (H, C, N, O, P, S, Zn); the
// excluded volumes and complex types of constants.hpp:61-81.  Tabulated
// physical constants, kept identical to the reference's.
const float kCM[7][9] = {
    {0.4930020f, 0.3229120f, 0.1401910f, 4.0810000E-02f, 3.0380001E-03f, 10.51090f, 26.12570f, 3.142360f, 57.79970f},
    {2.310000f, 1.020000f, 1.588600f, 0.8650000f, 0.2156000f, 20.84390f, 10.20750f, 0.5687000f, 51.65120f},
    {12.21260f, 3.132200f, 2.012500f, 1.166300f, -11.52900f, 5.7000001E-03f, 9.893300f, 28.99750f, 0.5826000f},
    {3.048500f, 2.286800f, 1.546300f, 0.8670000f, 0.2508000f, 13.27710f, 5.701100f, 0.3239000f, 32.90890f},
    {6.434500f, 4.179100f, 1.780000f, 1.490800f, 1.114900f, 1.906700f, 27.15700f, 0.5260000f, 68.16450f},
    {6.905300f, 5.203400f, 1.437900f, 1.586300f, 0.8669000f, 1.467900f, 22.21510f, 0.2536000f, 56.17200f},
    {14.07430f, 7.031800f, 5.165200f, 2.410000f, 1.304100f, 3.265500f, 0.2333000f, 10.31630f, 58.70970f}};
const float kExclVol[7] = {5.15f, 16.44f, 2.49f, 9.13f, 5.73f, 19.86f, 9.85f};
const int kComplex[9][2] = {{1, 1}, {1, 2}, {1, 3}, {2, 1}, {2, 2}, {2, 3}, {3, 1}, {3, 2}, {5, 1}};
constexpr float kRho0 = 0.334f;
constexpr int ATOMIC_TYPES = 16;  // atom types of the default table (constants.hpp:28-51)
constexpr int ATOMIC_OH2 = 14;    // its water row, the OH2 group

// The exponentials and powers are evaluated in double from float arguments and
// rounded to float once per expression, the evaluation order of ff_coef.hpp:86-102
// under the C library's double exp/pow, so the table matches it bit for bit.
float excl_ff(int t, float k_sq)
{
    const double e = std::exp((double)(-(float)PI) * std::pow((double)kExclVol[t], (double)(2.0f / 3.0f)) * (double)k_sq);
    return (float)((double)(kRho0 * kExclVol[t]) * e);
}

void atomic_radii(float* out16)
{
    float fe[16];
    for (int t = 0; t < 7; ++t) fe[t] = excl_ff(t, 0.0f);
    for (int t = 7; t < 16; ++t) fe[t] = fe[kComplex[t - 7][0]] + (float)kComplex[t - 7][1] * fe[0];
    const float coef = (float)(3.0 / (4.0 * PI * 0.334f));
    for (int t = 0; t < 16; ++t) out16[t] = std::pow(coef * fe[t], (float)(1.0 / 3));
}

void atomic_ff(const float* q, int nq, float avg_radius, float c, float* out)
{
    const float k43 = (float)std::pow((double)(4.0f * (float)PI / 3.0f), (double)(2.0f / 3.0f));
    for (int i = 0; i < nq; ++i) {
        float* page = out + (size_t)i * 16;
        const float k = q[i] / (4.0f * (float)PI);
        const float k_sq = (float)(k * k);
        float E = 1.0f;
        if (c != 1.0f) {
            const float c_sq = c * c;
            const float arg = -(float)PI * k43 * k_sq * avg_radius * avg_radius * (c_sq - 1.0f);
            E = (float)((double)(c_sq * c) * std::exp((double)arg));
        }
        for (int t = 0; t < 7; ++t) {
            float f = kCM[t][4];
            for (int a = 0; a < 4; ++a) f = (float)((double)f + (double)kCM[t][a] * std::exp((double)(-kCM[t][5 + a] * k_sq)));
            page[t] = f - excl_ff(t, k_sq) * E;
        }
        for (int t = 7; t < 16; ++t) page[t] = page[kComplex[t - 7][0]] + (float)kComplex[t - 7][1] * page[0];
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
struct rsx_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of incremental updates (default priority, like the first): the
    // new-records launch runs there beside the old-records launch
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_skip = nullptr, ev_new = nullptr;
    hipEvent_t ev_out = nullptr;  // recorded after the download of I(q)
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    std::string err;
    int state = 0;  // 0 blank, 1 initialized, 3 args set, 4 computing

    // q grid
    std::vector<float> q;
    int nq = 0;
    bool uniform_grid = false;  // q_k = q_0 + k dq
    bool uniform = false;       // the recurrence path is active
    double centre[3] = {0, 0, 0};
    double radius = 0.0;        // bounding radius of the current structure
    double dq = 0.0;
    int q0_index = -1;

    // factors
    std::vector<float> table;
    int n_types = 0;
    int water = -1;
    float w = 0.0f;
    bool atomic_table = false;  // table = the default atomic factors (rsx_set_atomic_factors)
    float avg_radius = 1.62f;   // their mean atomic radius
    float expansion = 1.0f;     // excluded-volume expansion factor c
    bool params_dirty = true;  // factors / weight changed since the last full calc
    bool w_dirty = false;
    bool no_rsasa_recalc = false;

    // SASA set-up
    bool sasa_set = false;
    std::vector<float> radii;  // per type
    float probe = 1.8f;
    int dot_kind = RSX_DOTS_DODECA_ICOSA;
    int n_half = 0;
    int n_dots = 0;
    std::vector<float> dots;  // [n_types][n_dots][4]
    std::vector<float> r2, rext;

    // tuning
    int k_req = 0, range_req = 0;
    int k_max = 0;  // cap on K while the gradient forward is active (0: none)
    int variant = rsx::VARIANT_FP32R;  // recurrence kernel (rsx::VARIANT_*)
    int K = 0, n_chunks = 0, ldf = 0;
    double path_radius = 0.0;  // bounding radius the current path and series length were chosen for
    // full recomputes of the running incremental I(q): its FP64 sum of FP32 increments was
    // measured not to drift (50k atoms, 2000 committed 500-atom moves with and without
    // hydration: 2.6e-6 / 3.0e-6 after 199 moves, 1.9e-6 / 2.3e-6 after 1999, against a fresh
    // FP64 full recompute; the error is the first forward's), so the period only bounds
    // drift over very long runs
    int rebase_every = 10000;

    // bodies (host mirror, caller order)
    int n = 0;
    std::vector<float> bodies;  // n*4, type bits in w
    std::vector<uint32_t> type;
    std::vector<uint32_t> counts;
    std::vector<float> rsasa;
    bool have_rsasa = false;
    std::vector<uint32_t> cls;      // class per caller index
    std::vector<int> slot_of, caller_of;
    int d1 = 1;                      // classes per type
    int n_class = 0;
    std::vector<int> cls_count;      // atoms per class (built on demand from cls)
    bool cls_count_ok = false;
    std::vector<double> f64;         // [n_class][nq] FP64 class factors
    bool hydrate = false;

    // device buffers
    DBuf<float4> d_cols, d_saved_cols, d_sasa_atoms, d_sorted_atoms;
    bool sasa_mirror = false;        // d_sasa_atoms and d_counts hold exactly the host bodies and counts
    // local SASA recount: the listed atoms, and the block read back (two counter sets of
    // rsx::LOCAL_HDR_INTS, then (atom, count) x LOCAL_CHANGED_CAP)
    DBuf<int> d_local_list, d_local;
    int l_parity = 0;                // counter set of the next local recount
    int l_used_parity = 0;           // counter set of the last one enqueued
    // slab cell list of the local recounts (rsx::SasaArgs::slab): records, per-bucket counts,
    // each atom's record.  slab_age counts the recounts since it was built from the device
    // atoms (-1: not built for them); slab_off: the atoms overflowed a bucket right after a
    // build, so local recounts rebuild the hashed cell list instead
    DBuf<float4> d_slab;
    DBuf<int> d_slab_count, d_slab_pos;
    int slab_age = -1;
    bool slab_off = false;
    const void* slab_built[3] = {nullptr, nullptr, nullptr};  // slab, counts, positions built
    int slab_built_nb = 0, slab_built_n = 0;
    float slab_built_cell = 0.0f;
    hipEvent_t ev_sasa = nullptr;    // recorded after the local recount's read-back
    // the local recount's stage (rsx::LocalStage + moved atoms) and its captured graph,
    // kept while its arguments and buffers are unchanged
    DBuf<unsigned char> d_lstage;
    HBuf<unsigned char> pin_lstage;
    hipGraph_t l_graph = nullptr;
    hipGraphExec_t l_exec = nullptr;
    rsx::SasaArgs l_args;
    const void* l_pin_stage = nullptr;
    const void* l_pin_local = nullptr;
    DBuf<unsigned int> h_u32;        // distance histograms (rsx_calc_curve_hist)
    DBuf<unsigned long long> h_u64;
    DBuf<double> h_dbl, h_out, h_q, h_part;
    DBuf<int> h_pidx;
    DBuf<rsx::HistTile> h_tiles;
    DBuf<float> d_ftab, d_qv, d_qsplit, d_r2, d_rext;
    DBuf<double> d_qrev;
    DBuf<ChunkConst> d_chunks;
    DBuf<double> d_invq, d_partials, d_iq, d_saved_iq, d_out, d_wq, d_q0;
    DBuf<double> d_red;              // chunk sums of the two-stage reduction (single structure)
    DBuf<PairJob> d_fjobs;          // forward (triangle) jobs
    DBuf<double> d_f0col;           // FP64 F_class(q = 0) per class
    DBuf<int> d_slot_begin_f;  // [0, n forward jobs]
    DBuf<int> d_slot_begin, d_bucket_of, d_bucket_count, d_bucket_start, d_bucket_fill, d_sorted, d_s2c,
        d_tile_sum;
    DBuf<uint32_t> d_counts;
    DBuf<float4> d_dots;
    DBuf<float> d_grad;
    // adjoint layout (adjoint_layout), rebuilt when layout_stamp changes
    DBuf<float> a_ftab, a_rowpart, a_colpart, a_invq32, a_cb, a_cg;
    bool adj_fp32 = true;  // adjoint kernel precision (rsx_set_adjoint_tuning)
    bool adj_pair_ok = true;  // FP32: class-pair coefficient form where the layout allows it
    int adj_k_req = 0, adj_range_req = 0;
    bool a_fp32 = false;   // precision the cached adjoint layout was built for
    bool a_pair = false;   // the cached layout runs the class-pair form (a_coef)
    DBuf<float> a_coef;
    DBuf<ChunkConst> a_chunks;
    DBuf<double> a_invq;
    DBuf<PairJob> a_jobs;
    DBuf<int> a_tri;
    int a_K = 0, a_nch = 0, a_ldf = 0, a_L = 0, a_njobs = 0, a_stride = 0, a_n = -1;
    uint64_t a_stamp = ~0ull;
    const float4* a_cols = nullptr;  // the column buffer the adjoint jobs point into
    bool a_have_w = false;           // d_wq holds the weights of the last gradient call
    bool persistent = true;          // batched launch: persistent work-queue kernel (FP32 variant)
    int n_cu = 256;                  // compute units of the device
    DBuf<int> d_queue;               // work counter of the persistent launch
    DBuf<int> d_guard;               // checks of device-resident bodies (rsx::GUARD_*)
    bool dev_pending = false;        // a device-resident forward ran since the last synchronize
    DBuf<int> d_cs_table, d_cs_tot;  // scratch of the device class sort
    // a device-resident forward overwrote the columns, slot map or running I(q): the host
    // bodies' state has to be laid out again before an incremental update or a commit
    bool host_stale = false;
    uint64_t layout_stamp = 0;  // bumped whenever the class table or the slot layout changes
    size_t ftab_rows = 0;

    std::string tune_report;
    int last_L = 0;  // column range of the last forward job layout
    // job tables of incremental updates, reused while their shape recurs (the changed set's
    // size varies from move to move with hydration, over a few row-tile counts)
    struct IncKey {
        int part = -1, nrows = -1, n = -1, L = -1, K = -1, n_chunks = -1, variant = -1, uniform = -1;
        const float4* rows = nullptr;
        const float4* cols = nullptr;
        bool operator==(const IncKey& o) const
        {
            return part == o.part && nrows == o.nrows && n == o.n && L == o.L && K == o.K && n_chunks == o.n_chunks &&
                   variant == o.variant && uniform == o.uniform && rows == o.rows && cols == o.cols;
        }
    };
    struct IncTable {
        IncKey key;
        DBuf<PairJob> jobs;
        int njobs = 0;
        uint64_t used = 0;
    };
    static constexpr int INC_TABLES = 12;  // least recently used one replaced
    IncTable inc_tab[INC_TABLES];
    uint64_t inc_clock = 0;
    std::vector<PairJob> inc_jobs;  // host table of the incremental jobs (capacity kept)
    // Incremental updates: the arrays of each launch packed into one pinned block and copied
    // with one transfer into a device block that only grows, so arrays at fixed offsets keep
    // their addresses from update to update (the cached job tables point into them).
    struct Staging {
        HBuf<unsigned char> host;
        DBuf<unsigned char> dev;
    };
    Staging st_old, st_new;
    DBuf<double> d_partials_new;  // partials of an update's second launch
    // pinned staging of the per-update transfers (job tables: one per launch of an update)
    HBuf<float4> pin_moved;
    HBuf<PairJob> pin_jobs[2];
    HBuf<int> pin_local;
    HBuf<double> pin_iq;
    std::vector<char> in_m;                       // membership of the changed set (all zero between calls)
    // host sources of small uploads that may still be in flight when launch_jobs returns
    int up_sb[2] = {0, 0};
    double up_q0 = 0.0;

    // last forward launch (for timing)
    std::vector<PairJob> last_jobs;
    // what the forward job table on the device (d_fjobs, last_jobs) was built for: atoms,
    // column range, column buffer, kernel (variant, K, chunks) and path
    struct FwdKey {
        int n = -1, L = 0, variant = -1, K = 0, n_chunks = 0, uniform = -1;
        const void* cols = nullptr;
        bool operator==(const FwdKey& o) const
        {
            return n == o.n && L == o.L && variant == o.variant && K == o.K && n_chunks == o.n_chunks &&
                   uniform == o.uniform && cols == o.cols;
        }
    } fjobs_key;
    int last_slots = 0;
    double last_q0 = 0.0;

    // transaction
    bool xaction = false;
    struct Saved {
        std::vector<float> bodies;
        std::vector<uint32_t> counts, cls;
        std::vector<float> rsasa;
        std::vector<int> slot_of, caller_of;
        bool have_rsasa = false;
        int since_rebase = 0;
    } saved;
    // bodies / counts / cls / rsasa entries written by incremental updates since the last
    // commit or revert, which then copy only those; dirty_full: a path rewrote whole
    // arrays or the slot layout since then
    std::vector<int> dirty;
    bool dirty_full = true;
    bool slots_dirty = false;  // incremental updates moved atoms between slots since the last transaction
    std::vector<char> slot_mark;                  // slots touched by a relocation (all zero between calls)
    int since_rebase = 0;

    // inputs the batch layout depends on, stamped when they change
    uint64_t factor_stamp = 1;  // factor table, water row or weight
    uint64_t sasa_stamp = 1;    // SASA radii, probe or dot sphere

    // batch: the frames as given, and a layout of its own (path, q chunks, classes, factor
    // table, jobs), so single-structure calls on the same engine never touch it
    int b_frames = 0;
    std::vector<int> b_offsets;
    std::vector<float> b_bodies;     // frames as given, {x, y, z, type bits}
    bool b_layout_ok = false;        // the layout matches q, the tuning and the SASA set-up
    uint64_t b_gen = 0;              // layouts built (the batched VJP's layout keys on it)
    std::vector<int> b_s2c;          // caller index (position in the frames as given) per batch column
    // batched VJP (rsx_batch_vjp): the adjoint's own q chunking and pair jobs over every frame,
    // per-frame triangle tables, the class tables 2 F_c,k and 2 F_c,k / q_k, partials
    uint64_t ba_gen = 0, ba_factor = 0;
    int ba_K = 0, ba_nch = 0, ba_ldf = 0, ba_L = 0, ba_njobs = 0, ba_n = 0;
    bool ba_pair = false;
    DBuf<float> ba_ftab, ba_cb, ba_cg, ba_invq32, ba_w, ba_rowpart, ba_colpart, ba_grad, ba_coef;
    DBuf<double> ba_invq, ba_ones;
    DBuf<ChunkConst> ba_chunks;
    DBuf<PairJob> ba_jobs;
    DBuf<int> ba_tri, ba_tri_off, ba_atom_frame, ba_frame_off, ba_s2c;
    bool ba_have_w = false;
    uint64_t b_factor_stamp = 0;     // factor_stamp the class table was built from
    uint64_t b_sasa_stamp = 0;
    int b_nq = 0, b_n_types = 0;
    bool b_hydrate = false;          // classes type x exposed-dot count
    int b_d1 = 1, b_n_class = 0;
    std::vector<uint32_t> b_counts;  // exposed-dot count per body (hydrated layouts)
    std::vector<int> b_pop_off;      // frame f's class populations: b_pop[b_pop_off[f] .. b_pop_off[f+1])
    std::vector<std::pair<int, int>> b_pop;  // (class, atoms)
    bool b_uniform = false;
    int b_variant = 0, b_K = 0, b_n_chunks = 0, b_ldf = 0, b_L = 0;
    double b_radius = 0.0;
    DBuf<ChunkConst> b_chunks;
    DBuf<double> b_invq, b_qrev;
    DBuf<float> b_qv, b_qsplit, b_ftab;
    HBuf<float> b_pin_ftab;
    HBuf<double> b_pin_q0;
    std::vector<double> b_f64;       // [class][n_q] FP64 class factors of the current table
    // batch tuning (rsx_batch_autotune / its cache): variant, K, column range, launch form;
    // -1 / 0 = follow the engine
    int b_var_req = -1, b_k_req = 0, b_range_req = 0, b_pers_req = -1;
    DBuf<float4> b_cols;
    DBuf<PairJob> b_jobs;
    DBuf<double> b_partials, b_out, b_mean, b_q0, b_red;
    DBuf<int> b_slot_begin;
    int b_njobs = 0, b_slots = 0, b_max_slots = 0;
    hipGraph_t b_graph = nullptr;
    hipGraphExec_t b_exec = nullptr;
    double* b_graph_out = nullptr;
    double* b_graph_mean = nullptr;
    int b_pers_waves = -1;  // launch shape of the captured batch graph (-1: none yet)
};

namespace {

void set_device(rsx_engine* h)
{
    hipchk(hipSetDevice(h->device), "hipSetDevice");
}

void mark_dirty_full(rsx_engine* h)
{
    h->dirty_full = true;
    h->dirty.clear();
    h->slots_dirty = false;  // the full copy of a transaction takes the slot maps too
}

// record entry i before an incremental update writes it
void mark_dirty(rsx_engine* h, int i)
{
    if (h->dirty_full) return;
    if (4 * h->dirty.size() >= (size_t)h->n) {
        mark_dirty_full(h);
        return;
    }
    h->dirty.push_back(i);
}

// ---- q grid and chunking -------------------------------------------------
// K list of a pair kernel: the path of a (uniform grid?, recurrence variant) choice
int pair_path(bool uniform, int variant)
{
    return !uniform                          ? rsx::PATH_DIRECT
           : variant == rsx::VARIANT_FP32  ? rsx::PATH_REC32
           : variant == rsx::VARIANT_FP32R ? rsx::PATH_REC32R
                                           : rsx::PATH_REC64;
}

int pair_path(const rsx_engine* h)
{
    return pair_path(h->uniform, h->variant);
}

// q per thread K (and the chunk count C) of a pair kernel over the engine's n_q: the
// requested K when k_req > 0, else the cheapest of the kernel's K list at most k_max (> 0)
void pick_chunks(const rsx_engine* h, bool uni, int variant, int k_req, int k_max, int& K_out, int& C_out)
{
    const int path = pair_path(uni, variant);
    const int nk = rsx::debye_k_count(path);
    int bestK = 0, bestC = 0;
    double best = 1e300;
    for (int i = 0; i < nk; ++i) {
        const int K = rsx::debye_k_value(path, i);
        if (k_req > 0 && K != k_req) continue;
        if (k_max > 0 && K > k_max) continue;
        if (k_req <= 0 && !uni && K > 48) continue;  // the direct kernel spills at K = 64
        const int C = (h->nq + K - 1) / K;
        // cost per (pair, chunk) in FP64-op units: the FP64 recurrence 2K + ~44 for distance
        // and seeds; the FP32 one 3K FP32 ops (~0.6 of an FP64 op each) + ~25
        const double cost = !uni ? C * (6.0 * K + 12.0)
                            : path == rsx::PATH_REC32 || path == rsx::PATH_REC32R ? C * (1.8 * K + 25.0)
                                                      : C * (2.0 * K + 44.0);
        if (cost < best) {
            best = cost;
            bestK = K;
            bestC = C;
        }
    }
    if (bestK == 0) fail(RSX_INVALID_ARG, "unsupported K (q values per thread)");
    K_out = bestK;
    C_out = bestC;
}

void choose_chunks(rsx_engine* h)
{
    pick_chunks(h, h->uniform, h->variant, h->k_req, h->k_max, h->K, h->n_chunks);
    h->ldf = h->K * h->n_chunks;
}

// Seed constants of the q chunk starting at q index k0: a = q_k0 and the phases per unit r
// of a and a + dq/2 in revolutions.
ChunkConst chunk_const(const rsx_engine* h, int k0)
{
    ChunkConst c;
    const double a = k0 < h->nq ? (double)h->q[k0] : 0.0;
    c.a = (float)a;
    c.a_rev = a / (2.0 * PI);
    c.mid_rev = (a + 0.5 * h->dq) / (2.0 * PI);
    c.pad = 0.0f;
    return c;
}

// Per-chunk constants and the q tables ([ldf], zero padded) of a chunking of the
// engine's q grid into n_chunks runs of K.
void upload_q_consts(const rsx_engine* h, int K, int n_chunks, DBuf<ChunkConst>& chunks, DBuf<double>& d_invq,
                     DBuf<double>& d_qrev, DBuf<float>& d_qsplit, DBuf<float>& d_qv)
{
    const int nq = h->nq, ldf = K * n_chunks;
    std::vector<ChunkConst> cc(n_chunks);
    for (int c = 0; c < n_chunks; ++c) cc[c] = chunk_const(h, c * K);
    std::vector<double> invq(ldf, 0.0), qrev(ldf, 0.0);
    std::vector<float> qv(ldf, 0.0f), qsplit(2 * (size_t)ldf, 0.0f);
    for (int k = 0; k < nq; ++k) {
        invq[k] = h->q[k] != 0.0f ? 1.0 / (double)h->q[k] : 0.0;
        qrev[k] = (double)h->q[k] / (2.0 * PI);
        qv[k] = h->q[k];
        qsplit[2 * k] = (float)qrev[k];
        qsplit[2 * k + 1] = (float)(qrev[k] - (double)qsplit[2 * k]);
    }
    chunks.ensure(cc.size());
    d_invq.ensure(ldf);
    d_qrev.ensure(ldf);
    d_qv.ensure(ldf);
    d_qsplit.ensure(2 * (size_t)ldf);
    hipchk(hipMemcpy(chunks.p, cc.data(), cc.size() * sizeof(ChunkConst), hipMemcpyHostToDevice), "upload chunks");
    hipchk(hipMemcpy(d_qsplit.p, qsplit.data(), 2 * (size_t)ldf * sizeof(float), hipMemcpyHostToDevice),
           "upload qsplit");
    hipchk(hipMemcpy(d_invq.p, invq.data(), ldf * sizeof(double), hipMemcpyHostToDevice), "upload invq");
    hipchk(hipMemcpy(d_qrev.p, qrev.data(), ldf * sizeof(double), hipMemcpyHostToDevice), "upload qrev");
    hipchk(hipMemcpy(d_qv.p, qv.data(), ldf * sizeof(float), hipMemcpyHostToDevice), "upload qv");
}

// Upload the per-chunk constants of the active path (uniform recurrence or direct).
void apply_path(rsx_engine* h, bool uniform)
{
    h->layout_stamp++;
    h->uniform = uniform;
    choose_chunks(h);
    upload_q_consts(h, h->K, h->n_chunks, h->d_chunks, h->d_invq, h->d_qrev, h->d_qsplit, h->d_qv);
    h->params_dirty = true;  // the class table is laid out with stride ldf
}

void setup_q(rsx_engine* h)
{
    const int nq = h->nq;
    h->q0_index = -1;
    for (int k = 0; k < nq; ++k)
        if (h->q[k] == 0.0f) {
            h->q0_index = k;
            break;
        }
    // uniform grid test: every q within 4 ulps (of the largest |q|) of q_0 + k dq
    h->uniform_grid = false;
    h->dq = 0.0;
    if (nq >= 2) {
        const double dq = ((double)h->q[nq - 1] - (double)h->q[0]) / (nq - 1);
        double qmax = 0.0;
        for (float v : h->q) qmax = std::max(qmax, (double)std::fabs(v));
        const double tol = 4.0 * std::ldexp(1.0, -23) * qmax;
        bool ok = dq > 0.0 && h->q[0] >= 0.0f;
        for (int k = 0; k < nq && ok; ++k)
            if (std::fabs((double)h->q[k] - ((double)h->q[0] + k * dq)) > tol) ok = false;
        h->uniform_grid = ok;
        h->dq = ok ? dq : 0.0;
    }
    apply_path(h, h->uniform_grid);
    h->d_iq.ensure(nq);
    h->d_saved_iq.ensure(nq);
    h->d_out.ensure(nq);
    h->d_q0.ensure(64);
}

// The recurrence's step angle dq r must stay below 3 rad (its FP64 sin
// polynomial covers dq r / 2 <= 1.5 and the three-term form loses accuracy
// as dq r -> pi).  r <= 2 R for R the largest distance from the centroid.
constexpr double MAX_STEP_ANGLE = 3.0;

bool fits_recurrence(const rsx_engine* h, double radius)
{
    return h->uniform_grid && h->dq * 2.0 * radius <= MAX_STEP_ANGLE;
}

// The direct kernel may form its phases in FP32 when every q r / (2 pi) with r <= 2 R
// stays within rsx::DIRECT_PHASE32_MAX_REV revolutions.
bool phase32_ok(const rsx_engine* h, double radius)
{
    double qmax = 0.0;
    for (float v : h->q) qmax = std::max(qmax, (double)std::fabs(v));
    return qmax * 2.0 * radius / (2.0 * PI) <= rsx::DIRECT_PHASE32_MAX_REV;
}

double bounding_radius(const float* bodies, int n, double* centre)
{
    double c[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) c[d] += bodies[4 * (size_t)i + d];
    for (int d = 0; d < 3; ++d) c[d] = n ? c[d] / n : 0.0;
    double r2 = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double v = bodies[4 * (size_t)i + d] - c[d];
            s += v * v;
        }
        r2 = std::max(r2, s);
    }
    if (centre)
        for (int d = 0; d < 3; ++d) centre[d] = c[d];
    return std::sqrt(r2) * (1.0 + 1e-6) + 1e-6;
}

// choose the path for a structure of the given bounding radius; true if it changed
bool select_path(rsx_engine* h, double radius)
{
    h->path_radius = radius;
    const bool want = fits_recurrence(h, radius);
    if (want == h->uniform) return false;
    apply_path(h, want);
    return true;
}

// ---- classes ---------------------------------------------------------------
bool hydration_on(const rsx_engine* h)
{
    return h->w != 0.0f && h->water >= 0 && h->sasa_set;
}

// Factor rows of the classes t * d1 + c (t < n_types, c < d1): FP64 [class][n_q] and
// FP32 [class][ldf] (ft must hold n_types * d1 rows of ldf, zero padded by the caller).
// With hydration (d1 = n_dots + 1) class count c has F = f_t + (c / n_dots) w f_water,
// the exposed fraction formed as the reference forms it, float(count) / n_dots; a table
// without a water row contributes nothing there.
void class_factor_rows(const rsx_engine* h, int d1, int ldf, std::vector<double>& f64, float* ft)
{
    const int nq = h->nq;
    const bool hyd = d1 > 1;
    f64.assign((size_t)h->n_types * d1 * nq, 0.0);
    for (int t = 0; t < h->n_types; ++t)
        for (int c = 0; c < d1; ++c) {
            const int id = t * d1 + c;
            const float s = hyd ? (float)c / (float)h->n_dots : 0.0f;
            for (int k = 0; k < nq; ++k) {
                double f = h->table[(size_t)k * h->n_types + t];
                if (hyd && h->water >= 0)
                    f += (double)s * ((double)h->w * (double)h->table[(size_t)k * h->n_types + h->water]);
                f64[(size_t)id * nq + k] = f;
                ft[(size_t)id * ldf + k] = (float)f;
            }
        }
}

void build_class_table(rsx_engine* h)
{
    h->layout_stamp++;
    h->cls_count_ok = false;
    for (auto& t : h->inc_tab) t.njobs = 0;
    h->hydrate = hydration_on(h);
    h->d1 = h->hydrate ? h->n_dots + 1 : 1;
    h->n_class = h->n_types * h->d1;
    const int nq = h->nq, ldf = h->ldf;
    // + an all-zero row; with hydration + 2 D + 1 delta rows (delta_class)
    const int n_delta = h->hydrate ? 2 * h->n_dots + 1 : 0;
    std::vector<float> ft((size_t)(h->n_class + 1 + n_delta) * ldf, 0.0f);
    class_factor_rows(h, h->d1, ldf, h->f64, ft.data());
    // delta rows: the factor change of an atom whose exposed-dot count changed by d,
    // (d / D) w f_water(q), for d in [-D, D]
    for (int d = -h->n_dots; h->hydrate && d <= h->n_dots; ++d) {
        const size_t id = (size_t)h->n_class + 1 + (size_t)(d + h->n_dots);
        for (int k = 0; k < nq; ++k)
            ft[id * ldf + k] = (float)((double)d / (double)h->n_dots *
                                       ((double)h->w * (double)h->table[(size_t)k * h->n_types + h->water]));
    }
    h->d_ftab.ensure(ft.size());
    hipchk(hipMemcpy(h->d_ftab.p, ft.data(), ft.size() * sizeof(float), hipMemcpyHostToDevice), "upload ftab");
    std::vector<double> f0(h->n_class, 0.0);
    if (h->q0_index >= 0)
        for (int c = 0; c < h->n_class; ++c) f0[c] = h->f64[(size_t)c * nq + h->q0_index];
    h->d_f0col.ensure(h->n_class);
    hipchk(hipMemcpy(h->d_f0col.p, f0.data(), f0.size() * sizeof(double), hipMemcpyHostToDevice), "upload f0");
}

// class id of a delta row: an exposed-dot count change of d (hydration layouts only)
uint32_t delta_class(const rsx_engine* h, int d)
{
    return (uint32_t)(h->n_class + 1 + d + h->n_dots);
}

uint32_t class_of(const rsx_engine* h, int i)
{
    return h->type[i] * (uint32_t)h->d1 + (h->hydrate ? h->counts[i] : 0u);
}

// stable sort of atoms by class -> slot maps; upload the column records
void assign_slots(rsx_engine* h)
{
    const int n = h->n;
    mark_dirty_full(h);
    h->cls.resize(n);
    for (int i = 0; i < n; ++i) h->cls[i] = class_of(h, i);
    h->cls_count_ok = false;
    std::vector<int> cnt(h->n_class + 1, 0);
    for (int i = 0; i < n; ++i) cnt[h->cls[i] + 1]++;
    for (int c = 0; c < h->n_class; ++c) cnt[c + 1] += cnt[c];
    mark_dirty_full(h);
    h->caller_of.resize(n);
    h->slot_of.resize(n);
    for (int i = 0; i < n; ++i) {
        const int s = cnt[h->cls[i]]++;
        h->caller_of[s] = i;
        h->slot_of[i] = s;
    }
    std::vector<float4> cols(n);
    for (int s = 0; s < n; ++s) {
        const int i = h->caller_of[s];
        cols[s] = make_float4(h->bodies[4 * i], h->bodies[4 * i + 1], h->bodies[4 * i + 2], bitsf(h->cls[i]));
    }
    h->d_cols.ensure(n);
    hipchk(hipMemcpyAsync(h->d_cols.p, cols.data(), n * sizeof(float4), hipMemcpyHostToDevice, h->stream),
           "upload cols");
    h->d_s2c.ensure(n);
    hipchk(hipMemcpyAsync(h->d_s2c.p, h->caller_of.data(), n * sizeof(int), hipMemcpyHostToDevice, h->stream),
           "upload slot map");
    hipchk(hipStreamSynchronize(h->stream), "sync");
}

// sum of F(q = 0) over the atoms, as sum over classes of (atoms in the class) x F_class(0)
// in class order: O(n_class), and the same bits for the same class populations however
// they were reached (full layouts and incremental updates agree)
double sum_f0(rsx_engine* h)
{
    if (h->q0_index < 0) return 0.0;
    if (!h->cls_count_ok || (int)h->cls_count.size() != h->n_class) {
        h->cls_count.assign(h->n_class, 0);
        for (uint32_t c : h->cls)
            if (c != CLS_SKIP && (int)c < h->n_class) h->cls_count[c]++;
        h->cls_count_ok = true;
    }
    double s = 0.0;
    for (int c = 0; c < h->n_class; ++c)
        if (h->cls_count[c]) s += (double)h->cls_count[c] * h->f64[(size_t)c * h->nq + h->q0_index];
    return s;
}

// ---- SASA ----------------------------------------------------------------
// changed counts a local recount reads back in its one transfer (more: all counts compared)
constexpr int LOCAL_CHANGED_CAP = 4096;
// local recounts between rebuilds of the slab (clears the tombstones moves leave in it)
constexpr int SLAB_REBUILD = 256;
// largest structure whose local recounts use the slab: it holds 64 float4 records per bucket
// and a power of two >= n buckets, 1 KB per atom (4 GiB at this size); larger structures
// rebuild the hashed cell list instead
constexpr int SLAB_MAX_ATOMS = 1 << 22;
// moved atoms a local recount's captured graph carries in its stage (more: eager launches)
constexpr int LOCAL_PATCH_CAP = 4096;

void upload_sasa_tables(rsx_engine* h)
{
    const int nt = (int)h->radii.size();
    h->r2.resize(nt);
    h->rext.resize(nt);
    for (int t = 0; t < nt; ++t) {
        const float R = h->probe + h->radii[t];
        h->r2[t] = R * R;
        h->rext[t] = (float)std::sqrt((double)h->r2[t]);
    }
    h->d_r2.ensure(nt);
    h->d_rext.ensure(nt);
    h->d_dots.ensure((size_t)nt * h->n_dots);
    hipchk(hipMemcpy(h->d_r2.p, h->r2.data(), nt * sizeof(float), hipMemcpyHostToDevice), "upload r2");
    hipchk(hipMemcpy(h->d_rext.p, h->rext.data(), nt * sizeof(float), hipMemcpyHostToDevice), "upload rext");
    hipchk(hipMemcpy(h->d_dots.p, h->dots.data(), h->dots.size() * sizeof(float), hipMemcpyHostToDevice),
           "upload dots");
}

static_assert(sizeof(rsx::LocalStage) <= rsx::LOCAL_STAGE_HEADER, "stage header");

void release_local_graph(rsx_engine* h)
{
    if (h->l_exec) (void)hipGraphExecDestroy(h->l_exec);
    if (h->l_graph) (void)hipGraphDestroy(h->l_graph);
    h->l_exec = nullptr;
    h->l_graph = nullptr;
}

// SASA arguments over the engine's buffers for n atoms in d_sasa_atoms (zeroed first, so
// that two argument sets compare byte for byte).  local_cap >= 0: a local recount (the
// changed counts listed in d_local, the box and moved range in d_lstage; with
// local_cap > 0 the first kernel also writes up to local_cap moved atoms from the stage),
// on the slab cell list when slab is set.
rsx::SasaArgs sasa_args(rsx_engine* h, int n, int local_cap = -1, bool slab = false)
{
    float rmax = 0.0f;
    for (float v : h->rext) rmax = std::max(rmax, v);
    int nb = 1024;
    while (nb < 2 * n) nb <<= 1;
    h->d_bucket_of.ensure(n);
    h->d_bucket_count.ensure(nb + 1);
    h->d_bucket_start.ensure(nb + 1);
    h->d_bucket_fill.ensure(nb);
    h->d_tile_sum.ensure(nb / 1024);
    h->d_sorted.ensure(n);
    h->d_sorted_atoms.ensure(n);
    h->d_counts.ensure(n);
    rsx::SasaArgs a;
    std::memset(static_cast<void*>(&a), 0, sizeof a);
    a.atoms = h->d_sasa_atoms.p;
    a.n = n;
    a.n_types = (int)h->rext.size();
    a.r2 = h->d_r2.p;
    a.rext = h->d_rext.p;
    a.dots = h->d_dots.p;
    a.n_dots = h->n_dots;
    a.cell = 2.0f * rmax + 0.02f;
    a.inv_cell = 1.0f / a.cell;
    a.n_buckets = nb;
    a.bucket_of = h->d_bucket_of.p;
    a.bucket_count = h->d_bucket_count.p;
    a.bucket_start = h->d_bucket_start.p;
    a.bucket_fill = h->d_bucket_fill.p;
    a.tile_sum = h->d_tile_sum.p;
    a.sorted = h->d_sorted.p;
    a.sorted_atoms = h->d_sorted_atoms.p;
    a.counts = h->d_counts.p;
    if (local_cap >= 0) {
        h->d_local_list.ensure(n);
        const size_t local_ints = rsx::LOCAL_HDR_TOTAL + 2 * (size_t)LOCAL_CHANGED_CAP;
        if (h->d_local.cap < local_ints) {
            h->d_local.ensure(local_ints);
            // counter sets start cleared (each recount then clears the next one's)
            hipchk(hipMemsetAsync(h->d_local.p, 0, rsx::LOCAL_HDR_TOTAL * sizeof(int), h->stream), "clear counters");
        }
        h->d_lstage.ensure(rsx::LOCAL_STAGE_HEADER + 16 * (size_t)LOCAL_PATCH_CAP);
        a.local_hdr = h->d_local.p;
        a.local_list = h->d_local_list.p;
        a.local_changed = h->d_local.p + rsx::LOCAL_HDR_TOTAL;
        a.changed_cap = LOCAL_CHANGED_CAP;
        a.local_stage = reinterpret_cast<const rsx::LocalStage*>(h->d_lstage.p);
        a.atoms_w = h->d_sasa_atoms.p;
        a.patch_cap = local_cap;
        if (slab) {
            int nbs = 1024;
            while (nbs < n) nbs <<= 1;
            h->d_slab.ensure((size_t)nbs * rsx::SLAB_CAP);
            h->d_slab_count.ensure(nbs);
            h->d_slab_pos.ensure(n);
            a.slab = h->d_slab.p;
            a.slab_count = h->d_slab_count.p;
            a.slab_pos = h->d_slab_pos.p;
            a.slab_buckets = nbs;
        }
    }
    return a;
}

// counts for n atoms already on the device in d_sasa_atoms
void run_sasa_device(rsx_engine* h, int n)
{
    hipchk(rsx::launch_sasa(sasa_args(h, n), h->stream), "sasa kernels");
}

// recompute counts / rsasa for the host mirror bodies (caller order)
void compute_sasa(rsx_engine* h)
{
    const int n = h->n;
    for (int i = 0; i < n; ++i)
        if (h->type[i] >= h->radii.size()) fail(RSX_INVALID_ARG, "body type index outside the SASA radius table");
    h->d_sasa_atoms.ensure(n);
    hipchk(hipMemcpyAsync(h->d_sasa_atoms.p, h->bodies.data(), (size_t)n * 16, hipMemcpyHostToDevice, h->stream),
           "upload sasa atoms");
    h->sasa_mirror = true;
    h->slab_age = -1;
    run_sasa_device(h, n);
    mark_dirty_full(h);
    h->counts.resize(n);
    hipchk(hipMemcpyAsync(h->counts.data(), h->d_counts.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream),
           "download counts");
    hipchk(hipStreamSynchronize(h->stream), "sync sasa");
    h->rsasa.resize(n);
    for (int i = 0; i < n; ++i) h->rsasa[i] = (float)h->counts[i] / (float)h->n_dots;
    h->have_rsasa = true;
}

// Local recount after bodies [start, start+len) moved (old positions in old_bodies),
// enqueued: only an atom within reach (2 max(r + p) + 0.02) of a moved atom's old or new
// position can change its count, so the atoms inside the moved set's bounding box widened by
// that reach are recounted on the device against a current cell list, and the counts that
// changed come back in one transfer (sasa_local_finish waits for it).  Every other count is
// unchanged by construction.  The cell list is the slab (moved atoms moved in it by one
// kernel; rebuilt from the device atoms every SLAB_REBUILD recounts and whenever those were
// written otherwise) or, without it, the hashed cell list rebuilt over all atoms.  Returns
// false (nothing enqueued) without current counts to compare with.
bool sasa_local_enqueue(rsx_engine* h, int start, int len, const std::vector<float>& old_bodies)
{
    const int n = h->n;
    if (len <= 0 || (int)h->counts.size() != n || !h->have_rsasa) return false;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int a = 0; a < len; ++a) {
        for (int d = 0; d < 3; ++d) {
            const float vo = old_bodies[4 * (size_t)a + d], vn = h->bodies[4 * (size_t)(start + a) + d];
            lo[d] = std::min(lo[d], std::min(vo, vn));
            hi[d] = std::max(hi[d], std::max(vo, vn));
        }
    }
    float rmax = 0.0f;
    for (float v : h->rext) rmax = std::max(rmax, v);
    const float reach = 2.0f * rmax + 0.02f;
    const size_t stage_bytes = rsx::LOCAL_STAGE_HEADER + 16 * (size_t)LOCAL_PATCH_CAP;
    h->pin_lstage.ensure(stage_bytes);
    rsx::LocalStage* st = reinterpret_cast<rsx::LocalStage*>(h->pin_lstage.p);
    for (int d = 0; d < 3; ++d) {
        st->lo[d] = lo[d] - reach;
        st->hi[d] = hi[d] + reach;
    }
    st->start = start;
    st->len = len;
    st->parity = h->l_parity;
    h->l_used_parity = h->l_parity;
    h->l_parity ^= 1;
    h->d_sasa_atoms.ensure(n);
    const size_t local_ints = rsx::LOCAL_HDR_TOTAL + 2 * (size_t)LOCAL_CHANGED_CAP;
    const size_t local_bytes = local_ints * sizeof(int);
    h->pin_local.ensure(local_ints);
    if (h->sasa_mirror && len <= LOCAL_PATCH_CAP) {
        // one replay of a captured graph: stage upload, the cell-list work (on the slab: one
        // kernel), the listed counts, the read-back
        std::memcpy(h->pin_lstage.p + rsx::LOCAL_STAGE_HEADER, &h->bodies[4 * (size_t)start], (size_t)len * 16);
        const bool slab = !h->slab_off && n <= SLAB_MAX_ATOMS;
        const rsx::SasaArgs a = sasa_args(h, n, LOCAL_PATCH_CAP, slab);
        if (slab && (h->slab_age < 0 || h->slab_age >= SLAB_REBUILD || h->slab_built[0] != a.slab ||
                     h->slab_built[1] != a.slab_count || h->slab_built[2] != a.slab_pos ||
                     h->slab_built_nb != a.slab_buckets || h->slab_built_n != n || h->slab_built_cell != a.cell)) {
            // from the device atoms as they are before this move (the graph patches them)
            hipchk(rsx::launch_slab_build(a, h->stream), "slab build");
            h->slab_age = 0;
            h->slab_built[0] = a.slab;
            h->slab_built[1] = a.slab_count;
            h->slab_built[2] = a.slab_pos;
            h->slab_built_nb = a.slab_buckets;
            h->slab_built_n = n;
            h->slab_built_cell = a.cell;
        }
        const bool same = h->l_exec && std::memcmp(&a, &h->l_args, sizeof a) == 0 &&
                          h->l_pin_stage == h->pin_lstage.p && h->l_pin_local == h->pin_local.p;
        if (!same) {
            release_local_graph(h);
            hipchk(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                hipchk(hipMemcpyAsync(h->d_lstage.p, h->pin_lstage.p, stage_bytes, hipMemcpyHostToDevice, h->stream),
                       "upload stage");
                hipchk(rsx::launch_sasa(a, h->stream), "sasa kernels");
                hipchk(hipMemcpyAsync(h->pin_local.p, h->d_local.p, local_bytes, hipMemcpyDeviceToHost, h->stream),
                       "download changed counts");
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(h->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hipchk(hipStreamEndCapture(h->stream, &h->l_graph), "end capture");
            hipchk(hipGraphInstantiate(&h->l_exec, h->l_graph, nullptr, nullptr, 0), "instantiate graph");
            h->l_args = a;
            h->l_pin_stage = h->pin_lstage.p;
            h->l_pin_local = h->pin_local.p;
        }
        hipchk(hipGraphLaunch(h->l_exec, h->stream), "graph launch");
        if (slab) h->slab_age++;
    } else {
        h->slab_age = -1;  // the device atoms are written here without the slab
        if (h->sasa_mirror) {
            h->pin_moved.ensure(len);
            std::memcpy(h->pin_moved.p, &h->bodies[4 * (size_t)start], (size_t)len * 16);
            hipchk(hipMemcpyAsync(h->d_sasa_atoms.p + start, h->pin_moved.p, (size_t)len * 16, hipMemcpyHostToDevice,
                                  h->stream),
                   "upload moved sasa atoms");
        } else {
            // the device copies hold other data: the host bodies and counts go up whole
            h->d_counts.ensure(n);
            hipchk(hipMemcpyAsync(h->d_sasa_atoms.p, h->bodies.data(), (size_t)n * 16, hipMemcpyHostToDevice,
                                  h->stream),
                   "upload sasa atoms");
            hipchk(hipMemcpyAsync(h->d_counts.p, h->counts.data(), (size_t)n * 4, hipMemcpyHostToDevice, h->stream),
                   "upload counts");
            hipchk(hipStreamSynchronize(h->stream), "sync sasa upload");  // pageable sources
        }
        const rsx::SasaArgs a = sasa_args(h, n, 0);
        hipchk(hipMemcpyAsync(h->d_lstage.p, h->pin_lstage.p, rsx::LOCAL_STAGE_HEADER, hipMemcpyHostToDevice,
                              h->stream),
               "upload stage");
        hipchk(rsx::launch_sasa(a, h->stream), "sasa kernels");
        hipchk(hipMemcpyAsync(h->pin_local.p, h->d_local.p, local_bytes, hipMemcpyDeviceToHost, h->stream),
               "download changed counts");
    }
    h->sasa_mirror = true;
    hipchk(hipEventRecord(h->ev_sasa, h->stream), "record sasa");
    return true;
}

// Waits for the local recount and applies it: changed receives the atoms whose count
// changed, ascending.  Returns false, with nothing applied, when a slab bucket overflowed
// (that recount missed atoms: the caller counts every atom instead).
bool sasa_local_finish(rsx_engine* h, std::vector<int>& changed)
{
    hipchk(hipEventSynchronize(h->ev_sasa), "sync sasa");
    changed.clear();
    const int* hdr = h->pin_local.p + rsx::LOCAL_HDR_INTS * h->l_used_parity;
    const int* pairs = h->pin_local.p + rsx::LOCAL_HDR_TOTAL;
    if (hdr[2] != 0) {
        // overflow in the first recount after a build: these atoms do not fit the slab
        h->slab_off = h->slab_age <= 1;
        h->slab_age = -1;
        return false;
    }
    const int n_changed = hdr[1];
    auto apply = [&](int i, uint32_t c) {
        changed.push_back(i);
        mark_dirty(h, i);
        h->counts[i] = c;
        h->rsasa[i] = (float)c / (float)h->n_dots;
    };
    if (n_changed <= LOCAL_CHANGED_CAP) {
        std::vector<std::pair<int, uint32_t>> e((size_t)n_changed);
        for (int t = 0; t < n_changed; ++t) e[t] = {pairs[2 * t], (uint32_t)pairs[2 * t + 1]};
        std::sort(e.begin(), e.end());
        for (const auto& x : e) apply(x.first, x.second);
        return true;
    }
    // more than the block holds: every count compared
    std::vector<uint32_t> c((size_t)h->n);
    hipchk(hipMemcpyAsync(c.data(), h->d_counts.p, (size_t)h->n * 4, hipMemcpyDeviceToHost, h->stream),
           "download counts");
    hipchk(hipStreamSynchronize(h->stream), "sync counts");
    for (int i = 0; i < h->n; ++i)
        if (c[i] != h->counts[i]) apply(i, c[i]);
    return true;
}

// ---- pair jobs ---------------------------------------------------------------
// Column range L of the pair jobs of a launch on a layout of the given path (uniform
// grid?, variant, K, chunks) and bounding radius.
int pick_range_for(const rsx_engine* h, double column_atoms_total, bool uniform, int variant, int K, int n_chunks,
                   double radius, double rounds)
{
    // resident waves of the pair kernel (the runtime's occupancy for the FP32 kernels,
    // two per SIMD otherwise)
    int per_cu = 0;
    if (uniform && (variant == rsx::VARIANT_FP32 || variant == rsx::VARIANT_FP32R))
        per_cu = rsx::rec32_blocks_per_cu(variant, K, rsx::sin_series_terms(h->dq * radius), false);
    if (per_cu <= 0) per_cu = 8;
    const double slots = (double)std::max(1, h->n_cu) * per_cu;
    // about `rounds` rounds of waves over the resident slots (default six): fewer leave a
    // partly filled last round unless the tail is split (measured on incremental updates
    // without tail splitting: L = 64-128 beats 512 by 1.3-2.2x), more pay the per-job
    // set-up and epilogue more often.  64..1024 columns, rounded down.
    const double jobs = std::max(1.0, rounds * slots / std::max(1, n_chunks));
    double L = column_atoms_total / 64.0 / jobs;
    L = std::max(64.0, std::min(1024.0, L));
    return std::max(64, (int)L / 64 * 64);
}

// the same on the single-structure layout (the caller's column range when set)
int pick_range(const rsx_engine* h, double column_atoms_total, int variant, int K, int n_chunks, double rounds = 6.0)
{
    if (h->range_req > 0) return h->range_req;
    return pick_range_for(h, column_atoms_total, h->uniform, variant, K, n_chunks, h->path_radius, rounds);
}

int pick_range(const rsx_engine* h, double column_atoms_total)
{
    return pick_range(h, column_atoms_total, h->variant, h->K, h->n_chunks);
}

// Kernel of the incremental launches: the forward's.  With hydration classes a 64-row tile
// of changed atoms holds tens of classes, so each flush loads the rows' factors from L2;
// that costs little while the columns keep long class runs (relocate_classes keeps them
// sorted): at 50k atoms a hydrated 500-atom move measured the same with it as with the
// per-lane-sum kernel at K = 64, and its FP64 row sums are the more accurate at deep minima
// of I(q).
void incremental_kernel(rsx_engine* h, int& variant, int& K, int& n_chunks, const ChunkConst*& chunks)
{
    variant = h->variant;
    K = h->K;
    n_chunks = h->n_chunks;
    chunks = h->d_chunks.p;
}

// triangle jobs over records [0, m) of cols (rows == cols); appends to jobs
void triangle_jobs(std::vector<PairJob>& jobs, const float4* cols, int m, int L, float scale)
{
    const int T = (m + 63) / 64;
    for (int I = 0; I < T; ++I) {
        const int r0 = I * 64;
        for (int c0 = r0; c0 < m; c0 += L) {
            PairJob j;
            j.rows = cols + r0;
            j.nrow = std::min(64, m - r0);
            j.cols = cols + c0;
            j.ncol = std::min(L, m - c0);
            j.diag = c0 == r0 ? 1 : 0;
            j.scale = scale;
            j.slot = 0;
            j.pad = 0;
            jobs.push_back(j);
        }
    }
}

// rows [0, nr) against cols [0, nc) (full rectangle)
void rect_jobs(std::vector<PairJob>& jobs, const float4* rows, int nr, const float4* cols, int nc, int L, float scale)
{
    jobs.reserve(jobs.size() + (size_t)((nr + 63) / 64) * ((nc + L - 1) / L));
    for (int r0 = 0; r0 < nr; r0 += 64)
        for (int c0 = 0; c0 < nc; c0 += L) {
            PairJob j;
            j.rows = rows + r0;
            j.nrow = std::min(64, nr - r0);
            j.cols = cols + c0;
            j.ncol = std::min(L, nc - c0);
            j.diag = 0;
            j.scale = scale;
            j.slot = 0;
            j.pad = 0;
            jobs.push_back(j);
        }
}

// order by cost (longest first, ties in build order) and number the slots in execution
// order.  The tables hold few distinct job sizes (full ranges, ragged ends, the M x M
// blocks), so a bucket pass over the sizes replaces a comparison sort.
void finalize_jobs(std::vector<PairJob>& jobs)
{
    std::vector<long> sizes;
    for (const PairJob& j : jobs) {
        const long c = (long)j.ncol * j.nrow;
        if (std::find(sizes.begin(), sizes.end(), c) == sizes.end()) sizes.push_back(c);
        if (sizes.size() > 64) break;
    }
    if (sizes.size() > 64) {
        std::stable_sort(jobs.begin(), jobs.end(), [](const PairJob& a, const PairJob& b) {
            return (long)a.ncol * a.nrow > (long)b.ncol * b.nrow;
        });
    } else {
        std::sort(sizes.begin(), sizes.end(), std::greater<long>());
        std::vector<PairJob> out;
        out.reserve(jobs.size());
        for (long c : sizes)
            for (const PairJob& j : jobs)
                if ((long)j.ncol * j.nrow == c) out.push_back(j);
        jobs.swap(out);
    }
    for (size_t s = 0; s < jobs.size(); ++s) jobs[s].slot = (int)s;
}

rsx::DebyeArgs debye_args(rsx_engine* h, const PairJob* d_jobs, double* partials)
{
    rsx::DebyeArgs a;
    a.jobs = d_jobs;
    a.n_chunks = h->n_chunks;
    a.ftab = h->d_ftab.p;
    a.ldf = h->ldf;
    a.zero_class = h->n_class;
    a.chunks = h->d_chunks.p;
    a.half_dq = 0.5 * h->dq;
    a.dq = (float)h->dq;
    a.invq = h->d_invq.p;
    a.qrev = h->d_qrev.p;
    a.qsplit = h->d_qsplit.p;
    a.phase32 = phase32_ok(h, h->path_radius) ? 1 : 0;
    a.qv = h->d_qv.p;
    a.partials = partials;
    a.variant = h->variant;
    a.queue = nullptr;
    a.persistent_waves = 0;
    // pairs satisfy r <= 2 R about the centre of the layout, so dq r / 2 <= dq R
    a.sin_terms = rsx::sin_series_terms(h->dq * h->path_radius);
    return a;
}

// upload jobs, launch the pair kernel and reduce into iq (single output)
struct KernelChoice {
    int variant, K, n_chunks;
    const ChunkConst* chunks;
};
void launch_jobs(rsx_engine* h, const PairJob* d_jobs, int nj, double* iq, int accumulate, double q0val,
                 bool sync = true, const KernelChoice* kc = nullptr);

void run_jobs(rsx_engine* h, DBuf<PairJob>& jbuf, const std::vector<PairJob>& jobs, double* iq, int accumulate,
              double q0val)
{
    const int nj = (int)jobs.size();
    jbuf.ensure(nj);
    if (nj > 0)
        hipchk(hipMemcpyAsync(jbuf.p, jobs.data(), nj * sizeof(PairJob), hipMemcpyHostToDevice, h->stream),
               "upload jobs");
    launch_jobs(h, jbuf.p, nj, iq, accumulate, q0val);
}

// launch nj jobs already on the device and reduce into iq (single output); without sync
// the work stays queued on the engine stream
void launch_jobs(rsx_engine* h, const PairJob* d_jobs, int nj, double* iq, int accumulate, double q0val, bool sync,
                 const KernelChoice* kc)
{
    h->d_partials.ensure((size_t)std::max(nj, 1) * h->ldf);
    h->d_red.ensure(std::max<size_t>(1, rsx::reduce_scratch_doubles(h->nq, 1, nj)));
    h->d_slot_begin.ensure(2);
    h->up_sb[0] = 0;
    h->up_sb[1] = nj;
    h->up_q0 = q0val;
    hipchk(hipMemcpyAsync(h->d_slot_begin.p, h->up_sb, sizeof(h->up_sb), hipMemcpyHostToDevice, h->stream),
           "upload slots");
    hipchk(hipMemcpyAsync(h->d_q0.p, &h->up_q0, sizeof(double), hipMemcpyHostToDevice, h->stream), "upload q0");
    rsx::DebyeArgs a = debye_args(h, d_jobs, h->d_partials.p);
    int K = h->K;
    if (kc) {
        a.variant = kc->variant;
        a.n_chunks = kc->n_chunks;
        a.chunks = kc->chunks;
        K = kc->K;
    }
    hipchk(rsx::launch_debye(a, nj, K, h->uniform, h->stream), "debye kernel");
    if (nj > 0) {
        hipchk(rsx::launch_reduce(h->d_partials.p, h->ldf, h->nq, h->d_slot_begin.p, 1, iq, h->nq, accumulate,
                                  h->d_q0.p, h->q0_index, nj, h->d_red.p, h->stream),
               "reduce kernel");
    } else if (!accumulate) {
        std::vector<double> z(h->nq, 0.0);
        if (h->q0_index >= 0) z[h->q0_index] = q0val;
        hipchk(hipMemcpyAsync(iq, z.data(), h->nq * sizeof(double), hipMemcpyHostToDevice, h->stream), "zero iq");
        hipchk(hipStreamSynchronize(h->stream), "sync zero iq");  // z is a local buffer
    }
    if (sync) hipchk(hipStreamSynchronize(h->stream), "sync pair kernels");
}

// Sort jobs by cost (most expensive first) and halve the cheapest splittable ones (at
// least 128 columns, not diagonal), `round` of them (0: none), twice, re-sorting after
// each pass: the final rounds of waves then consist of ever shorter jobs.
template <typename Cost>
void split_tail(std::vector<PairJob>& jobs, long round, Cost&& cost)
{
    auto by_cost = [&]() {
        std::vector<std::pair<double, int>> key(jobs.size());
        for (size_t k = 0; k < jobs.size(); ++k) key[k] = {-cost(jobs[k]), (int)k};
        std::stable_sort(key.begin(), key.end());
        std::vector<PairJob> out(jobs.size());
        for (size_t k = 0; k < jobs.size(); ++k) out[k] = jobs[key[k].second];
        jobs.swap(out);
    };
    by_cost();
    for (int level = 0; level < 2 && round > 0 && (long)jobs.size() > round; ++level) {
        long done = 0;
        const size_t m = jobs.size();
        for (size_t k = m; k-- > 0 && done < round;) {
            PairJob& j = jobs[k];
            if (j.diag || j.ncol < 128) continue;
            PairJob b = j;
            const int first = j.ncol / 2;
            j.ncol = first;
            b.cols = j.cols + first;
            b.ncol -= first;
            jobs.push_back(b);
            ++done;
        }
        by_cost();
    }
    for (size_t k = 0; k < jobs.size(); ++k) jobs[k].slot = (int)k;
}

// waves of the pair kernel resident at once, in jobs (the runtime's occupancy for the FP32
// kernels, eight one-wave blocks per CU otherwise), over the q chunks
long resident_jobs(const rsx_engine* h, int variant, int K, int n_chunks)
{
    int per_cu = 0;
    if (h->uniform && (variant == rsx::VARIANT_FP32 || variant == rsx::VARIANT_FP32R))
        per_cu = rsx::rec32_blocks_per_cu(variant, K, rsx::sin_series_terms(h->dq * h->path_radius), false);
    if (per_cu <= 0) per_cu = 8;
    return std::max<long>(1, (long)std::max(1, h->n_cu) * per_cu / std::max(1, n_chunks));
}

// Order of the forward (triangle) jobs, and a finer tail.  A job costs its columns plus
// the flushes of the class runs among them (in column equivalents; a run flush of the
// FP32 run-flush kernel is ~1.4 columns of work, of the other kernels less), and jobs
// run most expensive first.  Waves are handed out in launch order as resident slots
// free up, so with jobs of near-equal cost the last round is quantised: a partly
// filled round still takes a whole job's time.  The cheapest splittable jobs are
// therefore halved, twice, so the final rounds are made of ever shorter waves and the
// idle tail shrinks to about a quarter of a job.  (Incremental launches keep the plain
// longest-first order: their column range is already sized for about six rounds, and
// splitting measured slower there.)
void order_jobs(rsx_engine* h, std::vector<PairJob>& jobs)
{
    const int n = h->n;
    const bool have_cls = (int)h->caller_of.size() == n && (int)h->cls.size() == n;
    std::vector<int> P((size_t)n + 1, 0);  // P[s]: class changes at slots t < s (t >= 1)
    for (int t = 0; t < n; ++t)
        P[t + 1] = P[t] + (have_cls && t > 0 && h->cls[h->caller_of[t]] != h->cls[h->caller_of[t - 1]] ? 1 : 0);
    const double flush_w = h->variant == rsx::VARIANT_FP32R ? 1.4 : 0.6;
    const float4* base = h->d_cols.p;
    auto cost = [&](const PairJob& j) {
        const long c0 = (long)(j.cols - base), c1 = c0 + j.ncol;
        const double runs = 1.0 + (c1 - 1 > c0 ? P[c1] - P[c0 + 1] : 0);
        return ((double)j.ncol + flush_w * runs) * j.nrow;
    };
    const long round = resident_jobs(h, h->variant, h->K, h->n_chunks);
    split_tail(jobs, (long)jobs.size() > 2 * round ? round : 0, cost);
}

rsx_engine::FwdKey forward_key(const rsx_engine* h, int L)
{
    rsx_engine::FwdKey k;
    k.n = h->n;
    k.L = L;
    k.variant = h->variant;
    k.K = h->K;
    k.n_chunks = h->n_chunks;
    k.uniform = h->uniform ? 1 : 0;
    k.cols = h->d_cols.p;
    return k;
}

std::vector<PairJob> forward_jobs(rsx_engine* h)
{
    const int n = h->n;
    std::vector<PairJob> jobs;
    const int L = pick_range(h, (double)n * n / 2.0);
    h->last_L = L;
    triangle_jobs(jobs, h->d_cols.p, n, L, 2.0f);
    order_jobs(h, jobs);
    h->fjobs_key = forward_key(h, L);
    return jobs;
}

// The slot range {0, n jobs} the device-resident forward reduces over: written with every
// forward job table that reaches d_fjobs (full_forward, retune), so it always names that table.
void set_forward_slots(rsx_engine* h, int nj)
{
    int sbf[2] = {0, nj};
    h->d_slot_begin_f.ensure(2);
    hipchk(hipMemcpy(h->d_slot_begin_f.p, sbf, sizeof(sbf), hipMemcpyHostToDevice), "upload forward slots");
}

// The full forward into d_iq.  The job table is built (host triangle tiling, cost order,
// upload) only when the layout's key changed: a re-baseline of the incremental path repeats
// the forward on the same columns, whose table is still on the device.
void full_forward(rsx_engine* h)
{
    const double sf = sum_f0(h);
    const int L = pick_range(h, (double)h->n * h->n / 2.0);
    if (h->last_jobs.empty() || !(h->fjobs_key == forward_key(h, L))) {
        const std::vector<PairJob> jobs = forward_jobs(h);
        run_jobs(h, h->d_fjobs, jobs, h->d_iq.p, 0, sf * sf);
        h->last_jobs = jobs;
        set_forward_slots(h, (int)jobs.size());
    } else {
        h->last_L = L;
        launch_jobs(h, h->d_fjobs.p, (int)h->last_jobs.size(), h->d_iq.p, 0, sf * sf);
    }
    h->last_q0 = sf * sf;
    h->since_rebase = 0;
}

// I(q) to the host in two steps: the download is queued on the engine stream (work queued
// after it may still run when the host has the result), then the host waits for it alone
void enqueue_iq_download(rsx_engine* h)
{
    h->pin_iq.ensure(h->nq);
    hipchk(hipMemcpyAsync(h->pin_iq.p, h->d_iq.p, h->nq * sizeof(double), hipMemcpyDeviceToHost, h->stream),
           "download iq");
    hipchk(hipEventRecord(h->ev_out, h->stream), "record download iq");
}

void finish_iq_download(rsx_engine* h, float* out32, double* out64)
{
    hipchk(hipEventSynchronize(h->ev_out), "sync download iq");
    const double* tmp = h->pin_iq.p;
    for (int k = 0; k < h->nq; ++k) {
        if (out32) out32[k] = (float)tmp[k];
        if (out64) out64[k] = tmp[k];
    }
}

void copy_out(rsx_engine* h, float* out32, double* out64)
{
    enqueue_iq_download(h);
    finish_iq_download(h, out32, out64);
}

void check_args_set(rsx_engine* h)
{
    if (h->state < 3) fail(RSX_ARGS_NOT_SET, "engine arguments not set (rsx_initialize)");
}

void load_bodies(rsx_engine* h, const float* bodies, int start, int len)
{
    for (int i = start; i < start + len; ++i) {
        uint32_t t;
        std::memcpy(&t, &bodies[4 * i + 3], 4);
        if ((int)t >= h->n_types) fail(RSX_INVALID_ARG, "body type index outside the factor table");
        mark_dirty(h, i);
        std::memcpy(&h->bodies[4 * i], &bodies[4 * i], 16);
        h->type[i] = t;
    }
}

void calc_curve_impl(rsx_engine* h, const float* bodies, int n, float* out32, double* out64)
{
    check_args_set(h);
    if (n != h->n) fail(RSX_INVALID_ARG, "different number of bodies expected in calc_curve()");
    if (!bodies) fail(RSX_INVALID_ARG, "null bodies");
    load_bodies(h, bodies, 0, n);
    set_device(h);
    h->radius = bounding_radius(h->bodies.data(), n, h->centre);
    select_path(h, h->radius);
    if (h->params_dirty || h->d_ftab.p == nullptr) build_class_table(h);
    if (h->hydrate) compute_sasa(h);
    assign_slots(h);
    full_forward(h);
    h->params_dirty = false;
    h->w_dirty = false;
    h->host_stale = false;
    h->state = 4;
    copy_out(h, out32, out64);
}

// Keep the columns sorted by class through incremental updates: atom M[a] changed from
// class cls_old[a] to cls_new[a] while keeping its slot, which sits in the run of its old
// class; it is moved into the run of its new class by swaps across the runs in between,
// each swap exchanging the atom with the boundary slot of a run and shifting that boundary
// by one (so a count change of d dots costs d swaps).  run_start holds the old classes'
// run starts and is updated in place.  Appends to touched every slot whose atom changed and
// the final slots of all of M (whose records changed).
void relocate_classes(rsx_engine* h, const std::vector<int>& M, const std::vector<uint32_t>& cls_old,
                      const std::vector<uint32_t>& cls_new, std::vector<int>& run_start, std::vector<int>& touched)
{
    std::vector<char>& mark = h->slot_mark;
    mark.resize(h->n, 0);
    auto touch = [&](int sl) {
        if (!mark[sl]) {
            mark[sl] = 1;
            touched.push_back(sl);
        }
    };
    auto swap_slots = [&](int s1, int s2) {
        if (s1 == s2) return;
        const int i1 = h->caller_of[s1], i2 = h->caller_of[s2];
        h->caller_of[s1] = i2;
        h->caller_of[s2] = i1;
        h->slot_of[i1] = s2;
        h->slot_of[i2] = s1;
        touch(s1);
        touch(s2);
        h->slots_dirty = true;
    };
    for (size_t a = 0; a < M.size(); ++a) {
        const int i = M[a];
        int c = (int)cls_old[a];
        const int b = (int)cls_new[a];
        while (c < b) {  // to the last slot of run c, then across the boundary into run c + 1
            swap_slots(h->slot_of[i], run_start[c + 1] - 1);
            run_start[c + 1]--;
            ++c;
        }
        while (c > b) {  // to the first slot of run c, then across the boundary into run c - 1
            swap_slots(h->slot_of[i], run_start[c]);
            run_start[c]++;
            --c;
        }
    }
    for (int i : M) touch(h->slot_of[i]);
    for (int sl : touched) mark[sl] = 0;
}

// byte offsets of arrays packed into one staging block, each 16-byte aligned
struct Pack {
    size_t size = 0;
    size_t add(size_t bytes)
    {
        const size_t off = size;
        size += (bytes + 15) / 16 * 16;
        return off;
    }
};

// The staging block made to hold `bytes`; returns its host side to fill.  The device side
// grows to powers of two from 64 KiB, so its address rarely changes; before it is
// replaced, both streams are drained (work queued after an update's result, such as the
// relocation scatter, may still read the old block).
unsigned char* stage_begin(rsx_engine* h, rsx_engine::Staging& st, size_t bytes)
{
    st.host.ensure(bytes);
    if (bytes > st.dev.cap) {
        hipchk(hipStreamSynchronize(h->stream), "sync before staging growth");
        hipchk(hipStreamSynchronize(h->stream2), "sync before staging growth");
        size_t c = 65536;
        while (c < bytes) c <<= 1;
        st.dev.ensure(c);
    }
    return st.host.p;
}

void stage_upload(rsx_engine::Staging& st, size_t bytes, hipStream_t stream, const char* what)
{
    hipchk(hipMemcpyAsync(st.dev.p, st.host.p, bytes, hipMemcpyHostToDevice, stream), what);
}

// Job table of one launch of an update (part 0: old records, 1: new records): the nrows
// rows (on the device) against the columns with weight 2 w and against themselves with
// weight w.  Tables depend only on the shape and the buffers and are kept for the shapes
// that recur (the least recently used one is replaced).
const rsx_engine::IncTable* update_jobs(rsx_engine* h, int part, const float4* rows, int nrows, float w,
                                        const KernelChoice& kc, hipStream_t stream)
{
    // about two rounds of waves, the tail split (measured against six rounds unsplit:
    // 0.834 -> 0.824 ms plain, 2.23 -> 2.17 ms hydrated per 500-atom move at 50k atoms)
    const int L = pick_range(h, (double)nrows * (double)h->n, kc.variant, kc.K, kc.n_chunks, 2.0);
    rsx_engine::IncKey key;
    key.part = part;
    key.nrows = nrows;
    key.n = h->n;
    key.L = L;
    key.K = kc.K;
    key.n_chunks = kc.n_chunks;
    key.variant = kc.variant;
    key.uniform = h->uniform ? 1 : 0;
    key.rows = rows;
    key.cols = h->d_cols.p;
    rsx_engine::IncTable* tab = nullptr;
    for (auto& t : h->inc_tab)
        if (t.njobs > 0 && t.key == key) tab = &t;
    if (!tab) {
        tab = &h->inc_tab[0];
        for (auto& t : h->inc_tab)
            if (t.used < tab->used) tab = &t;
        std::vector<PairJob>& jobs = h->inc_jobs;
        jobs.clear();
        rect_jobs(jobs, rows, nrows, h->d_cols.p, h->n, L, 2.0f * w);
        rect_jobs(jobs, rows, nrows, rows, nrows, L, w);
        split_tail(jobs, resident_jobs(h, kc.variant, kc.K, kc.n_chunks),
                   [](const PairJob& j) { return (double)j.ncol * j.nrow; });
        tab->jobs.ensure(jobs.size());
        // staged through pinned memory, one buffer per part (the call waits for the
        // update's result before returning, so the buffer is free again at the next one)
        HBuf<PairJob>& pin = h->pin_jobs[part];
        pin.ensure(jobs.size());
        std::memcpy(pin.p, jobs.data(), jobs.size() * sizeof(PairJob));
        hipchk(hipMemcpyAsync(tab->jobs.p, pin.p, jobs.size() * sizeof(PairJob), hipMemcpyHostToDevice, stream),
               "upload jobs");
        tab->key = key;
        tab->njobs = (int)jobs.size();
    }
    tab->used = ++h->inc_clock;
    return tab;
}

// the pair kernel over nj jobs, partial sums in rows 0..nj-1 of partials
void launch_update_pairs(rsx_engine* h, const PairJob* d_jobs, int nj, double* partials, const KernelChoice& kc,
                         hipStream_t stream)
{
    rsx::DebyeArgs a = debye_args(h, d_jobs, partials);
    a.variant = kc.variant;
    a.n_chunks = kc.n_chunks;
    a.chunks = kc.chunks;
    hipchk(rsx::launch_debye(a, nj, kc.K, h->uniform, stream), "debye kernel");
}

// the sum of those partials added to d_iq (slot range and q = 0 value read from the
// device), on the engine stream (the two launches of an update add in a fixed order)
void reduce_update_part(rsx_engine* h, int nj, const double* partials, const int* d_sb, const double* d_q0)
{
    // the largest reduction scratch of one output: both launches of an update share it
    h->d_red.ensure(std::max<size_t>(1, rsx::reduce_scratch_doubles(h->nq, 1, 1 << 24)));
    hipchk(rsx::launch_reduce(partials, h->ldf, h->nq, d_sb, 1, h->d_iq.p, h->nq, 1, d_q0, h->q0_index, nj,
                              h->d_red.p, h->stream),
           "reduce kernel");
}

// The old-records launch of an update: S = [start, start + len) removed from the columns
// (skip records), its old records (old_bodies, old classes) sorted by class as rows, with
// weight -1 (nothing is read from the host afterwards, so the work stays queued).  Staged
// in two steps: the rows and job table uploaded and S skipped (ev_skip recorded), then the
// launch and its reduction.
struct OldPart {
    const rsx_engine::IncTable* tab = nullptr;
    const int* d_sb = nullptr;
    const double* d_q0 = nullptr;
};

OldPart stage_old_part(rsx_engine* h, int start, int len, const std::vector<float>& old_bodies, const KernelChoice& kc)
{
    OldPart op;
    if (len <= 0) return op;
    const int mp = (len + 63) / 64 * 64;
    std::vector<std::pair<uint32_t, int>> order((size_t)len);
    for (int a = 0; a < len; ++a) order[a] = {h->cls[start + a], a};
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<uint32_t, int>& x, const std::pair<uint32_t, int>& y) { return x.first < y.first; });
    Pack pk;
    const size_t off_rows = pk.add((size_t)mp * sizeof(float4));
    const size_t off_idx = pk.add((size_t)len * sizeof(int));
    const size_t off_sb = pk.add(2 * sizeof(int));
    const size_t off_q0 = pk.add(sizeof(double));
    unsigned char* hb = stage_begin(h, h->st_old, pk.size);
    float4* rows = reinterpret_cast<float4*>(hb + off_rows);
    for (int t = 0; t < len; ++t) {
        const float* ob = &old_bodies[4 * (size_t)order[t].second];
        rows[t] = make_float4(ob[0], ob[1], ob[2], bitsf(order[t].first));
    }
    for (int t = len; t < mp; ++t) rows[t] = make_float4(0.f, 0.f, 0.f, bitsf(CLS_SKIP));
    int* idx = reinterpret_cast<int*>(hb + off_idx);
    for (int a = 0; a < len; ++a) idx[a] = h->slot_of[start + a];
    unsigned char* db = h->st_old.dev.p;
    const rsx_engine::IncTable* tab =
        update_jobs(h, 0, reinterpret_cast<const float4*>(db + off_rows), mp, -1.0f, kc, h->stream);
    reinterpret_cast<int*>(hb + off_sb)[0] = 0;
    reinterpret_cast<int*>(hb + off_sb)[1] = tab->njobs;
    *reinterpret_cast<double*>(hb + off_q0) = 0.0;
    stage_upload(h->st_old, pk.size, h->stream, "upload old rows");
    hipchk(rsx::launch_scatter_value(h->d_cols.p, reinterpret_cast<const int*>(db + off_idx),
                                     make_float4(0.f, 0.f, 0.f, bitsf(CLS_SKIP)), len, h->stream),
           "scatter skip");
    hipchk(hipEventRecord(h->ev_skip, h->stream), "record skip");
    op.tab = tab;
    op.d_sb = reinterpret_cast<const int*>(db + off_sb);
    op.d_q0 = reinterpret_cast<const double*>(db + off_q0);
    return op;
}

void launch_old_part(rsx_engine* h, const OldPart& op, const KernelChoice& kc)
{
    if (!op.tab) return;
    h->d_partials.ensure((size_t)op.tab->njobs * h->ldf);
    launch_update_pairs(h, op.tab->jobs.p, op.tab->njobs, h->d_partials.p, kc, h->stream);
    reduce_update_part(h, op.tab->njobs, h->d_partials.p, op.d_sb, op.d_q0);
}

void incremental_update(rsx_engine* h, int start, int len, const std::vector<float>& old_bodies, bool sasa_now,
                        float* out32, double* out64);

void recalc_curve_impl(rsx_engine* h, const float* bodies, int n, int start, int len, float* out32, double* out64)
{
    if (h->state != 4) fail(RSX_NO_INITIAL_CALC, "recalc_curve() called before an initial call to calc_curve()");
    if (start < 0 || len < 0 || (long)start + len > n)
        fail(RSX_INVALID_ARG, "bodies vector too small in recalc_curve()");
    if ((long)start + len > h->n) fail(RSX_INVALID_ARG, "update region exceeds the number of bodies in recalc_curve()");
    set_device(h);
    const std::vector<float> old_bodies(h->bodies.begin() + 4 * (size_t)start,
                                        h->bodies.begin() + 4 * (size_t)(start + len));
    load_bodies(h, bodies, start, len);
    for (int i = start; i < start + len; ++i) {
        double s2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double v = h->bodies[4 * (size_t)i + d] - h->centre[d];
            s2 += v * v;
        }
        h->radius = std::max(h->radius, std::sqrt(s2) * (1.0 + 1e-6) + 1e-6);
    }
    if (select_path(h, h->radius)) h->params_dirty = true;

    // hydration update (saxs_alg_cl_gpu_pt_wf.cpp:333-401)
    const bool hyd_before = h->hydrate;
    bool sasa_now = false;
    if (hydration_on(h)) {
        if (len > 0)
            sasa_now = true;
        else
            sasa_now = h->w_dirty && !h->have_rsasa;
        if (h->no_rsasa_recalc && h->have_rsasa) sasa_now = false;
    }

    if (h->params_dirty || hyd_before != hydration_on(h) || h->host_stale) {
        // new parameters (or device state left by a device-resident forward) force a full
        // recalculation
        if (h->params_dirty || hyd_before != hydration_on(h)) build_class_table(h);
        if (h->hydrate && (sasa_now || !h->have_rsasa)) compute_sasa(h);
        assign_slots(h);
        full_forward(h);
        h->params_dirty = false;
        h->w_dirty = false;
        h->host_stale = false;
        copy_out(h, out32, out64);
        return;
    }

    // moved atoms' types against the SASA radius table before any device work
    if (sasa_now)
        for (int i = start; i < start + len; ++i)
            if (h->type[i] >= h->radii.size()) fail(RSX_INVALID_ARG, "body type index outside the SASA radius table");
    try {
        incremental_update(h, start, len, old_bodies, sasa_now, out32, out64);
    } catch (...) {
        // a failure part way leaves the columns and the running I(q) part updated: the
        // next call lays the host bodies out again
        h->host_stale = true;
        h->sasa_mirror = false;
        (void)hipStreamSynchronize(h->stream2);
        (void)hipStreamSynchronize(h->stream);
        throw;
    }
}

// The incremental part of recalc_curve_impl (bodies loaded, full-recalculation cases
// handled).
void incremental_update(rsx_engine* h, int start, int len, const std::vector<float>& old_bodies, bool sasa_now,
                        float* out32, double* out64)
{
    // An update over the moved atoms S = [start, start + len) and the atoms C whose
    // exposed-dot count (so class) changed without moving runs as two launches against the
    // columns as they were, S's slots skipped and C's records still the old ones:
    //   dI = 2 P1 x cols + P1 x P1  -  (2 P2 x cols + P2 x P2)
    // P2: the old records of S.  P1: the new records of S and, per atom of C, one record of
    // the delta class (factor F' - F) at its position.  Pairs within the rest cancel; a pair
    // (c, j), j unchanged, gets 2 (F'_c - F_c) F_j; a pair (c, d) of C gets
    // (F'_c - F_c) F_d + (F'_d - F_d) F_c from P1 x cols and (F'_c - F_c)(F'_d - F_d) from
    // P1 x P1, in all F'_c F'_d - F_c F_d; a pair (s, c) gets 2 (F'_s F_c - F_s F_c) from the
    // rectangles and 2 F'_s (F'_c - F_c) from P1 x P1.  The P2 launch needs no new count:
    // its rows are uploaded and S skipped ahead of the SASA recount, and it is queued on the
    // engine stream right behind the recount (which it would otherwise hold up), running
    // while the host reads the counts and builds P1; the P1 launch joins it on the second
    // stream, and both sums are added to I(q) on the engine stream, P2's first.  The records
    // that move between class runs travel with P1's upload and are written after the sums,
    // once I(q) is on its way to the host.  (Uploading P2 and skipping S on the second
    // stream, beside the recount, measured slower: the cross-stream wait costs more than it
    // overlaps.)
    const bool rebase = h->rebase_every > 0 && h->since_rebase + 1 >= h->rebase_every;
    // sasa_local_enqueue recounts exactly when these hold
    const bool local = sasa_now && len > 0 && (int)h->counts.size() == h->n && h->have_rsasa;
    KernelChoice kc;
    incremental_kernel(h, kc.variant, kc.K, kc.n_chunks, kc.chunks);
    const bool old_early = len > 0 && !rebase && (local || !sasa_now);
    OldPart op;
    if (old_early) op = stage_old_part(h, start, len, old_bodies, kc);
    if (local && !sasa_local_enqueue(h, start, len, old_bodies)) fail(RSX_HIP_ERROR, "local recount not queued");
    launch_old_part(h, op, kc);

    // atoms whose exposed-dot count changed
    std::vector<int> changed;
    const bool recounted = local && sasa_local_finish(h, changed);
    if (!recounted && sasa_now) {  // (a local recount that overflowed the slab lands here too)
        const std::vector<uint32_t> old_counts = h->counts;
        compute_sasa(h);
        for (int i = 0; i < h->n; ++i)
            if (old_counts[i] != h->counts[i]) changed.push_back(i);
    }
    // changed set M: S, then C ascending
    std::vector<int> M;
    M.reserve((size_t)len + changed.size());
    std::vector<char>& in_m = h->in_m;
    in_m.resize(h->n, 0);
    for (int i = start; i < start + len; ++i) {
        M.push_back(i);
        in_m[i] = 1;
    }
    for (int i : changed)
        if (!in_m[i]) {
            M.push_back(i);
            in_m[i] = 1;
        }
    for (int i : M) in_m[i] = 0;
    if (M.empty()) {
        copy_out(h, out32, out64);
        return;
    }
    if (rebase) {
        assign_slots(h);
        full_forward(h);
        copy_out(h, out32, out64);
        return;
    }
    if (!old_early) launch_old_part(h, stage_old_part(h, start, len, old_bodies, kc), kc);

    const int m = (int)M.size();
    std::vector<uint32_t> cls_new(m), cls_old(m);
    for (int a = 0; a < m; ++a) {
        cls_old[a] = h->cls[M[a]];
        cls_new[a] = class_of(h, M[a]);
    }
    // P1, sorted by class (stable), padded with skip records to a multiple of 64
    std::vector<std::pair<uint32_t, int>> p1_order((size_t)m);
    for (int a = 0; a < m; ++a) {
        const bool moved = M[a] >= start && M[a] < start + len;
        const uint32_t c = moved ? cls_new[a] : delta_class(h, (int)cls_new[a] - (int)cls_old[a]);
        p1_order[a] = {c, a};
    }
    std::stable_sort(p1_order.begin(), p1_order.end(),
                     [](const std::pair<uint32_t, int>& x, const std::pair<uint32_t, int>& y) { return x.first < y.first; });
    const int p1 = (m + 63) / 64 * 64;

    // q = 0: (sum F_new)^2 - (sum F_old)^2 from the class populations
    const double s_old = sum_f0(h);
    // run starts of the classes before the update (the populations are current: sum_f0)
    std::vector<int> run_start(h->n_class + 1, 0);
    for (int c = 0; c < h->n_class; ++c) run_start[c + 1] = run_start[c] + h->cls_count[c];
    for (int a = 0; a < m; ++a) {
        mark_dirty(h, M[a]);
        h->cls[M[a]] = cls_new[a];
        if (h->cls_count_ok) {
            h->cls_count[cls_old[a]]--;
            h->cls_count[cls_new[a]]++;
        }
    }
    const double s_new = sum_f0(h);
    // the new records, each in the run of its class (the slots of atoms whose class changed
    // are exchanged across runs)
    std::vector<int> touched;
    touched.reserve(2 * (size_t)m);
    relocate_classes(h, M, cls_old, cls_new, run_start, touched);
    const int nt = (int)touched.size();

    Pack pk;
    const size_t off_rows = pk.add((size_t)p1 * sizeof(float4));
    const size_t off_sb = pk.add(2 * sizeof(int));
    const size_t off_q0 = pk.add(sizeof(double));
    const size_t off_slot = pk.add((size_t)nt * sizeof(int));
    const size_t off_caller = pk.add((size_t)nt * sizeof(int));
    const size_t off_vals = pk.add((size_t)nt * sizeof(float4));
    unsigned char* hb = stage_begin(h, h->st_new, pk.size);
    float4* rows = reinterpret_cast<float4*>(hb + off_rows);
    for (int t = 0; t < m; ++t) {
        const float* nb = &h->bodies[4 * (size_t)M[p1_order[t].second]];
        rows[t] = make_float4(nb[0], nb[1], nb[2], bitsf(p1_order[t].first));
    }
    for (int t = m; t < p1; ++t) rows[t] = make_float4(0.f, 0.f, 0.f, bitsf(CLS_SKIP));
    for (int t = 0; t < nt; ++t) {
        const int sl = touched[t], i = h->caller_of[sl];
        const float* bb = &h->bodies[4 * (size_t)i];
        reinterpret_cast<int*>(hb + off_slot)[t] = sl;
        reinterpret_cast<int*>(hb + off_caller)[t] = i;
        reinterpret_cast<float4*>(hb + off_vals)[t] = make_float4(bb[0], bb[1], bb[2], bitsf(h->cls[i]));
    }
    unsigned char* db = h->st_new.dev.p;
    // the new-records launch reads the columns once S is skipped there and the engine
    // stream's earlier writes (the last update's relocation scatter) are done
    if (len == 0) hipchk(hipEventRecord(h->ev_skip, h->stream), "record skip");
    hipchk(hipStreamWaitEvent(h->stream2, h->ev_skip, 0), "wait skip");
    const rsx_engine::IncTable* tab =
        update_jobs(h, 1, reinterpret_cast<const float4*>(db + off_rows), p1, 1.0f, kc, h->stream2);
    reinterpret_cast<int*>(hb + off_sb)[0] = 0;
    reinterpret_cast<int*>(hb + off_sb)[1] = tab->njobs;
    *reinterpret_cast<double*>(hb + off_q0) = s_new * s_new - s_old * s_old;
    stage_upload(h->st_new, pk.size, h->stream2, "upload new rows");
    h->d_partials_new.ensure((size_t)tab->njobs * h->ldf);
    launch_update_pairs(h, tab->jobs.p, tab->njobs, h->d_partials_new.p, kc, h->stream2);
    hipchk(hipEventRecord(h->ev_new, h->stream2), "record new");
    hipchk(hipStreamWaitEvent(h->stream, h->ev_new, 0), "wait new");
    reduce_update_part(h, tab->njobs, h->d_partials_new.p, reinterpret_cast<const int*>(db + off_sb),
                       reinterpret_cast<const double*>(db + off_q0));
    enqueue_iq_download(h);
    // written after both launches (same stream, behind the wait for P1); the next call's
    // device work is queued behind it on this stream
    hipchk(rsx::launch_scatter_slots(h->d_cols.p, h->d_s2c.p, reinterpret_cast<const int*>(db + off_slot),
                                     reinterpret_cast<const float4*>(db + off_vals),
                                     reinterpret_cast<const int*>(db + off_caller), nt, h->stream),
           "scatter new");
    h->since_rebase++;
    finish_iq_download(h, out32, out64);
}

// the saved arrays have the current sizes (entry-wise copies are possible)
bool saved_matches_sizes(const rsx_engine* h)
{
    return h->saved.bodies.size() == h->bodies.size() && h->saved.counts.size() == h->counts.size() &&
           h->saved.cls.size() == h->cls.size() && h->saved.rsasa.size() == h->rsasa.size() &&
           h->saved.slot_of.size() == h->slot_of.size() && h->saved.caller_of.size() == h->caller_of.size();
}

// entry i of bodies / counts / cls / rsasa from src to dst (same sizes)
void copy_entry(std::vector<float>& db, std::vector<uint32_t>& dc, std::vector<uint32_t>& dk, std::vector<float>& dr,
                const std::vector<float>& sb, const std::vector<uint32_t>& sc, const std::vector<uint32_t>& sk,
                const std::vector<float>& sr, int i)
{
    std::memcpy(&db[4 * (size_t)i], &sb[4 * (size_t)i], 16);
    if ((size_t)i < dc.size()) dc[i] = sc[i];
    if ((size_t)i < dk.size()) dk[i] = sk[i];
    if ((size_t)i < dr.size()) dr[i] = sr[i];
}

void commit_impl(rsx_engine* h)
{
    if (h->state != 4) fail(RSX_NO_INITIAL_CALC, "commit() called before an initial call to calc_curve()");
    set_device(h);
    if (h->host_stale) {
        // the device holds a device-resident forward's bodies: commit the host bodies' state
        // (their exposed-dot counts on the host are current)
        assign_slots(h);
        full_forward(h);
        h->host_stale = false;
    }
    h->d_saved_cols.ensure(h->n);
    hipchk(hipMemcpyAsync(h->d_saved_cols.p, h->d_cols.p, (size_t)h->n * sizeof(float4), hipMemcpyDeviceToDevice,
                          h->stream),
           "save cols");
    hipchk(hipMemcpyAsync(h->d_saved_iq.p, h->d_iq.p, h->nq * sizeof(double), hipMemcpyDeviceToDevice, h->stream),
           "save iq");
    // no wait: the copies are ordered before every later use of these buffers (all on the
    // engine stream; the second stream reads the columns only behind an event recorded there)
    if (saved_matches_sizes(h) && !h->dirty_full) {
        for (int i : h->dirty) copy_entry(h->saved.bodies, h->saved.counts, h->saved.cls, h->saved.rsasa, h->bodies,
                                          h->counts, h->cls, h->rsasa, i);
        if (h->slots_dirty) {
            h->saved.slot_of = h->slot_of;
            h->saved.caller_of = h->caller_of;
        }
    } else {
        h->saved.bodies = h->bodies;
        h->saved.counts = h->counts;
        h->saved.cls = h->cls;
        h->saved.rsasa = h->rsasa;
        h->saved.slot_of = h->slot_of;
        h->saved.caller_of = h->caller_of;
    }
    h->dirty.clear();
    h->dirty_full = false;
    h->slots_dirty = false;
    h->saved.have_rsasa = h->have_rsasa;
    h->saved.since_rebase = h->since_rebase;
    h->xaction = true;
}

void revert_impl(rsx_engine* h)
{
    if (h->state != 4) fail(RSX_NO_INITIAL_CALC, "revert() called before an initial call to calc_curve()");
    if (!h->xaction) return;
    set_device(h);
    const bool partial = saved_matches_sizes(h) && !h->dirty_full;
    const bool layout_changed = !partial && h->saved.caller_of != h->caller_of;
    h->sasa_mirror = false;  // the device copy may hold the reverted positions
    auto restore_type = [h](int i) {
        uint32_t t;
        std::memcpy(&t, &h->bodies[4 * (size_t)i + 3], 4);
        h->type[i] = t;
    };
    const bool slots_moved = partial && h->slots_dirty;
    if (partial) {
        for (int i : h->dirty) {
            copy_entry(h->bodies, h->counts, h->cls, h->rsasa, h->saved.bodies, h->saved.counts, h->saved.cls,
                       h->saved.rsasa, i);
            restore_type(i);
        }
        if (slots_moved) {
            h->slot_of = h->saved.slot_of;
            h->caller_of = h->saved.caller_of;
        }
    } else {
        h->bodies = h->saved.bodies;
        h->counts = h->saved.counts;
        h->cls = h->saved.cls;
        h->rsasa = h->saved.rsasa;
        h->slot_of = h->saved.slot_of;
        h->caller_of = h->saved.caller_of;
        for (int i = 0; i < h->n; ++i) restore_type(i);
    }
    h->dirty.clear();
    h->dirty_full = false;
    h->slots_dirty = false;
    h->cls_count_ok = false;
    h->have_rsasa = h->saved.have_rsasa;
    h->since_rebase = h->saved.since_rebase;
    hipchk(hipMemcpyAsync(h->d_cols.p, h->d_saved_cols.p, (size_t)h->n * sizeof(float4), hipMemcpyDeviceToDevice,
                          h->stream),
           "restore cols");
    hipchk(hipMemcpyAsync(h->d_iq.p, h->d_saved_iq.p, h->nq * sizeof(double), hipMemcpyDeviceToDevice, h->stream),
           "restore iq");
    if (layout_changed || slots_moved || h->host_stale)
        hipchk(hipMemcpyAsync(h->d_s2c.p, h->caller_of.data(), h->n * sizeof(int), hipMemcpyHostToDevice, h->stream),
               "restore slot map");
    hipchk(hipStreamSynchronize(h->stream), "sync revert");
    h->host_stale = false;
}

// ---- standalone SASA -----------------------------------------------------------
// Distance-histogram I(q) (approximate): pair distances of the bodies binned at width dr
// per type pair (and, with hydration, weighted by the exposed-dot counts), each histogram
// transformed once per q; the factors enter only in the final FP64 combination
//   I = sum_i F_i^2 + 2 sum_b [sum_{t<=t'} f_t f_t' H0 + (g/D) sum_t f_t H1_t + (g/D)^2 H2](b) sinc(q r_b)
// with F_i = f_t + (count_i / D) g, g = w f_water.  The bins are the only approximation.
void calc_hist_impl(rsx_engine* h, const float* bodies, int n, float dr, double* out)
{
    if (h->nq <= 0 || h->n_types <= 0) fail(RSX_ARGS_NOT_SET, "histogram: q and factors must be set first");
    if (!bodies || n < 0 || !out || !(dr > 0.0f)) fail(RSX_INVALID_ARG, "histogram: invalid argument");
    const int nt = h->n_types, nq = h->nq;
    std::vector<int> type(n);
    std::vector<double> cen(3, 0.0);
    for (int i = 0; i < n; ++i) {
        uint32_t t;
        std::memcpy(&t, &bodies[4 * (size_t)i + 3], 4);
        if ((int)t >= nt) fail(RSX_INVALID_ARG, "histogram: body type index outside the factor table");
        type[i] = (int)t;
        for (int d = 0; d < 3; ++d) cen[d] += bodies[4 * (size_t)i + d];
    }
    set_device(h);
    const bool hyd = hydration_on(h);
    double r_max = 0.0;
    for (int i = 0; i < n; ++i) {
        double s2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double v = bodies[4 * (size_t)i + d] - cen[d] / std::max(1, n);
            s2 += v * v;
        }
        r_max = std::max(r_max, std::sqrt(s2));
    }
    const int n_bins = (int)std::min(4.0e6, 2.0 * r_max / (double)dr) + 2;
    const int n_pairs = nt * (nt + 1) / 2;
    std::vector<int> pidx((size_t)nt * nt);
    for (int a = 0, p = 0; a < nt; ++a)
        for (int b = a; b < nt; ++b, ++p) pidx[(size_t)a * nt + b] = pidx[(size_t)b * nt + a] = p;
    const int n_hist = n_pairs + (hyd ? nt + 1 : 0);
    // device buffers (engine-owned scratch of this entry point)
    h->d_sasa_atoms.ensure(std::max(n, 1));
    h->sasa_mirror = false;
    if (n > 0)
        hipchk(hipMemcpyAsync(h->d_sasa_atoms.p, bodies, (size_t)n * 16, hipMemcpyHostToDevice, h->stream),
               "upload bodies");
    std::vector<uint32_t> counts;
    if (hyd && n > 0) {
        run_sasa_device(h, n);
        counts.resize(n);
        hipchk(hipMemcpyAsync(counts.data(), h->d_counts.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream),
               "download counts");
    }
    const size_t n0 = (size_t)n_pairs * n_bins, n1 = hyd ? (size_t)nt * n_bins : 0, n2 = hyd ? (size_t)n_bins : 0;
    h->h_u32.ensure(std::max<size_t>(n0, 1));
    h->h_u64.ensure(std::max<size_t>(n1 + n2, 1));
    h->h_dbl.ensure((size_t)n_hist * n_bins);
    h->h_out.ensure((size_t)n_hist * nq);
    h->h_pidx.ensure(pidx.size());
    h->h_q.ensure(nq);
    std::vector<double> qd(h->q.begin(), h->q.end());
    hipchk(hipMemcpyAsync(h->h_pidx.p, pidx.data(), pidx.size() * sizeof(int), hipMemcpyHostToDevice, h->stream),
           "upload pair index");
    hipchk(hipMemcpyAsync(h->h_q.p, qd.data(), nq * sizeof(double), hipMemcpyHostToDevice, h->stream), "upload q");
    hipchk(rsx::launch_fill_i32((int*)h->h_u32.p, 0, (long long)n0, h->stream), "zero h0");
    if (hyd) hipchk(rsx::launch_fill_i32((int*)h->h_u64.p, 0, 2LL * (long long)(n1 + n2), h->stream), "zero h1 h2");
    const int n_win = (n_bins + rsx::HIST_WINDOW - 1) / rsx::HIST_WINDOW;
    if (!hyd && n > 1 && n_win <= 16) {
        // plain: atoms sorted by type, tiles of one type pair, LDS bin windows
        std::vector<int> seg(nt + 1, 0);
        for (int i = 0; i < n; ++i) seg[type[i] + 1]++;
        for (int t = 0; t < nt; ++t) seg[t + 1] += seg[t];
        std::vector<float4> sorted(n);
        std::vector<int> fill(seg.begin(), seg.end() - 1);
        for (int i = 0; i < n; ++i) {
            const float* b = &bodies[4 * (size_t)i];
            sorted[fill[type[i]]++] = make_float4(b[0], b[1], b[2], b[3]);
        }
        constexpr int CB = 4096;  // columns per tile
        std::vector<rsx::HistTile> tiles;
        for (int ta = 0; ta < nt; ++ta)
            for (int tb = ta; tb < nt; ++tb)
                for (int r0 = seg[ta]; r0 < seg[ta + 1]; r0 += rsx::HIST_ROWS) {
                    const int nr = std::min(rsx::HIST_ROWS, seg[ta + 1] - r0);
                    const int cs = ta == tb ? r0 : seg[tb];  // same type: columns from the block start on
                    for (int c0 = cs; c0 < seg[tb + 1]; c0 += CB)
                        tiles.push_back({r0, nr, c0, std::min(CB, seg[tb + 1] - c0), pidx[(size_t)ta * nt + tb],
                                         ta == tb ? 1 : 0});
                }
        hipchk(hipMemcpyAsync(h->d_sasa_atoms.p, sorted.data(), (size_t)n * 16, hipMemcpyHostToDevice, h->stream),
               "upload sorted bodies");
        h->h_tiles.ensure(tiles.size());
        hipchk(hipMemcpyAsync(h->h_tiles.p, tiles.data(), tiles.size() * sizeof(rsx::HistTile), hipMemcpyHostToDevice,
                              h->stream),
               "upload tiles");
        hipchk(rsx::launch_hist_tiles(h->d_sasa_atoms.p, h->h_tiles.p, (int)tiles.size(), 1.0f / dr, n_bins,
                                      h->h_u32.p, h->stream),
               "histogram tiles");
        hipchk(hipStreamSynchronize(h->stream), "sync histogram tiles");  // host tables above
    } else {
        rsx::HistArgs a;
        a.atoms = h->d_sasa_atoms.p;
        a.counts = hyd ? h->d_counts.p : nullptr;
        a.n = n;
        a.n_types = nt;
        a.inv_dr = 1.0f / dr;
        a.n_bins = n_bins;
        a.pair_index = h->h_pidx.p;
        a.h0 = h->h_u32.p;
        a.h1 = h->h_u64.p;
        a.h2 = h->h_u64.p + n1;
        hipchk(rsx::launch_hist_pairs(a, h->stream), "histogram");
    }
    hipchk(rsx::launch_widen_u32(h->h_u32.p, h->h_dbl.p, (long long)n0, n_bins, (double)dr, h->stream), "widen h0");
    if (hyd)
        hipchk(rsx::launch_widen_u64(h->h_u64.p, h->h_dbl.p + n0, (long long)(n1 + n2), n_bins, (double)dr, h->stream),
               "widen h1 h2");
    h->h_part.ensure(std::max<size_t>(1, rsx::hist_transform_scratch(n_hist, n_bins, nq)));
    hipchk(rsx::launch_hist_transform(h->h_dbl.p, n_hist, n_bins, (double)dr, h->h_q.p, nq, h->h_part.p, h->h_out.p,
                                      h->stream),
           "histogram transform");
    std::vector<double> A((size_t)n_hist * nq);
    hipchk(hipMemcpyAsync(A.data(), h->h_out.p, A.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream),
           "download transforms");
    hipchk(hipStreamSynchronize(h->stream), "sync histogram");
    // self terms per type: n_t, sum of counts, sum of squared counts
    std::vector<double> n_t(nt, 0.0), c1(nt, 0.0);
    double c2 = 0.0;
    for (int i = 0; i < n; ++i) {
        n_t[type[i]] += 1.0;
        if (hyd) {
            c1[type[i]] += counts[i];
            c2 += (double)counts[i] * counts[i];
        }
    }
    const double inv_d = hyd ? 1.0 / (double)h->n_dots : 0.0;
    for (int k = 0; k < nq; ++k) {
        const float* row = &h->table[(size_t)k * nt];
        const double g = hyd ? (double)h->w * (double)row[h->water] * inv_d : 0.0;  // per unit count
        double I = 0.0;
        for (int t = 0; t < nt; ++t) I += n_t[t] * (double)row[t] * row[t];
        if (hyd) {
            for (int t = 0; t < nt; ++t) I += 2.0 * g * row[t] * c1[t];
            I += g * g * c2;
        }
        double pair = 0.0;
        for (int t = 0; t < nt; ++t)
            for (int u = t; u < nt; ++u) pair += (double)row[t] * row[u] * A[(size_t)pidx[(size_t)t * nt + u] * nq + k];
        if (hyd) {
            for (int t = 0; t < nt; ++t) pair += g * row[t] * A[(size_t)(n_pairs + t) * nq + k];
            pair += g * g * A[(size_t)(n_pairs + nt) * nq + k];
        }
        out[k] = I + 2.0 * pair;
    }
}

void calc_rsasa_impl(rsx_engine* h, const float* bodies, int n, const float* radii, int nt, float probe,
                     const float* dots, int n_dots, unsigned* out_counts, float* out_frac)
{
    if (!bodies || !radii || !dots || n < 0 || nt <= 0 || n_dots <= 0 || n_dots > 512)
        fail(RSX_INVALID_ARG, "calc_rsasa: invalid argument");
    for (int i = 0; i < n; ++i) {
        uint32_t t;
        std::memcpy(&t, &bodies[4 * i + 3], 4);
        if ((int)t >= nt) fail(RSX_INVALID_ARG, "calc_rsasa: type index outside the radius table");
    }
    set_device(h);
    // use a private copy of the SASA tables
    const std::vector<float> keep_radii = h->radii, keep_dots = h->dots, keep_r2 = h->r2, keep_rext = h->rext;
    const float keep_probe = h->probe;
    const int keep_nd = h->n_dots;
    h->radii.assign(radii, radii + nt);
    h->probe = probe;
    h->n_dots = n_dots;
    h->dots.assign(dots, dots + (size_t)nt * n_dots * 4);
    upload_sasa_tables(h);
    if (n > 0) {
        h->d_sasa_atoms.ensure(n);
        h->sasa_mirror = false;
        hipchk(hipMemcpyAsync(h->d_sasa_atoms.p, bodies, (size_t)n * 16, hipMemcpyHostToDevice, h->stream),
               "upload atoms");
        run_sasa_device(h, n);
        std::vector<uint32_t> c(n);
        hipchk(hipMemcpyAsync(c.data(), h->d_counts.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream),
               "download counts");
        hipchk(hipStreamSynchronize(h->stream), "sync");
        for (int i = 0; i < n; ++i) {
            if (out_counts) out_counts[i] = c[i];
            if (out_frac) out_frac[i] = (float)c[i] / (float)n_dots;
        }
    }
    h->radii = keep_radii;
    h->dots = keep_dots;
    h->probe = keep_probe;
    h->n_dots = keep_nd;
    h->r2 = keep_r2;
    h->rext = keep_rext;
    if (h->sasa_set) upload_sasa_tables(h);
}

// ---- batch --------------------------------------------------------------------
void batch_release_graph(rsx_engine* h)
{
    if (h->b_exec) (void)hipGraphExecDestroy(h->b_exec);
    if (h->b_graph) (void)hipGraphDestroy(h->b_graph);
    h->b_exec = nullptr;
    h->b_graph = nullptr;
}

// the batch is laid out again (from the frames kept on the host) on its next launch
void batch_invalidate(rsx_engine* h)
{
    batch_release_graph(h);
    h->b_layout_ok = false;
}

// Exposed-dot counts of every frame's bodies (one SASA count per frame, each frame on
// its own: atoms of other frames never occlude), into b_counts (body order).
void batch_counts(rsx_engine* h)
{
    const int total = h->b_offsets[h->b_frames];
    h->b_counts.assign(total, 0u);
    for (int i = 0; i < total; ++i) {
        uint32_t t;
        std::memcpy(&t, &h->b_bodies[4 * (size_t)i + 3], 4);
        if (t >= h->radii.size()) fail(RSX_INVALID_ARG, "batch: body type index outside the SASA radius table");
    }
    int nmax = 0;
    for (int f = 0; f < h->b_frames; ++f) nmax = std::max(nmax, h->b_offsets[f + 1] - h->b_offsets[f]);
    h->d_sasa_atoms.ensure(std::max(nmax, 1));
    h->sasa_mirror = false;  // it holds the frames now, not the single structure
    for (int f = 0; f < h->b_frames; ++f) {
        const int b = h->b_offsets[f], m = h->b_offsets[f + 1] - b;
        if (m <= 0) continue;
        hipchk(hipMemcpyAsync(h->d_sasa_atoms.p, &h->b_bodies[4 * (size_t)b], (size_t)m * 16, hipMemcpyHostToDevice,
                              h->stream),
               "upload frame atoms");
        run_sasa_device(h, m);
        hipchk(hipMemcpyAsync(&h->b_counts[b], h->d_counts.p, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              h->stream),
               "download frame counts");
        hipchk(hipStreamSynchronize(h->stream), "sync frame sasa");
    }
}

// Lay the frames out: the path and q chunks for their largest bounding radius, classes
// (type, or type x exposed-dot count with hydration), columns class-sorted within each
// frame, and one job table over every frame's triangle (longest first).  The class
// factors come later (batch_factors): they change with the factor table, the layout
// does not.
void batch_layout(rsx_engine* h)
{
    batch_release_graph(h);
    h->b_layout_ok = false;
    const int nf = h->b_frames;
    const std::vector<int>& off = h->b_offsets;
    const float* bodies = h->b_bodies.data();
    for (int i = 0; i < off[nf]; ++i) {
        uint32_t t;
        std::memcpy(&t, &bodies[4 * (size_t)i + 3], 4);
        if ((int)t >= h->n_types) fail(RSX_INVALID_ARG, "batch: type index outside the factor table");
    }
    // path: recurrence where every frame fits it
    double rmax = 0.0;
    for (int f = 0; f < nf; ++f)
        if (off[f + 1] > off[f]) rmax = std::max(rmax, bounding_radius(bodies + 4 * (size_t)off[f], off[f + 1] - off[f], nullptr));
    h->b_radius = rmax;
    h->b_uniform = fits_recurrence(h, rmax);
    h->b_variant = h->b_var_req >= 0 ? h->b_var_req : h->variant;
    pick_chunks(h, h->b_uniform, h->b_variant, h->b_k_req > 0 ? h->b_k_req : h->k_req, 0, h->b_K, h->b_n_chunks);
    h->b_ldf = h->b_K * h->b_n_chunks;
    upload_q_consts(h, h->b_K, h->b_n_chunks, h->b_chunks, h->b_invq, h->b_qrev, h->b_qsplit, h->b_qv);
    // classes
    h->b_hydrate = hydration_on(h);
    if (h->b_hydrate) batch_counts(h);
    h->b_d1 = h->b_hydrate ? h->n_dots + 1 : 1;
    h->b_n_types = h->n_types;
    h->b_n_class = h->n_types * h->b_d1;
    std::vector<float4> cols(std::max(off[nf], 1));
    h->b_s2c.assign(std::max(off[nf], 1), 0);
    h->b_pop_off.assign(nf + 1, 0);
    h->b_pop.clear();
    std::vector<int> pop(h->b_n_class, 0);
    for (int f = 0; f < nf; ++f) {
        const int b = off[f], e = off[f + 1];
        std::vector<uint32_t> cl(e - b);
        for (int i = b; i < e; ++i) {
            uint32_t t;
            std::memcpy(&t, &bodies[4 * (size_t)i + 3], 4);
            cl[i - b] = t * (uint32_t)h->b_d1 + (h->b_hydrate ? h->b_counts[i] : 0u);
        }
        std::vector<int> idx(e - b);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return cl[x] < cl[y]; });
        for (int a = 0; a < e - b; ++a) {
            const size_t i = (size_t)(b + idx[a]);
            h->b_s2c[b + a] = (int)i;
            cols[b + a] = make_float4(bodies[4 * i], bodies[4 * i + 1], bodies[4 * i + 2], bitsf(cl[idx[a]]));
            pop[cl[idx[a]]]++;
        }
        for (int c = 0; c < h->b_n_class; ++c)
            if (pop[c]) {
                h->b_pop.emplace_back(c, pop[c]);
                pop[c] = 0;
            }
        h->b_pop_off[f + 1] = (int)h->b_pop.size();
    }
    h->b_cols.ensure(cols.size());
    hipchk(hipMemcpy(h->b_cols.p, cols.data(), cols.size() * sizeof(float4), hipMemcpyHostToDevice), "upload frames");
    // jobs grouped by frame: slots of frame f are [sb[f], sb[f+1])
    double work = 0.0;
    for (int f = 0; f < nf; ++f) {
        const double m = off[f + 1] - off[f];
        work += m * m / 2.0;
    }
    const int L = h->b_range_req > 0   ? h->b_range_req
                  : h->range_req > 0 ? h->range_req
                                     : pick_range_for(h, work, h->b_uniform, h->b_variant, h->b_K, h->b_n_chunks,
                                                      h->b_radius, 6.0);
    h->b_L = L;
    std::vector<PairJob> jobs;
    std::vector<int> sb(nf + 1, 0);
    for (int f = 0; f < nf; ++f) {
        std::vector<PairJob> fj;
        triangle_jobs(fj, h->b_cols.p + off[f], off[f + 1] - off[f], L, 2.0f);
        for (auto& j : fj) {
            j.slot = (int)jobs.size();
            jobs.push_back(j);
        }
        sb[f + 1] = (int)jobs.size();
    }
    // execution order: longest first; slots keep the frame grouping
    std::stable_sort(jobs.begin(), jobs.end(),
                     [](const PairJob& a, const PairJob& b) { return (long)a.ncol * a.nrow > (long)b.ncol * b.nrow; });
    h->b_njobs = (int)jobs.size();
    h->b_slots = h->b_njobs;
    h->b_max_slots = 0;
    for (int f = 0; f < nf; ++f) h->b_max_slots = std::max(h->b_max_slots, sb[f + 1] - sb[f]);
    h->b_jobs.ensure(jobs.size());
    h->b_partials.ensure((size_t)std::max(1, h->b_slots) * h->b_ldf);
    h->b_red.ensure(std::max<size_t>(1, rsx::reduce_scratch_doubles(h->nq, nf, h->b_max_slots)));
    h->b_slot_begin.ensure(nf + 1);
    h->b_out.ensure((size_t)nf * h->nq);
    h->b_mean.ensure(h->nq);
    h->b_q0.ensure(nf);
    h->b_ftab.ensure((size_t)(h->b_n_class + 1) * h->b_ldf);
    hipchk(hipMemcpy(h->b_jobs.p, jobs.data(), jobs.size() * sizeof(PairJob), hipMemcpyHostToDevice), "upload jobs");
    hipchk(hipMemcpy(h->b_slot_begin.p, sb.data(), sb.size() * sizeof(int), hipMemcpyHostToDevice), "upload sb");
    h->b_nq = h->nq;
    h->b_sasa_stamp = h->sasa_stamp;
    h->b_factor_stamp = 0;  // class factors still to come
    h->b_gen++;
    h->b_layout_ok = true;
}

// Class factors of the batch layout from the engine's current factor table and water
// weight, and each frame's exact I(0) = (sum of F(0))^2 from its class populations.
// Buffers keep their addresses, so a captured batch graph stays valid.
void batch_factors(rsx_engine* h)
{
    const int nf = h->b_frames, ldf = h->b_ldf, nc = h->b_n_class;
    hipchk(hipStreamSynchronize(h->stream), "sync before factor upload");  // pinned buffers free
    h->b_pin_ftab.ensure((size_t)(nc + 1) * ldf);
    h->b_pin_q0.ensure(nf);
    std::memset(h->b_pin_ftab.p, 0, (size_t)(nc + 1) * ldf * sizeof(float));
    class_factor_rows(h, h->b_d1, ldf, h->b_f64, h->b_pin_ftab.p);
    for (int f = 0; f < nf; ++f) {
        double s = 0.0;
        if (h->q0_index >= 0)
            for (int p = h->b_pop_off[f]; p < h->b_pop_off[f + 1]; ++p)
                s += (double)h->b_pop[p].second * h->b_f64[(size_t)h->b_pop[p].first * h->nq + h->q0_index];
        h->b_pin_q0.p[f] = s * s;
    }
    hipchk(hipMemcpyAsync(h->b_ftab.p, h->b_pin_ftab.p, (size_t)(nc + 1) * ldf * sizeof(float), hipMemcpyHostToDevice,
                          h->stream),
           "upload batch factors");
    hipchk(hipMemcpyAsync(h->b_q0.p, h->b_pin_q0.p, nf * sizeof(double), hipMemcpyHostToDevice, h->stream),
           "upload batch q0");
    h->b_factor_stamp = h->factor_stamp;
}

void batch_set_frames_impl(rsx_engine* h, const float* bodies, const int* offsets, int nf)
{
    if (h->nq <= 0 || h->n_types <= 0) fail(RSX_ARGS_NOT_SET, "batch: q and factors must be set first");
    if (nf <= 0 || !offsets || !bodies) fail(RSX_INVALID_ARG, "batch: invalid frames");
    if (offsets[0] != 0) fail(RSX_INVALID_ARG, "batch: offsets[0] must be 0");
    for (int f = 0; f < nf; ++f)
        if (offsets[f + 1] < offsets[f]) fail(RSX_INVALID_ARG, "batch: offsets must not decrease");
    set_device(h);
    batch_invalidate(h);
    h->b_bodies.assign(bodies, bodies + 4 * (size_t)offsets[nf]);
    h->b_offsets.assign(offsets, offsets + nf + 1);
    h->b_frames = nf;
    try {
        batch_layout(h);
    } catch (...) {
        h->b_frames = 0;
        throw;
    }
}

// waves a persistent launch keeps resident: the CUs times the kernel's resident
// one-wave blocks per CU (from the runtime's occupancy calculator); 0 (one wave per
// item instead) where the work-queue loop would cost the kernel occupancy
int persistent_waves(const rsx_engine* h, int variant, int K, double radius)
{
    const int st = rsx::sin_series_terms(h->dq * radius);
    const int pb = rsx::rec32_blocks_per_cu(variant, K, st, true);
    const int ib = rsx::rec32_blocks_per_cu(variant, K, st, false);
    return pb >= ib ? h->n_cu * pb : 0;
}

rsx::DebyeArgs batch_debye_args(rsx_engine* h)
{
    rsx::DebyeArgs a;
    a.jobs = h->b_jobs.p;
    a.n_chunks = h->b_n_chunks;
    a.ftab = h->b_ftab.p;
    a.ldf = h->b_ldf;
    a.zero_class = h->b_n_class;
    a.chunks = h->b_chunks.p;
    a.half_dq = 0.5 * h->dq;
    a.dq = (float)h->dq;
    a.invq = h->b_invq.p;
    a.qrev = h->b_qrev.p;
    a.qsplit = h->b_qsplit.p;
    a.phase32 = phase32_ok(h, h->b_radius) ? 1 : 0;
    a.qv = h->b_qv.p;
    a.partials = h->b_partials.p;
    a.variant = h->b_variant;
    a.queue = nullptr;
    a.persistent_waves = 0;
    a.sin_terms = rsx::sin_series_terms(h->dq * h->b_radius);
    return a;
}

void batch_enqueue(rsx_engine* h, double* out_frames, double* out_mean, int pers_waves)
{
    rsx::DebyeArgs a = batch_debye_args(h);
    if (pers_waves > 0) {
        a.queue = h->d_queue.p;
        a.persistent_waves = pers_waves;
    }
    hipchk(rsx::launch_debye(a, h->b_njobs, h->b_K, h->b_uniform, h->stream), "batch debye");
    hipchk(rsx::launch_reduce(h->b_partials.p, h->b_ldf, h->nq, h->b_slot_begin.p, h->b_frames, out_frames, h->nq, 0,
                              h->b_q0.p, h->q0_index, h->b_max_slots, h->b_red.p, h->stream),
           "batch reduce");
    hipchk(rsx::launch_frame_mean(out_frames, h->b_frames, h->nq, out_mean, h->stream), "frame mean");
}

// Launch the batch (graph-captured) into out_frames [n_frames][n_q] and out_mean [n_q] on
// the device.  The layout is redone when q, the tuning, the factor table's type count,
// the SASA set-up or a switch to hydration made it stale; the class factors are
// refreshed when the factor table or the water weight changed (the graph stays).
void batch_prepare(rsx_engine* h)
{
    if (h->b_frames <= 0) fail(RSX_ARGS_NOT_SET, "batch: no frames set");
    set_device(h);
    const bool stale = !h->b_layout_ok || h->b_nq != h->nq || h->b_n_types != h->n_types ||
                       (hydration_on(h) && !h->b_hydrate) || (h->b_hydrate && h->b_sasa_stamp != h->sasa_stamp);
    if (stale) batch_layout(h);
    if (h->b_factor_stamp != h->factor_stamp) batch_factors(h);
}

void batch_run_dev(rsx_engine* h, double* out_frames, double* out_mean)
{
    batch_prepare(h);
    if (!h->b_exec || h->b_graph_out != out_frames || h->b_graph_mean != out_mean) {
        batch_release_graph(h);
        const bool fp32 = h->b_variant == rsx::VARIANT_FP32 || h->b_variant == rsx::VARIANT_FP32R;
        const bool pers = h->b_uniform && fp32 && (h->b_pers_req >= 0 ? h->b_pers_req != 0 : h->persistent);
        const int pw = pers ? persistent_waves(h, h->b_variant, h->b_K, h->b_radius) : 0;
        h->b_pers_waves = pw;
        hipchk(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            batch_enqueue(h, out_frames, out_mean, pw);
        } catch (...) {
            hipGraph_t g;
            (void)hipStreamEndCapture(h->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipchk(hipStreamEndCapture(h->stream, &h->b_graph), "end capture");
        hipchk(hipGraphInstantiate(&h->b_exec, h->b_graph, nullptr, nullptr, 0), "instantiate graph");
        h->b_graph_out = out_frames;
        h->b_graph_mean = out_mean;
    }
    hipchk(hipGraphLaunch(h->b_exec, h->stream), "graph launch");
}

// ---- batched VJP ---------------------------------------------------------------
// grad [atoms of all frames][3] = sum_k w_f,k dI_f,k / dx_i, f the frame of atom i: one launch
// of the FP32 adjoint kernel over every frame's triangle of pair jobs (each job carrying its
// frame, whose weights scale the class tables inside the kernel) and one fixed-order
// reduction.  The layout (the adjoint's q chunking and column range, the jobs, the per-frame
// triangle tables, the class tables 2 F_c,k and 2 F_c,k / q_k) is kept until the batch layout
// or its class factors change.
// Class-pair coefficient form of the FP32 adjoint: at most this many classes, and its table
// (groups x classes^2 x chunks x stride floats) at most this many floats; the per-lane form
// otherwise
constexpr int ADJ_PAIR_MAX_CLASSES = 64;
constexpr size_t ADJ_PAIR_MAX_FLOATS = (size_t)1 << 26;

struct AdjChoice {
    int K = 0;
    bool pair = false;
};

// q per thread of the FP32 adjoint and its form: the class-pair form where the layout is plain
// (one class per type, d1 == 1), the classes few, the table small and a requested K one of its
// own; the per-lane form otherwise.  Cost per (pair, chunk) in FP32-op units: class-pair
// 2 K + 40 (its coefficients come through scalar loads, so K = 128 keeps 6 waves per SIMD),
// per-lane 2.4 K + 30 (tables in vector registers: 3 waves per SIMD up to K = 67 and 2 at
// K = 100, where the per-pair seeds spread over more q make up for it: 1.08 vs 1.09-1.12 ms
// at 10k x 200).
AdjChoice choose_adjoint32_k(const rsx_engine* h, int nq, int d1, int n_class, size_t groups)
{
    AdjChoice ch;
    double best = 1e300;
    if (h->adj_pair_ok && d1 == 1 && n_class >= 1 && n_class <= ADJ_PAIR_MAX_CLASSES) {
        for (int i = 0; i < rsx::adjoint_pair_k_count(); ++i) {
            const int k = rsx::adjoint_pair_k_value(i);
            if (h->adj_k_req > 0 && k != h->adj_k_req) continue;
            const int c = (nq + k - 1) / k;
            const size_t floats = groups * n_class * n_class * (size_t)c * rsx::adjoint_pair_stride(k);
            if (floats > ADJ_PAIR_MAX_FLOATS) continue;
            const double cost = c * (2.0 * k + 40.0);
            if (cost < best) {
                best = cost;
                ch.K = k;
                ch.pair = true;
            }
        }
        if (ch.K > 0) return ch;
    }
    for (int i = 0; i < rsx::adjoint_k_count(true); ++i) {
        const int k = rsx::adjoint_k_value(true, i);
        if (h->adj_k_req > 0 && k != h->adj_k_req) continue;
        const double cost = ((nq + k - 1) / k) * (2.4 * k + 30.0);
        if (cost < best) {
            best = cost;
            ch.K = k;
        }
    }
    if (ch.K == 0) fail(RSX_INVALID_ARG, "adjoint: the requested K is not available for this layout");
    return ch;
}

// Columns per adjoint job: the requested range, else for the class-pair form 64 when the
// launch is small (under 2^20 waves at 128 columns: the finer jobs balance the SIMDs, 0.98 vs
// 1.08 ms at 10k x 200) and 128 otherwise (the same speed, half the row partials: 512 frames of
// 1k-15k atoms x 128 q); 128 for the per-lane form.  waves128 = (job, chunk) items at 128.
int adjoint_col_range(const rsx_engine* h, bool pair, double waves128)
{
    if (h->adj_range_req > 0) return h->adj_range_req;
    return pair && waves128 < (double)(1 << 20) ? 64 : 128;
}

void batch_adjoint_layout(rsx_engine* h)
{
    if (h->ba_gen == h->b_gen && h->ba_factor == h->b_factor_stamp && h->ba_njobs >= 0 && h->ba_K > 0) return;
    const int nf = h->b_frames, nq = h->nq;
    const std::vector<int>& off = h->b_offsets;
    const int n = off[nf];
    const AdjChoice choice = choose_adjoint32_k(h, nq, h->b_d1, h->b_n_class, (size_t)std::max(1, nf));
    const int K = choice.K;
    const int nch = (nq + K - 1) / K, ldfa = nch * K;
    std::vector<ChunkConst> cc(nch);
    for (int c = 0; c < nch; ++c) cc[c] = chunk_const(h, c * K);
    const int ncl = std::max(1, h->b_n_class);
    std::vector<float> ft((size_t)ncl * ldfa, 0.0f);
    for (int c = 0; c < h->b_n_class; ++c)
        for (int k = 0; k < nq; ++k) ft[(size_t)c * ldfa + k] = (float)h->b_f64[(size_t)c * nq + k];
    std::vector<double> invq(ldfa, 0.0), ones(ldfa, 0.0);
    std::vector<float> invq32(ldfa, 0.0f);
    for (int k = 0; k < nq; ++k) {
        invq[k] = h->q[k] != 0.0f ? 1.0 / (double)h->q[k] : 0.0;
        invq32[k] = (float)invq[k];
        ones[k] = 1.0;
    }
    // every frame's triangle (creation order: row block I, then its column ranges m)
    double waves128 = 0.0;
    for (int f = 0; f < nf; ++f) {
        const double m = off[f + 1] - off[f];
        waves128 += std::ceil(m / 64.0) * std::ceil(m / 128.0) * 0.5 * nch;
    }
    const int L = adjoint_col_range(h, choice.pair, waves128);
    std::vector<PairJob> jobs;
    std::vector<int> jframe, jblock, jm, tri_off(std::max(nf, 1), 0);
    size_t tri_n = 0;
    for (int f = 0; f < nf; ++f) {
        const int m = off[f + 1] - off[f];
        tri_off[f] = (int)tri_n;
        if (m <= 0) continue;
        triangle_jobs(jobs, h->b_cols.p + off[f], m, L, 1.0f);
        const int T = (m + 63) / 64, stride = (m + L - 1) / L;
        for (int I = 0; I < T; ++I)
            for (int c0 = I * 64, mm = 0; c0 < m; c0 += L, ++mm) {
                jframe.push_back(f);
                jblock.push_back(I);
                jm.push_back(mm);
            }
        tri_n += (size_t)T * stride;
    }
    if (tri_n > (size_t)INT32_MAX || jobs.size() > (size_t)INT32_MAX / 64) fail(RSX_INVALID_ARG, "batch vjp: too many jobs");
    for (size_t j = 0; j < jobs.size(); ++j) jobs[j].pad = (int)j;
    finalize_jobs(jobs);
    std::vector<int> tri(std::max<size_t>(tri_n, 1), 0);
    for (size_t sl = 0; sl < jobs.size(); ++sl) {
        const int o = jobs[sl].pad, f = jframe[o];
        const int stride = (off[f + 1] - off[f] + L - 1) / L;
        tri[(size_t)tri_off[f] + (size_t)jblock[o] * stride + jm[o]] = (int)sl;
        jobs[sl].pad = f;
    }
    std::vector<int> atom_frame(std::max(n, 1), 0);
    for (int f = 0; f < nf; ++f)
        for (int i = off[f]; i < off[f + 1]; ++i) atom_frame[i] = f;
    const int nj = (int)jobs.size();
    h->ba_ftab.ensure(ft.size());
    h->ba_cb.ensure(ft.size());
    h->ba_cg.ensure(ft.size());
    h->ba_chunks.ensure(nch);
    h->ba_invq.ensure(ldfa);
    h->ba_ones.ensure(ldfa);
    h->ba_invq32.ensure(ldfa);
    h->ba_jobs.ensure(std::max(1, nj));
    h->ba_tri.ensure(tri.size());
    h->ba_tri_off.ensure(tri_off.size());
    h->ba_atom_frame.ensure(atom_frame.size());
    h->ba_frame_off.ensure(off.size());
    h->ba_s2c.ensure(h->b_s2c.size());
    h->ba_rowpart.ensure((size_t)std::max(1, nj) * nch * 64 * 3);
    h->ba_colpart.ensure((size_t)std::max(1, nj) * nch * L * 3);
    h->ba_grad.ensure((size_t)std::max(1, n) * 3);
    if (choice.pair)
        h->ba_coef.ensure((size_t)std::max(1, nf) * h->b_n_class * h->b_n_class * nch * rsx::adjoint_pair_stride(K));
    hipchk(hipMemcpy(h->ba_ftab.p, ft.data(), ft.size() * sizeof(float), hipMemcpyHostToDevice), "upload ftab");
    hipchk(hipMemcpy(h->ba_chunks.p, cc.data(), nch * sizeof(ChunkConst), hipMemcpyHostToDevice), "upload chunks");
    hipchk(hipMemcpy(h->ba_invq.p, invq.data(), ldfa * sizeof(double), hipMemcpyHostToDevice), "upload invq");
    hipchk(hipMemcpy(h->ba_ones.p, ones.data(), ldfa * sizeof(double), hipMemcpyHostToDevice), "upload ones");
    hipchk(hipMemcpy(h->ba_invq32.p, invq32.data(), ldfa * sizeof(float), hipMemcpyHostToDevice), "upload invq32");
    if (nj > 0)
        hipchk(hipMemcpy(h->ba_jobs.p, jobs.data(), nj * sizeof(PairJob), hipMemcpyHostToDevice), "upload jobs");
    hipchk(hipMemcpy(h->ba_tri.p, tri.data(), tri.size() * sizeof(int), hipMemcpyHostToDevice), "upload tri");
    hipchk(hipMemcpy(h->ba_tri_off.p, tri_off.data(), tri_off.size() * sizeof(int), hipMemcpyHostToDevice),
           "upload tri offsets");
    hipchk(hipMemcpy(h->ba_atom_frame.p, atom_frame.data(), atom_frame.size() * sizeof(int), hipMemcpyHostToDevice),
           "upload atom frames");
    hipchk(hipMemcpy(h->ba_frame_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice),
           "upload frame offsets");
    hipchk(hipMemcpy(h->ba_s2c.p, h->b_s2c.data(), h->b_s2c.size() * sizeof(int), hipMemcpyHostToDevice),
           "upload batch slot map");
    // class tables of the kernel's per-frame-weight form: 2 F_c,k and 2 F_c,k / q_k
    hipchk(rsx::launch_adjoint_coef(h->ba_ftab.p, h->ba_ones.p, h->ba_invq.p, ncl, ldfa, h->ba_cb.p, h->ba_cg.p,
                                    h->stream),
           "adjoint class tables");
    h->ba_K = K;
    h->ba_pair = choice.pair;
    h->ba_nch = nch;
    h->ba_ldf = ldfa;
    h->ba_L = L;
    h->ba_njobs = nj;
    h->ba_n = n;
    h->ba_gen = h->b_gen;
    h->ba_factor = h->b_factor_stamp;
    h->ba_have_w = false;
}

// the batched adjoint on the frame weights in ba_w: grad into ba_grad (caller order)
void launch_batch_adjoint_pass(rsx_engine* h)
{
    rsx::AdjointArgs a;
    a.jobs = h->ba_jobs.p;
    a.n_chunks = h->ba_nch;
    a.ftab = h->ba_ftab.p;
    a.ldf = h->ba_ldf;
    a.chunks = h->ba_chunks.p;
    a.half_dq = 0.5 * h->dq;
    a.wq = h->ba_ones.p;
    a.invq = h->ba_invq.p;
    a.invq32 = h->ba_invq32.p;
    a.cb = h->ba_cb.p;
    a.cg = h->ba_cg.p;
    a.dq_rev = h->dq / (2.0 * PI);
    a.fp32 = 1;
    a.sin_terms = rsx::sin_series_terms(h->dq * h->b_radius);
    a.col_range = h->ba_L;
    a.row_partial = h->ba_rowpart.p;
    a.col_partial = h->ba_colpart.p;
    a.n_class = h->b_n_class;
    if (h->ba_pair) {  // the frame weights go into the class-pair table
        rsx::AdjointPairCoefArgs pc;
        pc.ftab = h->ba_ftab.p;
        pc.cb = h->ba_cb.p;
        pc.cg = h->ba_cg.p;
        pc.wframe = h->ba_w.p;
        pc.groups = std::max(1, h->b_frames);
        pc.n_class = h->b_n_class;
        pc.n_chunks = h->ba_nch;
        pc.K = h->ba_K;
        pc.ldf = h->ba_ldf;
        pc.coef = h->ba_coef.p;
        hipchk(rsx::launch_adjoint_pair_coef(pc, h->stream), "batch adjoint class-pair table");
        a.coef = h->ba_coef.p;
        a.coef_grouped = 1;
    } else {
        a.wframe = h->ba_w.p;
    }
    hipchk(rsx::launch_adjoint(a, h->ba_njobs, h->ba_K, h->stream), "batch adjoint kernel");
    rsx::AdjointBatchReduceArgs r;
    r.row_partial = h->ba_rowpart.p;
    r.col_partial = h->ba_colpart.p;
    r.tri_slot = h->ba_tri.p;
    r.tri_off = h->ba_tri_off.p;
    r.atom_frame = h->ba_atom_frame.p;
    r.frame_off = h->ba_frame_off.p;
    r.n = h->ba_n;
    r.n_chunks = h->ba_nch;
    r.col_range = h->ba_L;
    r.slot_to_caller = h->ba_s2c.p;
    r.grad = h->ba_grad.p;
    hipchk(rsx::launch_adjoint_batch_reduce(r, h->stream), "batch adjoint reduce");
}

void batch_vjp_impl(rsx_engine* h, const double* w, int n_frames, int n_q, float* out_grad)
{
    if (!w || !out_grad) fail(RSX_INVALID_ARG, "batch vjp: null argument");
    batch_prepare(h);
    if (n_frames != h->b_frames || n_q != h->nq) fail(RSX_INVALID_ARG, "batch vjp: weights must be [frames][n_q]");
    if (!h->b_uniform) fail(RSX_NOT_SUPPORTED, "batch vjp needs a uniform q grid the recurrence covers");
    batch_adjoint_layout(h);
    const int ldfa = h->ba_ldf;
    std::vector<float> wf((size_t)std::max(1, n_frames) * ldfa, 0.0f);
    for (int f = 0; f < n_frames; ++f)
        for (int k = 0; k < n_q; ++k)
            if (h->q[k] != 0.0f) wf[(size_t)f * ldfa + k] = (float)w[(size_t)f * n_q + k];
    h->ba_w.ensure(wf.size());
    hipchk(hipMemcpyAsync(h->ba_w.p, wf.data(), wf.size() * sizeof(float), hipMemcpyHostToDevice, h->stream),
           "upload frame weights");
    launch_batch_adjoint_pass(h);
    h->ba_have_w = true;
    if (h->ba_n > 0)
        hipchk(hipMemcpyAsync(out_grad, h->ba_grad.p, (size_t)h->ba_n * 3 * sizeof(float), hipMemcpyDeviceToHost,
                              h->stream),
               "download grad");
    hipchk(hipStreamSynchronize(h->stream), "sync batch vjp");  // wf and out_grad are the caller's memory
}

// ---- adjoint -----------------------------------------------------------------
// Layout of the adjoint (its own q chunking; FP64 recurrence, so its own K list):
// rebuilt when the forward layout or the class table changes, kept otherwise, so a
// fitting loop allocates nothing per call.
void adjoint_layout(rsx_engine* h)
{
    const int n = h->n;
    const uint64_t stamp = h->layout_stamp;
    const bool fp32 = h->adj_fp32;
    if (h->a_stamp == stamp && h->a_n == n && h->a_cols == h->d_cols.p && h->a_fp32 == fp32) return;
    // q per thread: FP32 by choose_adjoint32_k; FP64 in FP64-op units per (pair, chunk), 4K ops
    // (two recurrences, combine, accumulate) + ~60 for distance and seeds
    int K = 0;
    bool pair = false;
    if (fp32) {
        const AdjChoice choice = choose_adjoint32_k(h, h->nq, h->d1, h->n_class, 1);
        K = choice.K;
        pair = choice.pair;
    } else {
        double best = 1e300;
        for (int i = 0; i < rsx::adjoint_k_count(false); ++i) {
            const int k = rsx::adjoint_k_value(false, i);
            if (h->adj_k_req > 0 && k != h->adj_k_req) continue;
            const double cost = ((h->nq + k - 1) / k) * (4.0 * k + 60.0);
            if (cost < best) {
                best = cost;
                K = k;
            }
        }
        if (K == 0) fail(RSX_INVALID_ARG, "adjoint: the requested K is not available for the FP64 kernel");
    }
    const int nch = (h->nq + K - 1) / K;
    const int ldfa = nch * K;
    std::vector<ChunkConst> cc(nch);
    for (int c = 0; c < nch; ++c) cc[c] = chunk_const(h, c * K);
    std::vector<float> ft((size_t)std::max(1, h->n_class) * ldfa, 0.0f);
    for (int c = 0; c < h->n_class; ++c)
        for (int k = 0; k < h->nq; ++k) ft[(size_t)c * ldfa + k] = (float)h->f64[(size_t)c * h->nq + k];
    std::vector<double> invq(ldfa, 0.0);
    std::vector<float> invq32(ldfa, 0.0f);
    for (int k = 0; k < h->nq; ++k) {
        invq[k] = h->q[k] != 0.0f ? 1.0 / (double)h->q[k] : 0.0;
        invq32[k] = (float)invq[k];
    }
    // triangle jobs over the slot-ordered atoms; remember each job's (row block, range)
    // (per-lane form measured: 128 <= 256 < 512 at 10k x 200)
    const int L = adjoint_col_range(h, pair, std::ceil(n / 64.0) * std::ceil(n / 128.0) * 0.5 * nch);
    std::vector<PairJob> jobs;
    triangle_jobs(jobs, h->d_cols.p, n, L, 1.0f);
    for (size_t j = 0; j < jobs.size(); ++j) jobs[j].pad = (int)j;
    const int T = (n + 63) / 64;
    const int stride = (n + L - 1) / L;
    std::vector<int> orig_block(jobs.size()), orig_m(jobs.size());
    {
        size_t j = 0;
        for (int I = 0; I < T; ++I)
            for (int c0 = I * 64, m = 0; c0 < n; c0 += L, ++m, ++j) {
                orig_block[j] = I;
                orig_m[j] = m;
            }
    }
    finalize_jobs(jobs);
    std::vector<int> tri((size_t)T * stride, 0);
    for (size_t s = 0; s < jobs.size(); ++s) {
        const int o = jobs[s].pad;
        tri[(size_t)orig_block[o] * stride + orig_m[o]] = (int)s;
    }
    const int nj = (int)jobs.size();
    h->a_ftab.ensure(ft.size());
    h->a_cb.ensure(ft.size());
    h->a_cg.ensure(ft.size());
    h->a_chunks.ensure(nch);
    h->a_invq.ensure(ldfa);
    h->a_invq32.ensure(ldfa);
    h->d_wq.ensure(ldfa);
    h->a_jobs.ensure(std::max(1, nj));
    h->a_tri.ensure(tri.size());
    h->a_rowpart.ensure((size_t)std::max(1, nj) * nch * 64 * 3);
    h->a_colpart.ensure((size_t)std::max(1, nj) * nch * L * 3);
    h->d_grad.ensure((size_t)n * 3);
    if (pair) h->a_coef.ensure((size_t)h->n_class * h->n_class * nch * rsx::adjoint_pair_stride(K));
    hipchk(hipMemcpy(h->a_ftab.p, ft.data(), ft.size() * sizeof(float), hipMemcpyHostToDevice), "upload ftab");
    hipchk(hipMemcpy(h->a_chunks.p, cc.data(), nch * sizeof(ChunkConst), hipMemcpyHostToDevice), "upload chunks");
    hipchk(hipMemcpy(h->a_invq.p, invq.data(), ldfa * sizeof(double), hipMemcpyHostToDevice), "upload invq");
    hipchk(hipMemcpy(h->a_invq32.p, invq32.data(), ldfa * sizeof(float), hipMemcpyHostToDevice), "upload invq32");
    hipchk(hipMemcpy(h->a_jobs.p, jobs.data(), nj * sizeof(PairJob), hipMemcpyHostToDevice), "upload jobs");
    hipchk(hipMemcpy(h->a_tri.p, tri.data(), tri.size() * sizeof(int), hipMemcpyHostToDevice), "upload tri");
    h->a_K = K;
    h->a_nch = nch;
    h->a_ldf = ldfa;
    h->a_L = L;
    h->a_njobs = nj;
    h->a_stride = stride;
    h->a_n = n;
    h->a_stamp = stamp;
    h->a_cols = h->d_cols.p;
    h->a_fp32 = fp32;
    h->a_pair = pair;
}

// adjoint kernel + reduction on the weights already in d_wq; grad [n][3] (caller order) -> grad
// (d_grad when null)
void launch_adjoint_pass(rsx_engine* h, float* grad = nullptr)
{
    rsx::AdjointArgs a;
    a.jobs = h->a_jobs.p;
    a.n_chunks = h->a_nch;
    a.ftab = h->a_ftab.p;
    a.ldf = h->a_ldf;
    a.chunks = h->a_chunks.p;
    a.half_dq = 0.5 * h->dq;
    a.wq = h->d_wq.p;
    a.invq = h->a_invq.p;
    a.invq32 = h->a_invq32.p;
    a.cb = h->a_cb.p;
    a.cg = h->a_cg.p;
    a.dq_rev = h->dq / (2.0 * PI);
    a.fp32 = h->a_fp32 ? 1 : 0;
    a.sin_terms = rsx::sin_series_terms(h->dq * h->path_radius);
    a.col_range = h->a_L;
    a.row_partial = h->a_rowpart.p;
    a.col_partial = h->a_colpart.p;
    a.n_class = h->n_class;
    if (h->a_fp32)
        hipchk(rsx::launch_adjoint_coef(h->a_ftab.p, h->d_wq.p, h->a_invq.p, std::max(1, h->n_class), h->a_ldf, h->a_cb.p,
                                        h->a_cg.p, h->stream),
               "adjoint coefficients");
    if (h->a_pair) {
        rsx::AdjointPairCoefArgs pc;
        pc.ftab = h->a_ftab.p;
        pc.cb = h->a_cb.p;
        pc.cg = h->a_cg.p;
        pc.wframe = nullptr;
        pc.groups = 1;
        pc.n_class = h->n_class;
        pc.n_chunks = h->a_nch;
        pc.K = h->a_K;
        pc.ldf = h->a_ldf;
        pc.coef = h->a_coef.p;
        hipchk(rsx::launch_adjoint_pair_coef(pc, h->stream), "adjoint class-pair table");
        a.coef = h->a_coef.p;
        a.coef_grouped = 0;
    }
    hipchk(rsx::launch_adjoint(a, h->a_njobs, h->a_K, h->stream), "adjoint kernel");
    rsx::AdjointReduceArgs r;
    r.row_partial = h->a_rowpart.p;
    r.col_partial = h->a_colpart.p;
    r.tri_slot = h->a_tri.p;
    r.tri_stride = h->a_stride;
    r.n = h->n;
    r.n_chunks = h->a_nch;
    r.col_range = h->a_L;
    r.slot_to_caller = h->d_s2c.p;
    r.grad = grad ? grad : h->d_grad.p;
    hipchk(rsx::launch_adjoint_reduce(r, h->stream), "adjoint reduce");
}

// queue the adjoint for weights wk (host) and leave grad [n][3] (caller order) in d_grad
void enqueue_adjoint(rsx_engine* h, const std::vector<double>& wk)
{
    adjoint_layout(h);
    std::vector<double> wq(h->a_ldf, 0.0);
    for (int k = 0; k < h->nq; ++k) wq[k] = h->q[k] == 0.0f ? 0.0 : wk[k];
    hipchk(hipMemcpyAsync(h->d_wq.p, wq.data(), h->a_ldf * sizeof(double), hipMemcpyHostToDevice, h->stream),
           "upload w");
    launch_adjoint_pass(h);
    h->a_have_w = true;
}

void adjoint_impl(rsx_engine* h, const std::vector<double>& wk, float* out_grad)
{
    set_device(h);
    enqueue_adjoint(h, wk);
    hipchk(hipMemcpyAsync(out_grad, h->d_grad.p, (size_t)h->n * 3 * sizeof(float), hipMemcpyDeviceToHost, h->stream),
           "download grad");
    hipchk(hipStreamSynchronize(h->stream), "sync adjoint");
}

// Device-resident bodies -> class records in slot order (d_cols, and d_s2c when sorted on
// the device), with the exposed-dot counts taken on the device for hydration layouts and the
// bodies' checks queued for rsx_synchronize.  No host synchronisation.
void enqueue_device_columns(rsx_engine* h, const float4* bodies_dev)
{
    const int n = h->n;
    const uint32_t* counts = nullptr;
    if (h->hydrate) {
        h->d_sasa_atoms.ensure(n);
        h->sasa_mirror = false;
        hipchk(hipMemcpyAsync(h->d_sasa_atoms.p, bodies_dev, (size_t)n * 16, hipMemcpyDeviceToDevice, h->stream),
               "copy atoms");
        run_sasa_device(h, n);
        counts = h->d_counts.p;
    }
    // checks of the device-resident bodies (read back by rsx_synchronize): types inside the
    // factor table, and positions inside the radius the layout's recurrence was chosen for
    // (its series covers 1.25 times that radius; the step angle dq r stays below 3)
    if (!h->dev_pending) {
        h->d_guard.ensure(1);
        hipchk(rsx::launch_fill_i32(h->d_guard.p, 0, 1, h->stream), "clear guard");
    }
    h->dev_pending = true;
    rsx::GatherGuard g;
    g.n_types = h->n_types;
    g.cx = (float)h->centre[0];
    g.cy = (float)h->centre[1];
    g.cz = (float)h->centre[2];
    const double r_ok = h->uniform ? std::min(1.25 * h->path_radius, 0.5 * MAX_STEP_ANGLE / h->dq) : 1e30;
    g.r2_max = (float)std::min(r_ok * r_ok, 1e30);
    g.flag = h->d_guard.p;
    h->host_stale = true;  // columns (and with hydration the slot map) now hold these bodies
    if (counts && h->n_class + 1 <= rsx::CLASS_SORT_MAX_KEYS) {
        // new positions change exposed-dot counts and so classes: the columns are sorted by
        // class again (stable, the host layout's order), so the pair kernel keeps long runs
        rsx::ClassSortArgs cs;
        cs.bodies = bodies_dev;
        cs.counts = counts;
        cs.d1 = h->d1;
        cs.n = n;
        cs.n_types = h->n_types;
        cs.n_class = h->n_class;
        h->d_cs_table.ensure((size_t)((n + 1023) / 1024) * (h->n_class + 1));
        h->d_cs_tot.ensure(h->n_class + 1);
        cs.table = h->d_cs_table.p;
        cs.tot = h->d_cs_tot.p;
        cs.cols = h->d_cols.p;
        cs.s2c = h->d_s2c.p;
        cs.g = g;
        hipchk(rsx::launch_class_sort(cs, h->stream), "class sort");
    } else {
        hipchk(rsx::launch_gather_cols(h->d_cols.p, bodies_dev, h->d_s2c.p, counts, h->d1, n, g, h->stream),
               "gather");
    }
}

// Device-resident forward step: (SASA) -> class records in slot order -> pair kernel ->
// fixed-order reduction into out [n_q].  No host synchronisation.
void enqueue_device_forward(rsx_engine* h, const float4* bodies_dev, double* out)
{
    const int n = h->n;
    enqueue_device_columns(h, bodies_dev);
    hipchk(rsx::launch_q0(h->d_cols.p, n, h->d_f0col.p, h->d_q0.p, h->stream), "q0 kernel");
    const int nj = (int)h->last_jobs.size();
    h->d_red.ensure(std::max<size_t>(1, rsx::reduce_scratch_doubles(h->nq, 1, nj)));  // sized for this table
    const rsx::DebyeArgs a = debye_args(h, h->d_fjobs.p, h->d_partials.p);
    hipchk(rsx::launch_debye(a, nj, h->K, h->uniform, h->stream), "debye kernel");
    hipchk(rsx::launch_reduce(h->d_partials.p, h->ldf, h->nq, h->d_slot_begin_f.p, 1, out, h->nq, 0, h->d_q0.p,
                              h->q0_index, nj, h->d_red.p, h->stream),
           "reduce kernel");
}

// ---- autotuner -------------------------------------------------------------
// Key of a tuning entry: device architecture, path, n_q, size bucket, hydration.
std::string tune_key(rsx_engine* h)
{
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) arch = prop.gcnArchName;
    for (char& c : arch)
        if (c == ' ' || c == ':') c = '_';
    int bucket = 0;
    while ((1 << (bucket + 1)) <= h->n) ++bucket;
    std::ostringstream k;
    k << arch << ":" << (h->uniform ? "rec" : "direct") << ":nq" << h->nq << ":n2^" << bucket << ":"
      << (h->hydrate ? "hyd" : "plain");
    return k.str();
}

// switch K / column range on a laid-out engine and rebuild the forward jobs
void retune(rsx_engine* h, int K, int L)
{
    h->k_req = K;
    h->range_req = L;
    apply_path(h, h->uniform);
    build_class_table(h);
    h->params_dirty = false;
    const std::vector<PairJob> jobs = forward_jobs(h);
    const int nj = (int)jobs.size();
    h->d_fjobs.ensure(nj);
    h->d_partials.ensure((size_t)std::max(nj, 1) * h->ldf);
    hipchk(hipMemcpy(h->d_fjobs.p, jobs.data(), nj * sizeof(PairJob), hipMemcpyHostToDevice), "upload jobs");
    h->last_jobs = jobs;
    set_forward_slots(h, nj);
}

float time_pair_kernel(rsx_engine* h, int reps)
{
    const rsx::DebyeArgs a = debye_args(h, h->d_fjobs.p, h->d_partials.p);
    const int nj = (int)h->last_jobs.size();
    hipchk(rsx::launch_debye(a, nj, h->K, h->uniform, h->stream), "warm-up");
    hipchk(hipEventRecord(h->ev0, h->stream), "event");
    for (int r = 0; r < reps; ++r) hipchk(rsx::launch_debye(a, nj, h->K, h->uniform, h->stream), "debye kernel");
    hipchk(hipEventRecord(h->ev1, h->stream), "event");
    hipchk(hipEventSynchronize(h->ev1), "event sync");
    float ms = 0.0f;
    hipchk(hipEventElapsedTime(&ms, h->ev0, h->ev1), "elapsed");
    return ms / reps;
}

// Look the current problem up in cache_path ("key K L ms" lines); if absent,
// time every (K, column range) candidate with HIP events, keep the fastest and
// append it.  The engine must hold a computed structure (calc_curve first).
void autotune_impl(rsx_engine* h, const char* cache_path)
{
    if (h->state != 4) fail(RSX_NO_INITIAL_CALC, "autotune needs a prior calc_curve");
    set_device(h);
    h->k_max = 0;  // the tuner chooses among all K; a later gradient call caps it again
    const std::string key = tune_key(h);
    if (cache_path && *cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            int K = 0, L = 0, V = 1;
            if ((ls >> k >> K >> L >> V) && k == key && V >= 0 && V < rsx::VARIANT_COUNT) {
                h->variant = V;
                retune(h, K, L);
                full_forward(h);
                return;
            }
        }
    }
    const int Ls[] = {256, 512, 768, 1024, 1536, 2048, 3072, 4096};
    struct Cand {
        float ms;
        int K, L, V;
    };
    std::vector<Cand> cands;
    // uniform grids: the two FP32 recurrence kernels (the FP64 ones stay selectable by
    // rsx_set_variant but are slower on every layout measured); the direct path has one
    const int uni_variants[] = {rsx::VARIANT_FP32R, rsx::VARIANT_FP32};
    const int n_var = h->uniform ? 2 : 1;
    for (int vi = 0; vi < n_var; ++vi) {
        if (h->uniform) h->variant = uni_variants[vi];
        const int path = pair_path(h);
        std::vector<int> Ks;
        for (int i = 0; i < rsx::debye_k_count(path); ++i) {
            const int k = rsx::debye_k_value(path, i);
            if (k >= 16 && !(k == 64 && !h->uniform)) Ks.push_back(k);
        }
        for (int K : Ks)
            for (int L : Ls) {
                if (L > std::max(256, h->n)) continue;
                retune(h, K, L);
                cands.push_back({time_pair_kernel(h, 3), K, L, h->variant});
            }
    }
    // the four fastest of the sweep are timed again, in turn, three times each (the
    // best of the three counts): run-to-run noise is ~1 %, about the spread among them
    std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
        return (a.V == rsx::VARIANT_FP32 ? a.ms * 1.03f : a.ms) < (b.V == rsx::VARIANT_FP32 ? b.ms * 1.03f : b.ms);
    });
    if (cands.size() > 4) cands.resize(4);
    for (auto& c : cands) c.ms = 1e30f;
    for (int round = 0; round < 3; ++round)
        for (auto& c : cands) {
            h->variant = c.V;
            retune(h, c.K, c.L);
            c.ms = std::min(c.ms, time_pair_kernel(h, 5));
        }
    // the per-lane-sum kernel keeps its run totals in FP32 per lane, which loses about
    // twice the accuracy of the run-flush kernel's FP64 row sums at deep minima of I(q)
    // (7e-6 against 3e-6 relative on a hydrated 3000-atom case): it has to be 3 % faster
    const auto eff = [](const Cand& c) { return c.V == rsx::VARIANT_FP32 ? c.ms * 1.03f : c.ms; };
    const Cand top = *std::min_element(cands.begin(), cands.end(),
                                       [&](const Cand& a, const Cand& b) { return eff(a) < eff(b); });
    const int bestK = top.K, bestL = top.L, bestV = top.V;
    const float best = top.ms;
    h->variant = bestV;
    retune(h, bestK, bestL);
    full_forward(h);
    if (cache_path && *cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << key << " " << bestK << " " << bestL << " " << bestV << " " << best << "\n";
    }
    h->tune_report = key + " K=" + std::to_string(bestK) + " L=" + std::to_string(bestL) + " ms=" + std::to_string(best);
}

// ---- batch autotuner --------------------------------------------------------------
// Key of a batch tuning entry: architecture, n_q, frame count and atom total (powers of
// two), class kind.
std::string batch_tune_key(rsx_engine* h)
{
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) arch = prop.gcnArchName;
    for (char& c : arch)
        if (c == ' ' || c == ':') c = '_';
    auto log2i = [](long long v) {
        int b = 0;
        while ((1LL << (b + 1)) <= v) ++b;
        return b;
    };
    std::ostringstream k;
    k << arch << ":batch:nq" << h->nq << ":f2^" << log2i(h->b_frames) << ":a2^" << log2i(h->b_offsets[h->b_frames])
      << ":" << (hydration_on(h) ? "hyd" : "plain");
    return k.str();
}

// average ms of `reps` replays of the batch graph (after one warm-up), HIP events
float time_batch(rsx_engine* h, int reps)
{
    batch_run_dev(h, h->b_out.p, h->b_mean.p);
    hipchk(hipEventRecord(h->ev0, h->stream), "event");
    for (int r = 0; r < reps; ++r) batch_run_dev(h, h->b_out.p, h->b_mean.p);
    hipchk(hipEventRecord(h->ev1, h->stream), "event");
    hipchk(hipEventSynchronize(h->ev1), "event sync");
    float ms = 0.0f;
    hipchk(hipEventElapsedTime(&ms, h->ev0, h->ev1), "elapsed");
    return ms / reps;
}

void set_batch_tuning(rsx_engine* h, int V, int K, int L, int P)
{
    h->b_var_req = V;
    h->b_k_req = K;
    h->b_range_req = L;
    h->b_pers_req = P;
    batch_invalidate(h);
}

// Look the batch up in cache_path ("key K L variant persistent ms" lines); if absent, time the
// FP32 kernels over K, column range and launch form on the laid-out frames, keep the
// fastest, append it.  Needs frames (rsx_batch_set_frames) on a uniform q grid.
void batch_autotune_impl(rsx_engine* h, const char* cache_path)
{
    if (h->b_frames <= 0) fail(RSX_ARGS_NOT_SET, "batch autotune needs frames");
    set_device(h);
    const std::string key = batch_tune_key(h);
    if (cache_path && *cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            int K = 0, L = 0, V = 0, P = 0;
            if ((ls >> k >> K >> L >> V >> P) && k == key && V >= 0 && V < rsx::VARIANT_COUNT) {
                set_batch_tuning(h, V, K, L, P);
                return;
            }
        }
    }
    if (h->b_layout_ok && !h->b_uniform) fail(RSX_NOT_SUPPORTED, "batch autotune needs a uniform q grid");
    h->b_out.ensure((size_t)h->b_frames * h->nq);
    h->b_mean.ensure(h->nq);
    struct Cand {
        float ms;
        int V, K, L, P;
    };
    std::vector<Cand> cands;
    const int r32r[] = {96, 100, 128}, r32[] = {64, 86};
    const int Ls[] = {256, 512, 1024};
    for (int K : r32r)
        for (int L : Ls)
            for (int P = 0; P < 2; ++P) cands.push_back({0.0f, rsx::VARIANT_FP32R, K, L, P});
    for (int K : r32)
        for (int P = 0; P < 2; ++P) cands.push_back({0.0f, rsx::VARIANT_FP32, K, 512, P});
    for (auto& c : cands) {
        set_batch_tuning(h, c.V, c.K, c.L, c.P);
        c.ms = time_batch(h, 2);
    }
    // the per-lane-sum kernel keeps its run totals in FP32 per lane (about twice the error of
    // the run-flush kernel's FP64 row sums at deep minima of I(q)): it has to be 3 % faster
    const auto eff = [](const Cand& c) { return c.V == rsx::VARIANT_FP32 ? c.ms * 1.03f : c.ms; };
    std::sort(cands.begin(), cands.end(), [&](const Cand& a, const Cand& b) { return eff(a) < eff(b); });
    if (cands.size() > 3) cands.resize(3);
    for (auto& c : cands) c.ms = 1e30f;
    for (int round = 0; round < 2; ++round)
        for (auto& c : cands) {
            set_batch_tuning(h, c.V, c.K, c.L, c.P);
            c.ms = std::min(c.ms, time_batch(h, 3));
        }
    const Cand top = *std::min_element(cands.begin(), cands.end(),
                                       [&](const Cand& a, const Cand& b) { return eff(a) < eff(b); });
    set_batch_tuning(h, top.V, top.K, top.L, top.P);
    if (cache_path && *cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << key << " " << top.K << " " << top.L << " " << top.V << " " << top.P << " " << top.ms << "\n";
    }
}

template <typename F>
int guarded(rsx_engine* h, F&& f)
{
    if (!h) return RSX_INVALID_ARG;
    try {
        f();
        h->err.clear();
        return RSX_OK;
    } catch (const Fail& e) {
        h->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        h->err = e.what();
        return RSX_HIP_ERROR;
    }
}

// The chi^2 weights w_k = -2 c (I_exp - c I) / sigma^2 amplify forward errors by
// I / sigma (~100x at 1 % errors).  The FP32 recurrence's error grows with its run
// length K, so the gradient entry points run the per-lane-sum FP32 kernel with runs
// of at most GRAD_K_MAX q (I(q) error ~5e-7, gradient error at 1 % noise no worse
// than with the FP64 recurrence); an engine on the run-flush FP32 kernel or on longer
// runs moves there (and stays there) on its first gradient call.  An engine set to
// an FP64 variant keeps it.
constexpr int GRAD_K_MAX = 51;

void use_gradient_forward(rsx_engine* h)
{
    if (h->nq <= 0 || (h->variant != rsx::VARIANT_FP32 && h->variant != rsx::VARIANT_FP32R)) return;
    if (h->variant == rsx::VARIANT_FP32 && h->k_max == GRAD_K_MAX && h->K <= GRAD_K_MAX) return;
    h->variant = rsx::VARIANT_FP32;
    h->k_req = 0;
    h->k_max = GRAD_K_MAX;
    setup_q(h);
    h->params_dirty = true;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int rsx_device_count(int* count)
{
    if (!count) return RSX_INVALID_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return RSX_OK;
}

int rsx_create(int device, rsx_engine** out)
{
    if (!out) return RSX_INVALID_ARG;
    *out = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || device < 0 || device >= c) return RSX_INVALID_DEVICE;
    rsx_engine* h = new rsx_engine();
    h->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
        hipEventCreate(&h->ev2) != hipSuccess || hipEventCreate(&h->ev3) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_sasa, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_skip, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_new, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming) != hipSuccess) {
        delete h;
        return RSX_HIP_ERROR;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        h->n_cu = prop.multiProcessorCount;
    try {
        h->d_queue.ensure(1);
    } catch (...) {
        delete h;
        return RSX_OUT_OF_MEMORY;
    }
    h->state = 1;
    *out = h;
    return RSX_OK;
}

int rsx_destroy(rsx_engine* h)
{
    if (!h) return RSX_INVALID_ARG;
    (void)hipSetDevice(h->device);
    batch_release_graph(h);
    release_local_graph(h);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->stream2) (void)hipStreamSynchronize(h->stream2);
    DBuf<float4>* f4[] = {&h->d_cols, &h->d_saved_cols, &h->d_sasa_atoms, &h->d_dots, &h->b_cols, &h->d_sorted_atoms,
                          &h->d_slab};
    for (auto* b : f4) b->release();
    h->d_qrev.release();
    DBuf<float>* ff[] = {&h->d_ftab,    &h->d_qv,       &h->d_r2,      &h->d_rext,    &h->d_grad,    &h->a_ftab,
                         &h->a_rowpart, &h->a_colpart,  &h->a_invq32,  &h->a_cb,      &h->a_cg,      &h->b_qv,
                         &h->b_ftab,    &h->d_qsplit,   &h->b_qsplit,  &h->ba_ftab,   &h->ba_cb,     &h->ba_cg,
                         &h->ba_invq32, &h->ba_w,       &h->ba_rowpart, &h->ba_colpart, &h->ba_grad,
                         &h->a_coef,    &h->ba_coef};
    for (auto* b : ff) b->release();
    DBuf<double>* fd[] = {&h->d_invq, &h->d_partials, &h->d_iq, &h->d_saved_iq, &h->d_out, &h->d_wq, &h->d_q0,
                          &h->d_red, &h->b_partials, &h->b_out, &h->b_mean, &h->b_q0, &h->b_red, &h->b_invq,
                          &h->b_qrev, &h->d_partials_new, &h->ba_invq, &h->ba_ones};
    for (auto* b : fd) b->release();
    DBuf<int>* fi[] = {&h->d_slot_begin, &h->d_bucket_of, &h->d_bucket_count, &h->d_bucket_start, &h->d_bucket_fill,
                       &h->d_sorted, &h->d_s2c, &h->b_slot_begin, &h->d_local_list, &h->d_local, &h->d_tile_sum,
                       &h->d_slab_count, &h->d_slab_pos, &h->ba_tri, &h->ba_tri_off, &h->ba_atom_frame,
                       &h->ba_frame_off, &h->ba_s2c};
    for (auto* b : fi) b->release();
    h->d_chunks.release();
    h->b_chunks.release();
    h->ba_chunks.release();
    h->ba_jobs.release();
    h->b_pin_ftab.release();
    h->b_pin_q0.release();
    for (rsx_engine::Staging* st : {&h->st_old, &h->st_new}) {
        st->host.release();
        st->dev.release();
    }
    for (auto& b : h->pin_jobs) b.release();
    h->d_lstage.release();
    h->pin_lstage.release();
    h->pin_moved.release();
    h->pin_local.release();
    h->pin_iq.release();
    h->a_chunks.release();
    h->a_invq.release();
    h->a_jobs.release();
    h->a_tri.release();
    h->h_u32.release();
    h->h_u64.release();
    h->h_dbl.release();
    h->h_out.release();
    h->h_part.release();
    h->h_q.release();
    h->h_pidx.release();
    h->h_tiles.release();
    for (auto& t : h->inc_tab) t.jobs.release();
    h->d_fjobs.release();
    for (auto& t : h->inc_tab) t.njobs = 0;
    h->d_f0col.release();
    h->d_slot_begin_f.release();
    h->b_jobs.release();
    h->d_counts.release();
    h->d_queue.release();
    h->d_guard.release();
    h->d_cs_table.release();
    h->d_cs_tot.release();
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->ev2) (void)hipEventDestroy(h->ev2);
    if (h->ev3) (void)hipEventDestroy(h->ev3);
    if (h->ev_sasa) (void)hipEventDestroy(h->ev_sasa);
    if (h->ev_skip) (void)hipEventDestroy(h->ev_skip);
    if (h->ev_new) (void)hipEventDestroy(h->ev_new);
    if (h->ev_out) (void)hipEventDestroy(h->ev_out);
    if (h->stream2) (void)hipStreamDestroy(h->stream2);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return RSX_OK;
}

const char* rsx_last_error(const rsx_engine* h)
{
    return h ? h->err.c_str() : "null engine";
}

void* rsx_stream(rsx_engine* h)
{
    return h ? (void*)h->stream : nullptr;
}

int rsx_state(const rsx_engine* h)
{
    return h ? h->state : RSX_INVALID_ARG;
}

int rsx_set_q(rsx_engine* h, const float* q, int n_q)
{
    return guarded(h, [&] {
        if (!q || n_q <= 0) fail(RSX_INVALID_ARG, "q vector must not be empty");
        for (int k = 0; k < n_q; ++k)
            if (!(q[k] >= 0.0f) || !std::isfinite(q[k])) fail(RSX_INVALID_ARG, "q values must be finite and >= 0");
        set_device(h);
        h->q.assign(q, q + n_q);
        h->nq = n_q;
        setup_q(h);
        h->params_dirty = true;
        if (h->state > 3) h->state = 3;
        batch_invalidate(h);
    });
}

int rsx_set_factors(rsx_engine* h, const float* table, int n_types, int water_index)
{
    return guarded(h, [&] {
        if (h->nq <= 0) fail(RSX_ARGS_NOT_SET, "set q before the factor table");
        if (!table || n_types <= 0 || water_index >= n_types) fail(RSX_INVALID_ARG, "invalid factor table");
        h->table.assign(table, table + (size_t)h->nq * n_types);
        h->n_types = n_types;
        h->water = water_index;
        h->atomic_table = false;
        h->params_dirty = true;
        h->factor_stamp++;
    });
}

int rsx_set_atomic_factors(rsx_engine* h, float avg_radius)
{
    return guarded(h, [&] {
        if (h->nq <= 0) fail(RSX_ARGS_NOT_SET, "set q before the factor table");
        if (!(avg_radius > 0.0f) || !std::isfinite(avg_radius)) fail(RSX_INVALID_ARG, "mean radius must be > 0");
        h->table.assign((size_t)h->nq * ATOMIC_TYPES, 0.0f);
        atomic_ff(h->q.data(), h->nq, avg_radius, h->expansion, h->table.data());
        h->n_types = ATOMIC_TYPES;
        h->water = ATOMIC_OH2;
        h->atomic_table = true;
        h->avg_radius = avg_radius;
        h->params_dirty = true;
        h->factor_stamp++;
    });
}

int rsx_set_expansion_factor(rsx_engine* h, float c)
{
    return guarded(h, [&] {
        if (!(c > 0.0f) || !std::isfinite(c)) fail(RSX_INVALID_ARG, "expansion factor must be > 0");
        if (c == h->expansion) return;
        h->expansion = c;
        if (h->atomic_table) atomic_ff(h->q.data(), h->nq, h->avg_radius, c, h->table.data());
        h->params_dirty = true;
        h->factor_stamp++;
    });
}

int rsx_set_water_weight(rsx_engine* h, float w)
{
    return guarded(h, [&] {
        if (!std::isfinite(w)) fail(RSX_INVALID_ARG, "water weight must be finite");
        if (w != h->w) {
            h->w = w;
            h->params_dirty = true;
            h->w_dirty = true;
            h->factor_stamp++;
        }
    });
}

int rsx_set_no_rsasa_recalc(rsx_engine* h, int flag)
{
    return guarded(h, [&] { h->no_rsasa_recalc = flag != 0; });
}

int rsx_set_sasa_params(rsx_engine* h, const float* radii, int n_types, float probe, int kind, int n_half)
{
    return guarded(h, [&] {
        if (!radii || n_types <= 0 || !(probe >= 0.0f)) fail(RSX_INVALID_ARG, "invalid SASA parameters");
        int nd = 0;
        if (kind == RSX_DOTS_DODECA_ICOSA)
            nd = 32;
        else if (kind == RSX_DOTS_FIBONACCI) {
            if (n_half <= 0 || 2 * n_half > 512) fail(RSX_INVALID_ARG, "Fibonacci sphere: 1 <= n_half <= 256");
            nd = 2 * n_half;
        } else
            fail(RSX_INVALID_ARG, "unknown dot-sphere kind");
        set_device(h);
        h->radii.assign(radii, radii + n_types);
        h->probe = probe;
        h->dot_kind = kind;
        h->n_half = n_half;
        h->n_dots = nd;
        h->dots.assign((size_t)n_types * nd * 4, 0.0f);
        for (int t = 0; t < n_types; ++t) {
            float* d = &h->dots[(size_t)t * nd * 4];
            if (kind == RSX_DOTS_DODECA_ICOSA)
                dodeca_icosa_dots(radii[t], probe, d);
            else
                fibonacci_dots(n_half, radii[t], probe, d);
            // templates are offsets from the atom centre; w carries the radius
        }
        upload_sasa_tables(h);
        h->sasa_set = true;
        h->have_rsasa = false;
        h->params_dirty = true;
        h->sasa_stamp++;
    });
}

int rsx_set_rebaseline(rsx_engine* h, int every)
{
    return guarded(h, [&] {
        if (every < 0) fail(RSX_INVALID_ARG, "rebaseline period must be >= 0");
        h->rebase_every = every;
    });
}

int rsx_set_tuning(rsx_engine* h, int k, int col_range)
{
    return guarded(h, [&] {
        if (k < 0 || col_range < 0) fail(RSX_INVALID_ARG, "tuning values must be >= 0");
        if (col_range > 0 && col_range % 64 != 0) fail(RSX_INVALID_ARG, "col_range must be a multiple of 64");
        h->k_req = k;
        h->k_max = 0;
        h->range_req = col_range;
        if (h->nq > 0) {
            set_device(h);
            setup_q(h);
            h->params_dirty = true;
            if (h->state > 3) h->state = 3;
            batch_invalidate(h);
        }
    });
}

int rsx_initialize(rsx_engine* h, int n_bodies)
{
    return guarded(h, [&] {
        if (n_bodies < 0) fail(RSX_INVALID_ARG, "negative body count");
        if (h->nq <= 0 || h->n_types <= 0) fail(RSX_ARGS_NOT_SET, "q and factor table must be set before initialize");
        set_device(h);
        h->n = n_bodies;
        h->bodies.assign((size_t)n_bodies * 4, 0.0f);
        h->type.assign(n_bodies, 0);
        h->counts.assign(n_bodies, 0);
        h->rsasa.assign(n_bodies, 0.0f);
        mark_dirty_full(h);
        h->have_rsasa = false;
        h->xaction = false;
        h->params_dirty = true;
        h->d_cols.ensure(n_bodies);
        h->d_s2c.ensure(n_bodies);
        h->state = 3;
    });
}

int rsx_calc_curve(rsx_engine* h, const float* bodies, int n, float* out)
{
    return guarded(h, [&] { calc_curve_impl(h, bodies, n, out, nullptr); });
}

int rsx_calc_curve_f64(rsx_engine* h, const float* bodies, int n, double* out)
{
    return guarded(h, [&] { calc_curve_impl(h, bodies, n, nullptr, out); });
}

int rsx_recalc_curve(rsx_engine* h, const float* bodies, int n, int start, int len, float* out)
{
    return guarded(h, [&] { recalc_curve_impl(h, bodies, n, start, len, out, nullptr); });
}

int rsx_recalc_curve_f64(rsx_engine* h, const float* bodies, int n, int start, int len, double* out)
{
    return guarded(h, [&] { recalc_curve_impl(h, bodies, n, start, len, nullptr, out); });
}

int rsx_commit(rsx_engine* h)
{
    return guarded(h, [&] { commit_impl(h); });
}

int rsx_revert(rsx_engine* h)
{
    return guarded(h, [&] { revert_impl(h); });
}

int rsx_get_rsasa(rsx_engine* h, float* out, int n)
{
    return guarded(h, [&] {
        if (!out || n < h->n) fail(RSX_INVALID_ARG, "rsasa output too small");
        for (int i = 0; i < h->n; ++i) out[i] = h->have_rsasa ? h->rsasa[i] : 0.0f;
    });
}

int rsx_calc_curve_dev_async(rsx_engine* h, const void* bodies_dev, int n, double* out_dev)
{
    return guarded(h, [&] {
        check_args_set(h);
        if (n != h->n || !bodies_dev) fail(RSX_INVALID_ARG, "calc_curve_dev: body count mismatch");
        if (h->state != 4) fail(RSX_NO_INITIAL_CALC, "calc_curve_dev needs one host calc_curve to lay out the atoms");
        if (h->params_dirty)
            fail(RSX_ARGS_NOT_SET, "calc_curve_dev: parameters changed since the last calc_curve (call it first)");
        set_device(h);
        enqueue_device_forward(h, (const float4*)bodies_dev, out_dev ? out_dev : h->d_iq.p);
    });
}

int rsx_synchronize(rsx_engine* h)
{
    return guarded(h, [&] {
        hipchk(hipStreamSynchronize(h->stream), "synchronize");
        if (!h->dev_pending) return;
        h->dev_pending = false;
        int flag = 0;
        hipchk(hipMemcpy(&flag, h->d_guard.p, sizeof(int), hipMemcpyDeviceToHost), "read guard");
        if (flag & rsx::GUARD_TYPE)
            fail(RSX_INVALID_ARG, "calc_curve_dev: a body type index is outside the factor table (its pairs were skipped)");
        if (flag & rsx::GUARD_RADIUS)
            fail(RSX_INVALID_ARG, "calc_curve_dev: bodies moved beyond the radius the layout was chosen for "
                                  "(the result may be inaccurate; lay out again with calc_curve)");
    });
}

int rsx_calc_rsasa(rsx_engine* h, const float* bodies, int n, const float* radii, int n_types, float probe,
                   const float* dots, int n_dots, unsigned* out_counts, float* out_frac)
{
    return guarded(h, [&] { calc_rsasa_impl(h, bodies, n, radii, n_types, probe, dots, n_dots, out_counts, out_frac); });
}

int rsx_calc_curve_hist(rsx_engine* h, const float* bodies, int n, float bin_width, double* out)
{
    return guarded(h, [&] { calc_hist_impl(h, bodies, n, bin_width, out); });
}

int rsx_fibonacci_dots(int n_half, float radius, float probe, float* out)
{
    if (n_half <= 0 || !out) return RSX_INVALID_ARG;
    fibonacci_dots(n_half, radius, probe, out);
    return RSX_OK;
}

int rsx_dodeca_icosa_dots(float radius, float probe, float* out)
{
    if (!out) return RSX_INVALID_ARG;
    dodeca_icosa_dots(radius, probe, out);
    return RSX_OK;
}

int rsx_sphere_dots(float radius, float n_sect, float probe, float* out, int cap, int* n_dots)
{
    if (!n_dots || cap < 0 || (cap > 0 && !out) || !(n_sect >= 0.0f) || n_sect > 1e5f) return RSX_INVALID_ARG;
    *n_dots = sphere_dots(radius, n_sect, probe, out, cap);
    return RSX_OK;
}

int rsx_device_name(int device, char* buf, int len)
{
    if (!buf || len <= 0) return RSX_INVALID_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || device < 0 || device >= c) return RSX_INVALID_DEVICE;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return RSX_HIP_ERROR;
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs, %.0f GiB)", p.name, p.gcnArchName, p.multiProcessorCount,
                  (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    return RSX_OK;
}

int rsx_atomic_form_factors(const float* q, int n_q, float avg_radius, float expansion, float* out)
{
    if (!q || n_q <= 0 || !out) return RSX_INVALID_ARG;
    atomic_ff(q, n_q, avg_radius, expansion, out);
    return RSX_OK;
}

int rsx_atomic_radii(float* out16)
{
    if (!out16) return RSX_INVALID_ARG;
    atomic_radii(out16);
    return RSX_OK;
}

int rsx_batch_set_frames(rsx_engine* h, const float* bodies, const int* offsets, int n_frames)
{
    return guarded(h, [&] { batch_set_frames_impl(h, bodies, offsets, n_frames); });
}

int rsx_batch_calc_dev_async(rsx_engine* h, double* out_frames_dev, double* out_mean_dev)
{
    return guarded(h, [&] {
        double* of = out_frames_dev ? out_frames_dev : (h->b_out.ensure((size_t)h->b_frames * h->nq), h->b_out.p);
        double* om = out_mean_dev ? out_mean_dev : (h->b_mean.ensure(h->nq), h->b_mean.p);
        batch_run_dev(h, of, om);
    });
}

int rsx_batch_calc(rsx_engine* h, double* out_frames, double* out_mean)
{
    return guarded(h, [&] {
        if (h->b_frames <= 0) fail(RSX_ARGS_NOT_SET, "batch: no frames set");
        h->b_out.ensure((size_t)h->b_frames * h->nq);
        h->b_mean.ensure(h->nq);
        batch_run_dev(h, h->b_out.p, h->b_mean.p);
        if (out_frames)
            hipchk(hipMemcpyAsync(out_frames, h->b_out.p, (size_t)h->b_frames * h->nq * sizeof(double),
                                  hipMemcpyDeviceToHost, h->stream),
                   "download frames");
        if (out_mean)
            hipchk(hipMemcpyAsync(out_mean, h->b_mean.p, h->nq * sizeof(double), hipMemcpyDeviceToHost, h->stream),
                   "download mean");
        hipchk(hipStreamSynchronize(h->stream), "sync batch");
    });
}

int rsx_batch_vjp(rsx_engine* h, const double* w, int n_frames, int n_q, float* out_grad)
{
    return guarded(h, [&] { batch_vjp_impl(h, w, n_frames, n_q, out_grad); });
}

int rsx_time_batch_vjp(rsx_engine* h, int reps, float* ms)
{
    return guarded(h, [&] {
        if (!h->ba_have_w || h->ba_gen != h->b_gen || h->ba_factor != h->b_factor_stamp)
            fail(RSX_NO_INITIAL_CALC, "time_batch_vjp needs a prior batch_vjp call on the current frames");
        if (reps <= 0 || !ms) fail(RSX_INVALID_ARG, "reps must be > 0");
        set_device(h);
        launch_batch_adjoint_pass(h);  // warm-up
        hipchk(hipEventRecord(h->ev0, h->stream), "event");
        for (int r = 0; r < reps; ++r) launch_batch_adjoint_pass(h);
        hipchk(hipEventRecord(h->ev1, h->stream), "event");
        hipchk(hipEventSynchronize(h->ev1), "event sync");
        float t = 0.0f;
        hipchk(hipEventElapsedTime(&t, h->ev0, h->ev1), "elapsed");
        *ms = t / reps;
    });
}

int rsx_chi2_grad(rsx_engine* h, const float* bodies, int n, const float* i_exp, const float* sigma, double* out_chi2,
                  double* out_scale, float* out_grad, double* out_iq)
{
    return guarded(h, [&] {
        if (!i_exp || !sigma) fail(RSX_INVALID_ARG, "chi2_grad: null profile");
        if (!h->uniform) fail(RSX_NOT_SUPPORTED, "chi2_grad needs a uniform q grid");
        use_gradient_forward(h);
        std::vector<double> I(h->nq);
        calc_curve_impl(h, bodies, n, nullptr, I.data());
        double num = 0.0, den = 0.0;
        for (int k = 0; k < h->nq; ++k) {
            const double p2 = 1.0 / ((double)sigma[k] * (double)sigma[k]);
            num += (double)i_exp[k] * I[k] * p2;
            den += I[k] * I[k] * p2;
        }
        const double c = num / den;
        double x2 = 0.0;
        std::vector<double> wk(h->nq);
        for (int k = 0; k < h->nq; ++k) {
            const double s2 = (double)sigma[k] * (double)sigma[k];
            const double d = (double)i_exp[k] - c * I[k];
            x2 += d * d / s2;
            wk[k] = -2.0 * c * d / s2;
        }
        if (out_chi2) *out_chi2 = x2;
        if (out_scale) *out_scale = c;
        if (out_iq) std::copy(I.begin(), I.end(), out_iq);
        if (out_grad) adjoint_impl(h, wk, out_grad);
    });
}

int rsx_debye_vjp(rsx_engine* h, const float* bodies, int n, const double* w, float* out_grad)
{
    return guarded(h, [&] {
        if (!w || !out_grad) fail(RSX_INVALID_ARG, "vjp: null argument");
        if (!h->uniform) fail(RSX_NOT_SUPPORTED, "vjp needs a uniform q grid");
        use_gradient_forward(h);
        calc_curve_impl(h, bodies, n, nullptr, nullptr);
        std::vector<double> wk(w, w + h->nq);
        adjoint_impl(h, wk, out_grad);
    });
}

// the engine's forward runs in the gradient configuration (that of rsx_chi2_grad: FP32 runs
// of at most GRAD_K_MAX q, or an FP64 recurrence the caller chose)
static bool gradient_forward_ready(const rsx_engine* h)
{
    if (h->variant == rsx::VARIANT_WAVE || h->variant == rsx::VARIANT_SHARED) return true;
    return h->variant == rsx::VARIANT_FP32 && h->k_max == GRAD_K_MAX && h->K <= GRAD_K_MAX;
}

static void check_device_gradient(rsx_engine* h, const void* bodies_dev, int n)
{
    check_args_set(h);
    if (n != h->n || !bodies_dev) fail(RSX_INVALID_ARG, "gradient_dev: body count mismatch");
    if (h->state != 4) fail(RSX_NO_INITIAL_CALC, "gradient_dev needs a host call to lay out the atoms");
    if (h->params_dirty)
        fail(RSX_ARGS_NOT_SET, "gradient_dev: parameters changed since the last host call (call it first)");
    if (!h->uniform) fail(RSX_NOT_SUPPORTED, "gradient_dev needs a uniform q grid");
}

int rsx_debye_vjp_dev_async(rsx_engine* h, const void* bodies_dev, int n, const double* w_dev, float* grad_dev)
{
    return guarded(h, [&] {
        if (!w_dev || !grad_dev) fail(RSX_INVALID_ARG, "vjp_dev: null argument");
        check_device_gradient(h, bodies_dev, n);
        set_device(h);
        adjoint_layout(h);
        enqueue_device_columns(h, (const float4*)bodies_dev);
        hipchk(rsx::launch_adjoint_weights(w_dev, h->nq, h->q0_index, h->a_ldf, h->d_wq.p, h->stream),
               "adjoint weights");
        launch_adjoint_pass(h, grad_dev);
        h->a_have_w = true;
    });
}

int rsx_chi2_grad_dev_async(rsx_engine* h, const void* bodies_dev, int n, const float* i_exp_dev,
                            const float* sigma_dev, double* out_dev, float* grad_dev, double* iq_dev)
{
    return guarded(h, [&] {
        if (!i_exp_dev || !sigma_dev || !out_dev) fail(RSX_INVALID_ARG, "chi2_grad_dev: null argument");
        check_device_gradient(h, bodies_dev, n);
        if (!gradient_forward_ready(h))
            fail(RSX_ARGS_NOT_SET, "chi2_grad_dev: lay the engine out for gradients with a host rsx_chi2_grad first");
        set_device(h);
        adjoint_layout(h);
        double* iq = iq_dev ? iq_dev : h->d_out.p;
        enqueue_device_forward(h, (const float4*)bodies_dev, iq);
        hipchk(rsx::launch_chi2_weights(iq, i_exp_dev, sigma_dev, h->nq, h->q0_index, h->a_ldf, out_dev, h->d_wq.p,
                                        h->stream),
               "chi2 weights");
        if (grad_dev) launch_adjoint_pass(h, grad_dev);
        h->a_have_w = true;
    });
}

int rsx_batch_launch(rsx_engine* h, int* persistent_waves_out, int* n_jobs)
{
    return guarded(h, [&] {
        if (h->b_pers_waves < 0 || !h->b_exec) fail(RSX_NO_INITIAL_CALC, "batch_launch: no batch graph yet");
        if (persistent_waves_out) *persistent_waves_out = h->b_pers_waves;
        if (n_jobs) *n_jobs = h->b_njobs;
    });
}

int rsx_autotune(rsx_engine* h, const char* cache_path)
{
    return guarded(h, [&] { autotune_impl(h, cache_path); });
}

int rsx_batch_autotune(rsx_engine* h, const char* cache_path)
{
    return guarded(h, [&] { batch_autotune_impl(h, cache_path); });
}

int rsx_batch_get_tuning(rsx_engine* h, int* variant, int* k, int* col_range, int* waves)
{
    return guarded(h, [&] {
        if (h->b_frames <= 0) fail(RSX_ARGS_NOT_SET, "batch: no frames set");
        if (!h->b_layout_ok) {
            set_device(h);
            batch_layout(h);
        }
        if (variant) *variant = h->b_variant;
        if (k) *k = h->b_K;
        if (col_range) *col_range = h->b_L;
        if (waves) {
            const bool fp32 = h->b_variant == rsx::VARIANT_FP32 || h->b_variant == rsx::VARIANT_FP32R;
            const bool pers = h->b_uniform && fp32 && (h->b_pers_req >= 0 ? h->b_pers_req != 0 : h->persistent);
            *waves = pers ? persistent_waves(h, h->b_variant, h->b_K, h->b_radius) : 0;
        }
    });
}

int rsx_set_variant(rsx_engine* h, int variant)
{
    return guarded(h, [&] {
        if (variant < 0 || variant >= rsx::VARIANT_COUNT) fail(RSX_INVALID_ARG, "variant must be 0, 1, 2 or 3");
        h->k_max = 0;
        if (variant == h->variant) return;
        h->variant = variant;
        if (h->nq > 0) {  // the K list may differ: lay out again on the next calc_curve
            set_device(h);
            setup_q(h);
            h->params_dirty = true;
            if (h->state > 3) h->state = 3;
            batch_invalidate(h);
        }
    });
}

int rsx_get_tuning(rsx_engine* h, int* k, int* col_range, int* recurrence, int* variant)
{
    return guarded(h, [&] {
        if (k) *k = h->K;
        if (col_range) *col_range = h->last_L;
        if (recurrence) *recurrence = h->uniform ? 1 : 0;
        if (variant) *variant = h->variant;
    });
}

int rsx_time_forward(rsx_engine* h, int reps, float* ms_total, float* ms_pair)
{
    return guarded(h, [&] {
        if (h->state != 4) fail(RSX_NO_INITIAL_CALC, "time_forward needs a prior calc_curve");
        if (reps <= 0) fail(RSX_INVALID_ARG, "reps must be > 0");
        set_device(h);
        const int nj = (int)h->last_jobs.size();
        h->d_red.ensure(std::max<size_t>(1, rsx::reduce_scratch_doubles(h->nq, 1, nj)));
        hipchk(hipMemcpy(h->d_q0.p, &h->last_q0, sizeof(double), hipMemcpyHostToDevice), "upload q0");
        const rsx::DebyeArgs a = debye_args(h, h->d_fjobs.p, h->d_partials.p);
        float pair_ms = 0.0f;
        hipchk(hipEventRecord(h->ev0, h->stream), "event");
        for (int r = 0; r < reps; ++r) {
            hipchk(hipEventRecord(h->ev2, h->stream), "event");
            hipchk(rsx::launch_debye(a, nj, h->K, h->uniform, h->stream), "debye kernel");
            hipchk(hipEventRecord(h->ev3, h->stream), "event");
            hipchk(rsx::launch_reduce(h->d_partials.p, h->ldf, h->nq, h->d_slot_begin_f.p, 1, h->d_iq.p, h->nq, 0,
                                      h->d_q0.p, h->q0_index, nj, h->d_red.p, h->stream),
                   "reduce kernel");
            hipchk(hipEventSynchronize(h->ev3), "event sync");
            float t = 0.0f;
            hipchk(hipEventElapsedTime(&t, h->ev2, h->ev3), "elapsed");
            pair_ms += t;
        }
        hipchk(hipEventRecord(h->ev1, h->stream), "event");
        hipchk(hipEventSynchronize(h->ev1), "event sync");
        float tot = 0.0f;
        hipchk(hipEventElapsedTime(&tot, h->ev0, h->ev1), "elapsed");
        if (ms_total) *ms_total = tot / reps;
        if (ms_pair) *ms_pair = pair_ms / reps;
    });
}

int rsx_set_persistent(rsx_engine* h, int on)
{
    return guarded(h, [&] {
        h->persistent = on != 0;
        batch_release_graph(h);  // the captured launch changes
    });
}

int rsx_set_adjoint_tuning(rsx_engine* h, int fp32, int k, int col_range)
{
    return guarded(h, [&] {
        if (k < 0 || col_range < 0 || (col_range > 0 && col_range % 64 != 0))
            fail(RSX_INVALID_ARG, "adjoint tuning: k >= 0, col_range a multiple of 64");
        if (fp32 < 0 || fp32 > 2) fail(RSX_INVALID_ARG, "adjoint tuning: fp32 is 0, 1 or 2");
        if (k > 0) {
            bool ok = false;
            for (int i = 0; i < rsx::adjoint_k_count(fp32 != 0); ++i) ok = ok || rsx::adjoint_k_value(fp32 != 0, i) == k;
            if (fp32 == 1)
                for (int i = 0; i < rsx::adjoint_pair_k_count(); ++i) ok = ok || rsx::adjoint_pair_k_value(i) == k;
            if (!ok) fail(RSX_INVALID_ARG, "unsupported adjoint K for this precision");
        }
        h->adj_fp32 = fp32 != 0;
        h->adj_pair_ok = fp32 == 1;
        h->adj_k_req = k;
        h->adj_range_req = col_range;
        h->a_stamp = ~0ull;  // lay out again on the next gradient call
        h->ba_K = 0;         // and on the next batch vjp
        h->ba_have_w = false;
    });
}

int rsx_get_adjoint_layout(rsx_engine* h, int batch, int* form, int* k, int* col_range)
{
    return guarded(h, [&] {
        int f, kk, L;
        if (batch) {
            if (h->ba_K <= 0) fail(RSX_NO_INITIAL_CALC, "no batch vjp layout yet");
            f = h->ba_pair ? RSX_ADJOINT_PAIR : RSX_ADJOINT_PER_LANE;
            kk = h->ba_K;
            L = h->ba_L;
        } else {
            if (h->a_stamp == ~0ull || h->a_K <= 0) fail(RSX_NO_INITIAL_CALC, "no adjoint layout yet");
            f = !h->a_fp32 ? RSX_ADJOINT_FP64 : h->a_pair ? RSX_ADJOINT_PAIR : RSX_ADJOINT_PER_LANE;
            kk = h->a_K;
            L = h->a_L;
        }
        if (form) *form = f;
        if (k) *k = kk;
        if (col_range) *col_range = L;
    });
}

int rsx_time_adjoint(rsx_engine* h, int reps, float* ms)
{
    return guarded(h, [&] {
        if (!h->a_have_w || h->a_n != h->n || h->state != 4)
            fail(RSX_NO_INITIAL_CALC, "time_adjoint needs a prior chi2_grad or vjp call");
        if (reps <= 0 || !ms) fail(RSX_INVALID_ARG, "reps must be > 0");
        set_device(h);
        launch_adjoint_pass(h);  // warm-up
        hipchk(hipEventRecord(h->ev0, h->stream), "event");
        for (int r = 0; r < reps; ++r) launch_adjoint_pass(h);
        hipchk(hipEventRecord(h->ev1, h->stream), "event");
        hipchk(hipEventSynchronize(h->ev1), "event sync");
        float t = 0.0f;
        hipchk(hipEventElapsedTime(&t, h->ev0, h->ev1), "elapsed");
        *ms = t / reps;
    });
}

int rsx_time_sasa(rsx_engine* h, int reps, float* ms)
{
    return guarded(h, [&] {
        if (reps <= 0) fail(RSX_INVALID_ARG, "time_sasa: reps > 0");
        if (!h->hydrate || !h->have_rsasa || !h->d_sasa_atoms.p || h->d_sasa_atoms.cap < (size_t)h->n)
            fail(RSX_NO_INITIAL_CALC, "time_sasa needs a prior hydrated calc_curve");
        set_device(h);
        run_sasa_device(h, h->n);
        hipchk(hipEventRecord(h->ev0, h->stream), "event");
        for (int r = 0; r < reps; ++r) run_sasa_device(h, h->n);
        hipchk(hipEventRecord(h->ev1, h->stream), "event");
        hipchk(hipEventSynchronize(h->ev1), "event sync");
        float t = 0.0f;
        hipchk(hipEventElapsedTime(&t, h->ev0, h->ev1), "elapsed");
        if (ms) *ms = t / reps;
    });
}

int rsx_time_copy(rsx_engine* h, long long bytes, int reps, float* ms)
{
    return guarded(h, [&] {
        if (bytes < 16 || reps <= 0) fail(RSX_INVALID_ARG, "time_copy: bytes >= 16 and reps > 0");
        set_device(h);
        const long long n = bytes / 16;
        DBuf<float4> a, b;
        a.ensure(n);
        b.ensure(n);
        try {
            hipchk(rsx::launch_fill_i32((int*)a.p, 0, n * 4, h->stream), "fill");
            hipchk(rsx::launch_copy(b.p, a.p, n, h->stream), "copy warm-up");
            hipchk(hipEventRecord(h->ev0, h->stream), "event");
            for (int r = 0; r < reps; ++r) hipchk(rsx::launch_copy(b.p, a.p, n, h->stream), "copy");
            hipchk(hipEventRecord(h->ev1, h->stream), "event");
            hipchk(hipEventSynchronize(h->ev1), "event sync");
            float t = 0.0f;
            hipchk(hipEventElapsedTime(&t, h->ev0, h->ev1), "elapsed");
            if (ms) *ms = t / reps;
        } catch (...) {
            a.release();
            b.release();
            throw;
        }
        a.release();
        b.release();
    });
}

int rsx_dev_alloc(rsx_engine* h, void** ptr, long long bytes)
{
    return guarded(h, [&] {
        if (!ptr || bytes < 0) fail(RSX_INVALID_ARG, "dev_alloc");
        set_device(h);
        hipError_t e = hipMalloc(ptr, (size_t)std::max(bytes, 1LL));
        if (e != hipSuccess) fail(RSX_OUT_OF_MEMORY, hipGetErrorString(e));
    });
}

int rsx_dev_free(rsx_engine* h, void* ptr)
{
    return guarded(h, [&] {
        set_device(h);
        hipchk(hipFree(ptr), "hipFree");
    });
}

int rsx_memcpy_h2d(rsx_engine* h, void* dst, const void* src, long long bytes)
{
    return guarded(h, [&] {
        set_device(h);
        hipchk(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice), "memcpy h2d");
    });
}

int rsx_memcpy_d2h(rsx_engine* h, void* dst, const void* src, long long bytes)
{
    return guarded(h, [&] {
        set_device(h);
        hipchk(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost), "memcpy d2h");
    });
}

}  // extern "C"
