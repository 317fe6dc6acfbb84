// HIP kernels of the Debye-SAXS engine for gfx950 (CDNA4, wave64).
//
// Pair sums
// ---------
// I(q) = sum_i sum_j F_i(q) F_j(q) sin(q r_ij) / (q r_ij)           (Debye)
// The reference evaluates one sin + one divide per (pair, q) in 64x64 blocks,
// one q per NDRange slice (resaxslib/src/gpu_saxs_pt_wf.cl:96-181).  Here one
// wave owns 64 row atoms (one per lane) and a chunk of K consecutive q values
// held in registers, and walks a run of column atoms that every lane shares
// (the column record is wave-uniform, so it is fetched with scalar loads).
//
// Uniform q grids (q_k = a + k dq inside a chunk) use, per pair, the
// three-term recurrence of u_k = sin(q_k r) / r:
//     u_{k+1} = 2 cos(dq r) u_k - u_{k-1}
// in FP64 (full rate on CDNA4), so a (pair, q) term costs one FMA for the
// recurrence and one add into the accumulator.  The seeds are formed in
// FP32 with v_sin/v_cos: u_0 = sin(a r)/r and u_1 - u_0 = 2 cos((a+dq/2) r)
// sin(dq r/2) / r (product form, accurate for small dq r), and
// 2 cos(dq r) = 2 - 4 sin^2(dq r/2) in FP64.  Other grids take the direct
// path: one v_sin per term into FP32 accumulators.
//
// Columns are grouped by form-factor class (type, or type x exposed-dot
// count with hydration).  Within a run of one class the F_j(q) factor is
// common, so the run is summed unweighted and multiplied by F_class(q) when
// the class changes ("flush").  Row factors F_i(q) and 1/q are applied once
// per lane at the end, then the wave reduces its 64 lanes in FP64 and writes
// one row of the partial-sum table; a separate kernel sums the table in a
// fixed order (deterministic FP64 cross-block reduction).
//
// SASA
// ----
// Dot-sphere occlusion (resaxslib/src/rsasa_dot_spheres.cl:94-163): a dot of
// atom i is buried when (p + r_j)^2 - |dot - x_j|^2 > 1e-4 for any atom j.
// One wave per atom, dots spread over lanes, candidates from a hashed cell
// grid; the predicate is evaluated in FP32 without contraction (dot_occluded) so
// the integer counts equal the CPU restatement bit for bit.

#include "kernels.h"

#include <hip/hip_runtime.h>

namespace rsx {

namespace {

// sin / cos of 2 pi x for a phase x in revolutions held in FP64: the integer
// part is removed exactly in FP64, the fraction goes to v_sin / v_cos
// (which take revolutions; ~1.2e-7 absolute error measured on gfx950).
__device__ __forceinline__ float sin_rev(double x)
{
    return __builtin_amdgcn_sinf((float)__builtin_amdgcn_fract(x));
}

__device__ __forceinline__ float cos_rev(double x)
{
    return __builtin_amdgcn_cosf((float)__builtin_amdgcn_fract(x));
}

// Column records of a pair job are read through the constant address space: the index
// is uniform over the wave, so they come as scalar loads into scalar registers and the
// column loops are uniform control flow.  (Read as ordinary global loads, the loaded
// class bits made the loops divergent: vector loads whose latency each step waited for,
// and exec-mask bookkeeping; 3.5 % of the run-flush kernel's time.)
typedef float v4f __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) v4f cv4f;

struct ConstCols {
    cv4f* p;
    __device__ __forceinline__ float4 operator[](int j) const
    {
        const v4f v = p[j];
        return make_float4(v.x, v.y, v.z, v.w);
    }
};


__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// ---------------------------------------------------------------------------
// uniform-grid policy: FP64 accumulators, three-term recurrence
// ---------------------------------------------------------------------------
struct Seeds {
    double u0, u1, c2;
};

// sin(x) for small x >= 0 in FP64 from the odd Taylor series, cut after x^(2M+1)
// (M a template argument, so the Horner chain is unrolled and branch-free); the
// host picks M so that the first omitted term stays below 1e-16 relative for the
// layout's largest dq r / 2.  The recurrence's frequency 2 cos(dq r) =
// 2 - 4 sin^2(dq r / 2) must be far more accurate than v_sin's ~4e-7 absolute
// error: a relative error e in sin(dq r / 2) shifts every later term's phase by
// ~ e q r.
template <int M>
__device__ __forceinline__ double sin_series(double x)
{
    constexpr double c[9] = {1.0,
                             -1.6666666666666666e-01,
                             8.333333333333333e-03,
                             -1.984126984126984e-04,
                             2.7557319223985893e-06,
                             -2.505210838544172e-08,
                             1.6059043836821613e-10,
                             -7.647163731819816e-13,
                             2.8114572543455206e-15};
    static_assert(M >= 2 && M <= 8, "series length");
    const double x2 = x * x;
    double p = c[M];
#pragma unroll
    for (int n = M - 1; n >= 1; --n) p = fma(p, x2, c[n]);
    return x * fma(p, x2, 1.0);
}

__device__ __forceinline__ float pair_r2(float xi, float yi, float zi, const float4 c)
{
    const float dx = xi - c.x, dy = yi - c.y, dz = zi - c.z;
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// r and 1/r of a pair from r2 (both 0 for r2 <= 1e-30, the r == 0 pairs): v_rsq_f32, then
// one Newton step for each.  v_rsq_f32's error is a structured function of its argument
// (mean -7e-9, rms 3.8e-8 relative), so pairs at similar distances err alike and the errors
// add up coherently: unrefined, the direct kernel's I(q) at 20,000 atoms was off by 1.1e-5
// at a high-q minimum; refined (mean 2e-11, rms 2.5e-8), by 6.6e-7.
struct PairR {
    float r, rinv;
};
__device__ __forceinline__ PairR pair_r(float r2)
{
    const float nzf = r2 > 1e-30f ? 1.0f : 0.0f;
    const float y = __builtin_amdgcn_rsqf(fmaxf(r2, 1e-30f));
    const float r0 = r2 * y;
    PairR p;
    p.r = fmaf(fmaf(-r0, r0, r2), 0.5f * y, r0) * nzf;
    p.rinv = y * fmaf(-0.5f * r2 * y, y, 1.5f) * nzf;
    return p;
}

// Seeds of a pair, branch-free: every value stays finite and r == 0 (the i == j
// diagonal or coincident atoms) is blended in with a 0/1 mask (u_k -> w q_k, linear
// in k), so the seed computation of the next column pair can be scheduled into the
// recurrence block.  w is the pair weight (diagonal weights, 0 for padding).
template <int M>
__device__ __forceinline__ Seeds make_seeds(float r2, const ChunkConst& cc, double half_dq, float dq, float w)
{
    const float nzf = r2 > 1e-30f ? 1.0f : 0.0f;
    const PairR pr = pair_r(r2);
    const float rinv = pr.rinv;
    const double rd = (double)pr.r;
    const double sh = sin_series<M>(rd * half_dq);  // sin(dq r / 2)
    const float sa = sin_rev(cc.a_rev * rd);        // sin(a r)
    const float cm = cos_rev(cc.mid_rev * rd);      // cos((a + dq/2) r)
    const float wr = w * rinv;
    const float wz = w * (1.0f - nzf);
    Seeds s;
    s.c2 = fma(-4.0 * sh, sh, 2.0);
    s.u0 = (double)fmaf(sa, wr, wz * cc.a);
    s.u1 = fma((double)((cm + cm) * wr), sh, s.u0 + (double)(wz * dq));
    return s;
}

// the same from shared geometry (r, 1/r with 1/r = 0 for r == 0, sin(dq r / 2))
__device__ __forceinline__ Seeds chunk_seeds(float r, float rinv, double sh, const ChunkConst& cc, float dq, float w)
{
    const double rd = (double)r;
    const float sa = sin_rev(cc.a_rev * rd);
    const float cm = cos_rev(cc.mid_rev * rd);
    const float nzf = rinv > 0.0f ? 1.0f : 0.0f;
    const float wr = w * rinv;
    const float wz = w * (1.0f - nzf);
    Seeds s;
    s.c2 = fma(-4.0 * sh, sh, 2.0);
    s.u0 = (double)fmaf(sa, wr, wz * cc.a);
    s.u1 = fma((double)((cm + cm) * wr), sh, s.u0 + (double)(wz * dq));
    return s;
}

// two interleaved chains (two columns) added into the running sums
template <int K>
__device__ __forceinline__ void run2(double (&acc)[K], const Seeds& a, const Seeds& b)
{
    acc[0] += a.u0;
    acc[1] += a.u1;
    acc[0] += b.u0;
    acc[1] += b.u1;
    double ap = a.u0, ac = a.u1, bp = b.u0, bc = b.u1;
#pragma unroll
    for (int k = 2; k < K; ++k) {
        const double an = fma(a.c2, ac, -ap);
        const double bn = fma(b.c2, bc, -bp);
        acc[k] += an;
        acc[k] += bn;
        ap = ac;
        ac = an;
        bp = bc;
        bc = bn;
    }
}

// diagonal-tile weights: 1 above the diagonal (j > i), 1/2 on it, 0 below
__device__ __forceinline__ float diag_weight(int j, int lane)
{
    return j > lane ? 1.0f : (j == lane ? 0.5f : 0.0f);
}

template <int K, int M>
__global__ __launch_bounds__(64, 2) void debye_uniform_kernel(const DebyeArgs a)
{
    const int job_id = blockIdx.x / a.n_chunks;
    const int chunk = blockIdx.x - job_id * a.n_chunks;
    const PairJob jb = a.jobs[job_id];
    const int lane = threadIdx.x;
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const int k0 = chunk * K;
    const ChunkConst cc = a.chunks[chunk];
    const double h_rev = a.half_dq;
    const float dq = a.dq;
    const float* __restrict__ ftab = a.ftab;
    const int ldf = a.ldf;

    // acc: running FP64 sums of the current class run; X: per-lane sums over
    // flushed runs weighted by F_class (FP32)
    double acc[K];
    float X[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc[k] = 0.0;
        X[k] = 0.0f;
    }

    auto flush = [&](uint32_t cls) {
        // the class is wave-uniform: a scalar offset lets the row come in through
        // scalar loads instead of per-lane vector loads
        const float* __restrict__ fr = ftab + __builtin_amdgcn_readfirstlane((int)cls * ldf + k0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            X[k] = fmaf(fr[k], (float)acc[k], X[k]);
            acc[k] = 0.0;
        }
    };

    const ConstCols cols{(cv4f*)jb.cols};
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    // one weighted pair step per iteration: the second column is the next one when
    // it has the same class, otherwise a zero-weight copy.  Outside the diagonal tile
    // the weights are wave-uniform (0 for skipped records).
    uint32_t cur = CLS_SKIP;
    int j = 0;
    while (j < ncol) {
        const float4 c0 = cols[j];
        const uint32_t cl0 = __float_as_uint(c0.w);
        if (cl0 != cur) {
            if (cur != CLS_SKIP) flush(cur);
            cur = cl0;
        }
        const bool pair = j + 1 < ncol && __float_as_uint(cols[j + 1].w) == cl0;
        const float4 c1 = pair ? cols[j + 1] : c0;
        const float live = cl0 != CLS_SKIP ? 1.0f : 0.0f;
        float w0 = live, w1 = pair ? live : 0.0f;
        if (j < jd) {
            w0 *= diag_weight(j, lane);
            w1 *= diag_weight(j + 1, lane);
        }
        run2<K>(acc, make_seeds<M>(pair_r2(ri.x, ri.y, ri.z, c0), cc, h_rev, dq, w0),
                make_seeds<M>(pair_r2(ri.x, ri.y, ri.z, c1), cc, h_rev, dq, w1));
        j += pair ? 2 : 1;
    }
    if (cur != CLS_SKIP) flush(cur);

    // epilogue: v_ik = F_i(q_k) X_ik (the all-zero class row serves padding
    // lanes), transposed through LDS so lane k sums column k over the 64 rows in
    // FP64; then 1/q_k and the job scale.  Rows of T are padded to 65 words so
    // both the row writes and the column reads are free of bank conflicts.
    // K > 64 goes through the transpose in slices of 64 q
    constexpr int KS = K < 64 ? K : 64;
    __shared__ float T[KS][65];
    const uint32_t ci = cls_i == CLS_SKIP ? (uint32_t)a.zero_class : cls_i;
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
#pragma unroll
    for (int kb = 0; kb < K; kb += KS) {
#pragma unroll
        for (int k = 0; k < KS; ++k)
            if (kb + k < K) T[k][lane] = fi[kb + k] * X[kb + k];
        __syncthreads();
        if (lane < KS && kb + lane < K) {
            double sum = 0.0;
#pragma unroll 16
            for (int t = 0; t < 64; ++t) sum += (double)T[lane][t];
            a.partials[(size_t)jb.slot * ldf + k0 + kb + lane] = sum * a.invq[k0 + kb + lane] * (double)jb.scale;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// uniform-grid path, workgroup variant: the waves of one workgroup are the
// q chunks of one job (up to 4).  The pair geometry that does not depend on q
// (r, 1/r and the FP64 sin(dq r / 2)) is computed once per (row, column) for
// tiles of SH_TC columns, each wave doing every nw-th column, and shared through
// LDS.  Tiles are double-buffered: the geometry of tile t+1 is written while
// tile t is read, so one barrier per tile separates the phases.
// ---------------------------------------------------------------------------
constexpr int SH_TC = 32;   // columns per shared tile
constexpr int SH_MAXW = 4;  // waves (q chunks) per workgroup

template <int K, int M>
__global__ __launch_bounds__(64 * SH_MAXW, 2) void debye_shared_kernel(const DebyeArgs a)
{
    struct Tile {
        float r[SH_TC][64];
        float ri[SH_TC][64];  // 1/r, 0 for r == 0
        double sh[SH_TC][64];
    };
    union Shared {
        Tile t[2];
        float T[SH_MAXW][K][65];
    };
    __shared__ Shared sm;

    const int nw = blockDim.x >> 6;                 // waves in the workgroup
    const int ng = (a.n_chunks + nw - 1) / nw;      // workgroups per job
    const int job_id = blockIdx.x / ng;
    const int grp = blockIdx.x - job_id * ng;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int chunk = grp * nw + wave;
    const bool active = chunk < a.n_chunks;         // wave-uniform
    const PairJob jb = a.jobs[job_id];
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const int k0 = (active ? chunk : 0) * K;
    const ChunkConst cc = a.chunks[active ? chunk : 0];
    const double half_dq = a.half_dq;
    const float dq = a.dq;
    const float* __restrict__ ftab = a.ftab;
    const int ldf = a.ldf;

    double acc[K];
    float X[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc[k] = 0.0;
        X[k] = 0.0f;
    }
    auto flush = [&](uint32_t cls) {
        // the class is wave-uniform: a scalar offset lets the row come in through
        // scalar loads instead of per-lane vector loads
        const float* __restrict__ fr = ftab + __builtin_amdgcn_readfirstlane((int)cls * ldf + k0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            X[k] = fmaf(fr[k], (float)acc[k], X[k]);
            acc[k] = 0.0;
        }
    };

    const ConstCols cols{(cv4f*)jb.cols};
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    const int n_tiles = (ncol + SH_TC - 1) / SH_TC;

    // geometry of one tile into buffer b
    auto geometry = [&](int tile, Tile& b) {
        const int t0 = tile * SH_TC;
        const int tn = ncol - t0 < SH_TC ? ncol - t0 : SH_TC;
        for (int t = wave; t < tn; t += nw) {
            const PairR pr = pair_r(pair_r2(ri.x, ri.y, ri.z, cols[t0 + t]));
            b.r[t][lane] = pr.r;
            b.ri[t][lane] = pr.rinv;
            b.sh[t][lane] = sin_series<M>((double)pr.r * half_dq);
        }
    };

    if (n_tiles > 0) geometry(0, sm.t[0]);
    uint32_t cur = CLS_SKIP;
    for (int tile = 0; tile < n_tiles; ++tile) {
        __syncthreads();  // tile's geometry is complete; every wave is done with the other buffer
        if (tile + 1 < n_tiles) geometry(tile + 1, sm.t[(tile + 1) & 1]);
        const Tile& b = sm.t[tile & 1];
        if (active) {
            const int t0 = tile * SH_TC;
            const int tn = ncol - t0 < SH_TC ? ncol - t0 : SH_TC;
            // one weighted pair step per iteration (the second column is the next one
            // when it is in this tile and of the same class, otherwise a zero-weight
            // copy); skipped records and the lower half of the diagonal tile weigh 0
            int t = 0;
            while (t < tn) {
                const int j = t0 + t;
                const uint32_t cl0 = __float_as_uint(cols[j].w);
                if (cl0 != cur) {
                    if (cur != CLS_SKIP) flush(cur);
                    cur = cl0;
                }
                const bool pair = t + 1 < tn && __float_as_uint(cols[j + 1].w) == cl0;
                const int t1 = pair ? t + 1 : t;
                const float live = cl0 != CLS_SKIP ? 1.0f : 0.0f;
                float w0 = live, w1 = pair ? live : 0.0f;
                if (j < jd) {
                    w0 *= diag_weight(j, lane);
                    w1 *= diag_weight(j + 1, lane);
                }
                run2<K>(acc, chunk_seeds(b.r[t][lane], b.ri[t][lane], b.sh[t][lane], cc, dq, w0),
                        chunk_seeds(b.r[t1][lane], b.ri[t1][lane], b.sh[t1][lane], cc, dq, w1));
                t += pair ? 2 : 1;
            }
        }
    }
    if (active && cur != CLS_SKIP) flush(cur);
    __syncthreads();  // the tile buffers are reused for the epilogue

    // epilogue as in debye_uniform_kernel, through this wave's slice of the union
    const uint32_t ci = cls_i == CLS_SKIP ? (uint32_t)a.zero_class : cls_i;
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
#pragma unroll
    for (int k = 0; k < K; ++k) sm.T[wave][k][lane] = fi[k] * X[k];
    __syncthreads();
    if (active && lane < K) {
        double sum = 0.0;
#pragma unroll 16
        for (int t = 0; t < 64; ++t) sum += (double)sm.T[wave][lane][t];
        a.partials[(size_t)jb.slot * ldf + k0 + lane] = sum * a.invq[k0 + lane] * (double)jb.scale;
    }
}

// ---------------------------------------------------------------------------
// uniform-grid path, FP32 variant: the difference form of the recurrence
//     d_{k+1} = d_k - delta u_k,   u_{k+1} = u_k + d_{k+1},   delta = 4 sin^2(dq r / 2)
// (u_{k+1} = (2 - delta) u_k - u_{k-1}, the same three-term recurrence).  Unlike
// u_{k+1} = c u_k - u_{k-1} with c = 2 cos(dq r) rounded near 2, the small step
// delta keeps its full relative precision, so FP32 rounding grows only linearly in k
// for any dq r (~6e-7 absolute over 64 steps for 1e-4 <= dq r <= 1 in an FP32
// emulation).  FP32 VALU instructions issue ~1.7x as fast as FP64 ones on gfx950 at
// 2-3 waves per SIMD (measured), so three FP32 ops per term (FMA, add, add into the
// sum) cost less than the FP64 form's two FP64 ops.
// Seeds keep FP64 phases (the phase a r reaches ~80 revolutions, beyond FP32's
// resolution); sin(dq r / 2) only needs FP32 relative accuracy here (delta and d_1 are
// FP32), so it comes from an FP32 odd polynomial -- not v_sin, whose absolute error
// is large relative to small dq r.
// ---------------------------------------------------------------------------
struct Seeds32 {
    float u0, d1, delta;
};

// sin(x) for 0 <= x <= x_max in FP32: through x^7 when the layout's series length is
// the short one (x_max <= ~0.2, truncation < 1e-9 relative), otherwise through x^11
// (x <= 1.5, truncation < 3e-8)
template <int M>
__device__ __forceinline__ float sin_small_f32(float x)
{
    const float x2 = x * x;
    float p;
    if (M <= 5) {
        p = fmaf(x2, -1.98412698e-04f, 8.33333333e-03f);
    } else {
        p = fmaf(x2, -2.50521084e-08f, 2.75573192e-06f);
        p = fmaf(p, x2, -1.98412698e-04f);
        p = fmaf(p, x2, 8.33333333e-03f);
    }
    p = fmaf(p, x2, -1.66666667e-01f);
    return fmaf(x * x2, p, x);
}

template <int M>
__device__ __forceinline__ Seeds32 make_seeds32(float r2, const ChunkConst& cc, float half_dq, float dq, float w)
{
    const float nzf = r2 > 1e-30f ? 1.0f : 0.0f;
    const PairR pr = pair_r(r2);
    const float rinv = pr.rinv, r = pr.r;
    const double rd = (double)r;
    const float sh = sin_small_f32<M>(r * half_dq);  // sin(dq r / 2)
    const float sa = sin_rev(cc.a_rev * rd);         // sin(a r)
    const float cm = cos_rev(cc.mid_rev * rd);       // cos((a + dq/2) r)
    const float wr = w * rinv;
    const float wz = w * (1.0f - nzf);
    Seeds32 s;
    s.delta = 4.0f * sh * sh;
    s.u0 = fmaf(sa, wr, wz * cc.a);  // r == 0: u_k = w q_k, linear in k
    s.d1 = fmaf((cm + cm) * wr, sh, wz * dq);
    return s;
}

template <int K>
__device__ __forceinline__ void run2_32(float (&acc)[K], const Seeds32& a, const Seeds32& b)
{
    float ua = a.u0, da = a.d1, ub = b.u0, db = b.d1;
    acc[0] += ua + ub;
    ua += da;
    ub += db;
    acc[1] += ua + ub;
#pragma unroll
    for (int k = 2; k < K; ++k) {
        da = fmaf(-a.delta, ua, da);
        db = fmaf(-b.delta, ub, db);
        ua += da;
        ub += db;
        acc[k] += ua + ub;
    }
}

// The column sweep of one (job, q chunk) item of the FP32 recurrence: columns in
// pairs, acc[k] += sin(q_k r) / r over the current class run; flush(cls) is called when
// the run of class cls ends.  Skip records (the slots of moved atoms during an
// incremental update) join whatever run they sit in with weight 0 instead of ending it.
template <int K, int M, typename Flush>
__device__ __forceinline__ void rec32_columns(const PairJob& jb, const float4 ri, const ChunkConst& cc, float h_dq,
                                              float dq, float (&acc)[K], Flush&& flush)
{
    const int lane = threadIdx.x;
    const ConstCols cols{(cv4f*)jb.cols};
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    uint32_t cur = CLS_SKIP;
    int j = 0;
    while (j < ncol) {
        const float4 c0 = cols[j];
        const uint32_t cl0 = __float_as_uint(c0.w);
        const uint32_t cl1 = j + 1 < ncol ? __float_as_uint(cols[j + 1].w) : CLS_SKIP;
        const uint32_t cl = cl0 != CLS_SKIP ? cl0 : cl1;  // class of this step (skip if both are)
        if (cl != CLS_SKIP && cl != cur) {
            if (cur != CLS_SKIP) flush(cur);
            cur = cl;
        }
        const bool pair = j + 1 < ncol && (cl1 == cl || cl1 == CLS_SKIP);
        const float4 c1 = pair ? cols[j + 1] : c0;
        float w0 = cl0 != CLS_SKIP ? 1.0f : 0.0f, w1 = pair && cl1 != CLS_SKIP ? 1.0f : 0.0f;
        if (j < jd) {
            w0 *= diag_weight(j, lane);
            w1 *= diag_weight(j + 1, lane);
        }
        run2_32<K>(acc, make_seeds32<M>(pair_r2(ri.x, ri.y, ri.z, c0), cc, h_dq, dq, w0),
                   make_seeds32<M>(pair_r2(ri.x, ri.y, ri.z, c1), cc, h_dq, dq, w1));
        j += pair ? 2 : 1;
    }
    if (cur != CLS_SKIP) flush(cur);
}

// one (job, q chunk) work item of the FP32 recurrence; T is the wave's LDS transpose
template <int K, int M, int KS>
__device__ __forceinline__ void rec32_item(const DebyeArgs& a, int job_id, int chunk, float (&T)[KS][65])
{
    const PairJob jb = a.jobs[job_id];
    const int lane = threadIdx.x;
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const int k0 = chunk * K;
    const ChunkConst cc = a.chunks[chunk];
    const float* __restrict__ ftab = a.ftab;
    const int ldf = a.ldf;

    // acc: FP32 sums of the current class run; X: per-lane sums over flushed runs
    float acc[K], X[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc[k] = 0.0f;
        X[k] = 0.0f;
    }
    rec32_columns<K, M>(jb, ri, cc, (float)a.half_dq, a.dq, acc, [&](uint32_t cls) {
        const float* __restrict__ fr = ftab + __builtin_amdgcn_readfirstlane((int)cls * ldf + k0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            X[k] = fmaf(fr[k], acc[k], X[k]);
            acc[k] = 0.0f;
        }
    });

    // epilogue as in debye_uniform_kernel, in slices of up to KS (32) q so the LDS
    // transpose stays 8.3 KB for any K: LDS then does not limit the waves per CU below
    // what the registers allow (16.6 KB slices held it to nine per CU)
    const uint32_t ci = cls_i == CLS_SKIP ? (uint32_t)a.zero_class : cls_i;
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
#pragma unroll
    for (int kb = 0; kb < K; kb += KS) {
#pragma unroll
        for (int k = 0; k < KS; ++k)
            if (kb + k < K) T[k][lane] = fi[kb + k] * X[kb + k];
        __syncthreads();
        if (lane < KS && kb + lane < K) {
            double sum = 0.0;
#pragma unroll 16
            for (int t = 0; t < 64; ++t) sum += (double)T[lane][t];
            a.partials[(size_t)jb.slot * ldf + k0 + kb + lane] = sum * a.invq[k0 + kb + lane] * (double)jb.scale;
        }
        __syncthreads();
    }
}

// Run-flush form of the FP32 recurrence: no per-lane sums over runs.  Each class run
// is reduced over the 64 rows when it ends: the lane's row factor times its run sum
// goes through an LDS transpose in slices of RSLICE q, lanes (k, half) sum 32 rows
// each, a cross-half shuffle completes the column, and the run's class factor scales
// it into an FP64 per-q total.  Only the K run accumulators stay in registers, so K
// reaches 128 at three waves per SIMD; the flush costs ~4K lane operations, which
// pays off when a column range holds few class runs (type classes).
constexpr int RSLICE = 32;
// Row factors of a flush are loaded in groups of FGROUP between compiler memory
// barriers.  Larger groups keep more loads in flight per flush (each group waits for its
// loads) at the price of registers: at K = 128 and 168 VGPRs, groups of 16 spill 16
// values around the flush and still run ~2 % faster on hydration layouts (many flushes)
// than groups of 4 (no spills), and equal elsewhere.
constexpr int FGROUP = 16;
// row stride of the run flush's transpose: even, so a lane's 32 rows read as float2 pairs
constexpr int RPAD = 66;

template <int K, int M>
__device__ __forceinline__ void rec32r_item(const DebyeArgs& a, int job_id, int chunk, float (&T)[RSLICE][RPAD])
{
    constexpr int NS = (K + RSLICE - 1) / RSLICE;
    const PairJob jb = a.jobs[job_id];
    const int lane = threadIdx.x;
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const uint32_t ci = cls_i == CLS_SKIP ? (uint32_t)a.zero_class : cls_i;
    const int k0 = chunk * K;
    const ChunkConst cc = a.chunks[chunk];
    const float* __restrict__ ftab = a.ftab;
    const int ldf = a.ldf;
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
    const int ks = lane & (RSLICE - 1), half = lane >> 5;

    float acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0f;
    double Y[NS];  // lane (ks, half): total of q index s * RSLICE + ks over flushed runs
#pragma unroll
    for (int s = 0; s < NS; ++s) Y[s] = 0.0;
    rec32_columns<K, M>(jb, ri, cc, (float)a.half_dq, a.dq, acc, [&](uint32_t cls) {
        const float* __restrict__ fr = ftab + __builtin_amdgcn_readfirstlane((int)cls * ldf + k0);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
#pragma unroll
            for (int k = 0; k < RSLICE; ++k) {
                // row factors in groups of FGROUP loads: the compiler may not gather a
                // slice's loads into live registers on top of the K accumulators
                if (k % FGROUP == 0) asm volatile("" ::: "memory");
                if (s * RSLICE + k < K) {
                    T[k][lane] = fi[s * RSLICE + k] * acc[s * RSLICE + k];
                    acc[s * RSLICE + k] = 0.0f;
                }
            }
            __syncthreads();
            // the 64 rows' contributions cancel deeply near minima of I(q): summed in
            // FP64 (an FP32 sum here cost up to 2e-5 relative at such q on long runs),
            // each pair of rows first added in FP32 (one more rounding of the size of the
            // product's own: 3.9e-6 -> 4.4e-6 at the deepest minimum of the hydrated dip
            // test, for 2 % less time on hydration layouts)
            double sum = 0.0;
            const float2* row = reinterpret_cast<const float2*>(&T[ks][half * 32]);
#pragma unroll 8
            for (int t = 0; t < 16; ++t) {
                const float2 v = row[t];
                sum += (double)(v.x + v.y);
            }
            sum += __shfl_xor(sum, 32, 64);
            if (s * RSLICE + ks < K) Y[s] = fma((double)fr[s * RSLICE + ks], sum, Y[s]);
            __syncthreads();
        }
    });
    if (half == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int k = s * RSLICE + ks;
            if (k < K) a.partials[(size_t)jb.slot * ldf + k0 + k] = Y[s] * a.invq[k0 + k] * (double)jb.scale;
        }
    }
}

template <int K, int M>
__global__ __launch_bounds__(64, 2) void debye_rec32_kernel(const DebyeArgs a)
{
    constexpr int KS = K < 32 ? K : 32;
    __shared__ float T[KS][65];
    const int job_id = blockIdx.x / a.n_chunks;
    rec32_item<K, M, KS>(a, job_id, blockIdx.x - job_id * a.n_chunks, T);
}

// Persistent form: a grid sized to the resident waves pulls (job, chunk) items from
// a work counter in job order (longest first) until the list is exhausted; every
// wave exits once the counter passes the item count.  The counter is zeroed by a
// fill kernel on the same stream before the launch (graph-capturable).
template <int K, int M>
__global__ __launch_bounds__(64, 2) void debye_rec32_persistent(const DebyeArgs a, int n_items)
{
    constexpr int KS = K < 32 ? K : 32;
    __shared__ float T[KS][65];
    for (;;) {
        int item = 0;
        if (threadIdx.x == 0) item = atomicAdd(a.queue, 1);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= n_items) break;
        const int job_id = item / a.n_chunks;
        rec32_item<K, M, KS>(a, job_id, item - job_id * a.n_chunks, T);
    }
}

// three resident waves per SIMD: K = 128 fits in 168 registers
template <int K, int M>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void debye_rec32r_kernel(const DebyeArgs a)
{
    __shared__ float T[RSLICE][RPAD];
    const int job_id = blockIdx.x / a.n_chunks;
    rec32r_item<K, M>(a, job_id, blockIdx.x - job_id * a.n_chunks, T);
}

// (the work-queue loop needs a few more registers: K = 128 runs at two waves per SIMD)
template <int K, int M>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void debye_rec32r_persistent(const DebyeArgs a, int n_items)
{
    __shared__ float T[RSLICE][RPAD];
    for (;;) {
        int item = 0;
        if (threadIdx.x == 0) item = atomicAdd(a.queue, 1);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= n_items) break;
        const int job_id = item / a.n_chunks;
        rec32r_item<K, M>(a, job_id, item - job_id * a.n_chunks, T);
    }
}

// ---------------------------------------------------------------------------
// direct policy (any q grid): FP32 accumulators, one v_sin per term
// ---------------------------------------------------------------------------
// One column against the wave's rows at K q.  The phase q_k r / (2 pi) goes to v_sin_f32
// in revolutions.  PH32: formed in FP32 as fma(hi_k, r, lo_k r) from the split
// hi_k + lo_k = q_k / (2 pi), a single rounding of a phase that is otherwise exact for the
// FP32 r (v_sin_f32 reduces its argument exactly, so no fraction is taken); the caller
// guarantees the phase stays within DIRECT_PHASE32_MAX_REV.  Otherwise formed in FP64 and
// reduced to [0, 1) before the conversion.
template <int K, bool WEIGHTED, bool PH32>
__device__ __forceinline__ void direct_step(float (&acc)[K], float& zsum, float xi, float yi, float zi,
                                            const float4 c, const double* __restrict__ qrev,
                                            const float* __restrict__ qsplit, float w)
{
    const float dx = xi - c.x, dy = yi - c.y, dz = zi - c.z;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const bool nz = r2 > 1e-30f;
    const PairR pr = pair_r(r2);
    const float rinv = pr.rinv, r = pr.r;
    const float wr = WEIGHTED ? w * rinv : rinv;
    // r == 0 pairs contribute w * q_k; their weights are summed and applied at the flush
    zsum += nz ? 0.0f : (WEIGHTED ? w : 1.0f);
    if (PH32) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float ph = fmaf(qsplit[2 * k], r, qsplit[2 * k + 1] * r);
            acc[k] = fmaf(__builtin_amdgcn_sinf(ph), wr, acc[k]);
        }
    } else {
        const double rd = (double)r;
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fmaf(sin_rev(qrev[k] * rd), wr, acc[k]);
    }
}

// columns summed in FP32 before their sums go to the FP64 accumulators (a class run of
// plain factors can span a job's whole column range, up to 1024 columns)
constexpr int DIRECT_FLUSH = 32;

template <int K, bool PH32>
__global__ __launch_bounds__(64, 2) void debye_direct_kernel(const DebyeArgs a)
{
    const int job_id = blockIdx.x / a.n_chunks;
    const int chunk = blockIdx.x - job_id * a.n_chunks;
    const PairJob jb = a.jobs[job_id];
    const int lane = threadIdx.x;
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const int k0 = chunk * K;
    const float* __restrict__ ftab = a.ftab;
    const int ldf = a.ldf;
    const double* __restrict__ qrev = a.qrev + k0;
    const float* __restrict__ qsplit = a.qsplit + 2 * k0;
    const float* __restrict__ qv = a.qv + k0;

    float acc[K];
    double X[K];
    float zsum = 0.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc[k] = 0.0f;
        X[k] = 0.0;
    }
    auto flush = [&](uint32_t cls) {
        // the class is wave-uniform: a scalar offset lets the row come in through
        // scalar loads instead of per-lane vector loads
        const float* __restrict__ fr = ftab + __builtin_amdgcn_readfirstlane((int)cls * ldf + k0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            X[k] = fma((double)fr[k], (double)fmaf(zsum, qv[k], acc[k]), X[k]);
            acc[k] = 0.0f;
        }
        zsum = 0.0f;
    };

    const ConstCols cols{(cv4f*)jb.cols};
    const int ncol = jb.ncol;
    uint32_t cur = CLS_SKIP;
    int since = 0;  // columns in acc since the last flush (uniform)
    int j = 0;
    if (jb.diag) {
        const int jd = ncol < 64 ? ncol : 64;
        for (; j < jd; ++j) {
            const float4 c = cols[j];
            const uint32_t cl = __float_as_uint(c.w);
            if (cl != cur || since == DIRECT_FLUSH) {
                if (cur != CLS_SKIP) flush(cur);
                cur = cl;
                since = 0;
            }
            if (cl != CLS_SKIP) {
                const float w = j > lane ? 1.0f : (j == lane ? 0.5f : 0.0f);
                direct_step<K, true, PH32>(acc, zsum, ri.x, ri.y, ri.z, c, qrev, qsplit, w);
                ++since;
            }
        }
    }
    for (; j < ncol; ++j) {
        const float4 c = cols[j];
        const uint32_t cl = __float_as_uint(c.w);
        if (cl != cur || since == DIRECT_FLUSH) {
            if (cur != CLS_SKIP) flush(cur);
            cur = cl;
            since = 0;
        }
        if (cl != CLS_SKIP) {
            direct_step<K, false, PH32>(acc, zsum, ri.x, ri.y, ri.z, c, qrev, qsplit, 1.0f);
            ++since;
        }
    }
    if (cur != CLS_SKIP) flush(cur);

    const uint32_t ci = cls_i == CLS_SKIP ? 0u : cls_i;
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
    const double* __restrict__ invq = a.invq + k0;
    const double scale = (double)jb.scale;
    double mine = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double v = cls_i == CLS_SKIP ? 0.0 : (double)fi[k] * X[k];
        v = wave_sum(v) * invq[k] * scale;
        mine = lane == k ? v : mine;
    }
    if (lane < K) a.partials[(size_t)jb.slot * ldf + k0 + lane] = mine;
}

// ---------------------------------------------------------------------------
// fixed-order reduction of the partial table
// ---------------------------------------------------------------------------
constexpr int RED_GROUPS = 16;

__global__ __launch_bounds__(64 * RED_GROUPS) void reduce_kernel(const double* __restrict__ partials, int ldf, int n_q,
                                                               const int* __restrict__ slot_begin, double* out,
                                                               int out_ld, int accumulate,
                                                               const double* __restrict__ q0_val, int q0_index)
{
    __shared__ double part[RED_GROUPS][64];
    const int o = blockIdx.y;
    const int kl = threadIdx.x & 63;
    const int g = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + kl;
    const int s0 = slot_begin[o], s1 = slot_begin[o + 1];
    double acc = 0.0;
    if (k < n_q)
        for (int s = s0 + g; s < s1; s += RED_GROUPS) acc += partials[(size_t)s * ldf + k];
    part[g][kl] = acc;
    __syncthreads();
    if (g == 0 && k < n_q) {
        double tot = 0.0;
#pragma unroll
        for (int t = 0; t < RED_GROUPS; ++t) tot += part[t][kl];
        double* dst = out + (size_t)o * out_ld + k;
        if (k == q0_index) tot = q0_val[o];
        *dst = accumulate ? *dst + tot : tot;
    }
}

// Two-stage form of the fixed-order reduction for many slots: stage 1 sums slot chunk z
// of output o into scratch[(o * nsplit + z) * n_q + k] (16 interleaved groups of the
// chunk, then the groups in order); stage 2 sums the chunks in order.
__global__ __launch_bounds__(64 * RED_GROUPS) void reduce_part_kernel(const double* __restrict__ partials, int ldf,
                                                                    int n_q, const int* __restrict__ slot_begin,
                                                                    int chunk, int nsplit, double* scratch)
{
    __shared__ double part[RED_GROUPS][64];
    const int o = blockIdx.y, z = blockIdx.z;
    const int kl = threadIdx.x & 63;
    const int g = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + kl;
    const int sb = slot_begin[o], se = slot_begin[o + 1];
    const int s0 = sb + z * chunk;
    const int s1 = s0 + chunk < se ? s0 + chunk : se;
    double acc = 0.0;
    if (k < n_q)
        for (int s = s0 + g; s < s1; s += RED_GROUPS) acc += partials[(size_t)s * ldf + k];
    part[g][kl] = acc;
    __syncthreads();
    if (g == 0 && k < n_q) {
        double tot = 0.0;
#pragma unroll
        for (int t = 0; t < RED_GROUPS; ++t) tot += part[t][kl];
        scratch[((size_t)o * nsplit + z) * n_q + k] = tot;
    }
}

__global__ void reduce_final_kernel(const double* __restrict__ scratch, int nsplit, int n_q, double* out, int out_ld,
                                    int accumulate, const double* __restrict__ q0_val, int q0_index)
{
    const int o = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_q) return;
    double tot = 0.0;
    for (int z = 0; z < nsplit; ++z) tot += scratch[((size_t)o * nsplit + z) * n_q + k];
    double* dst = out + (size_t)o * out_ld + k;
    if (k == q0_index) tot = q0_val[o];
    *dst = accumulate ? *dst + tot : tot;
}

__global__ void frame_mean_kernel(const double* __restrict__ in, int n_frames, int n_q, double* out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_q) return;
    double s = 0.0;
    for (int f = 0; f < n_frames; ++f) s += in[(size_t)f * n_q + k];
    out[k] = s / (double)n_frames;
}

__global__ void scatter_value_kernel(float4* cols, const int* __restrict__ slot, float4 v, int m)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) cols[slot[t]] = v;
}

__global__ void scatter_slots_kernel(float4* cols, int* s2c, const int* __restrict__ slot,
                                     const float4* __restrict__ vals, const int* __restrict__ callers, int m)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        cols[slot[t]] = vals[t];
        s2c[slot[t]] = callers[t];
    }
}

__global__ void gather_cols_kernel(float4* cols, const float4* __restrict__ bodies, const int* __restrict__ src,
                                   const uint32_t* __restrict__ counts, int d1, int n, GatherGuard g)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int c = src[s];
    const float4 b = bodies[c];
    const uint32_t type = __float_as_uint(b.w);
    int bad = 0;
    uint32_t cls = CLS_SKIP;
    if (type < (uint32_t)g.n_types)
        cls = type * (uint32_t)d1 + (counts ? counts[c] : 0u);
    else
        bad |= GUARD_TYPE;
    const float dx = b.x - g.cx, dy = b.y - g.cy, dz = b.z - g.cz;
    if (!(dx * dx + dy * dy + dz * dz <= g.r2_max)) bad |= GUARD_RADIUS;
    if (bad) atomicOr(g.flag, bad);
    cols[s] = make_float4(b.x, b.y, b.z, __uint_as_float(cls));
}

// ---- adjoint weights on the device ----
__global__ void adjoint_weights_kernel(const double* __restrict__ w, int n_q, int q0_index, int ldfa, double* wq)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ldfa) wq[k] = (k < n_q && k != q0_index) ? w[k] : 0.0;
}

// one workgroup of 1024: each thread sums its q values in index order, then a fixed-shape
// tree over the lanes and waves, so the result repeats bit for bit
__device__ __forceinline__ double block_sum_1024(double v, double* part)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int u = 0; u < 16; ++u) t += part[u];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(1024) void chi2_weights_kernel(const double* __restrict__ iq,
                                                           const float* __restrict__ i_exp,
                                                           const float* __restrict__ sigma, int n_q, int q0_index,
                                                           int ldfa, double* out, double* wq)
{
    __shared__ double part[16];
    double num = 0.0, den = 0.0;
    for (int k = threadIdx.x; k < n_q; k += 1024) {
        const double p2 = 1.0 / ((double)sigma[k] * (double)sigma[k]);
        num += (double)i_exp[k] * iq[k] * p2;
        den += iq[k] * iq[k] * p2;
    }
    num = block_sum_1024(num, part);
    den = block_sum_1024(den, part);
    const double c = num / den;
    double x2 = 0.0;
    for (int k = threadIdx.x; k < ldfa; k += 1024) {
        double wk = 0.0;
        if (k < n_q) {
            const double s2 = (double)sigma[k] * (double)sigma[k];
            const double d = (double)i_exp[k] - c * iq[k];
            x2 += d * d / s2;
            if (k != q0_index) wk = -2.0 * c * d / s2;
        }
        wq[k] = wk;
    }
    x2 = block_sum_1024(x2, part);
    if (threadIdx.x == 0) {
        out[0] = x2;
        out[1] = c;
    }
}

// ---- stable counting sort by class (device-resident hydrated steps) ----
constexpr int CS_BLOCK = 1024;

__device__ __forceinline__ int sort_key(const ClassSortArgs& a, uint32_t type, int i)
{
    return type < (uint32_t)a.n_types ? (int)(type * (uint32_t)a.d1 + a.counts[i]) : a.n_class;
}

// per block of CS_BLOCK bodies: how many of each key (LDS histogram)
__global__ __launch_bounds__(CS_BLOCK) void class_hist_kernel(ClassSortArgs a)
{
    extern __shared__ int hist[];
    const int nk = a.n_class + 1;
    for (int c = threadIdx.x; c < nk; c += CS_BLOCK) hist[c] = 0;
    __syncthreads();
    const int i = blockIdx.x * CS_BLOCK + threadIdx.x;
    if (i < a.n) atomicAdd(&hist[sort_key(a, __float_as_uint(a.bodies[i].w), i)], 1);
    __syncthreads();
    for (int c = threadIdx.x; c < nk; c += CS_BLOCK) a.table[(size_t)blockIdx.x * nk + c] = hist[c];
}

// one workgroup: table[b][c] becomes the first slot of block b's bodies of key c, i.e. the
// bodies of all smaller keys plus those of key c in blocks before b
__global__ __launch_bounds__(1024) void class_scan_kernel(ClassSortArgs a, int n_blk)
{
    __shared__ int wave_tot[16];
    __shared__ int carry_s;
    const int nk = a.n_class + 1;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int c = tid; c < nk; c += 1024) {
        int off = 0;
        for (int b = 0; b < n_blk; ++b) {
            int* p = a.table + (size_t)b * nk + c;
            const int v = *p;
            *p = off;
            off += v;
        }
        a.tot[c] = off;
    }
    __syncthreads();
    // exclusive scan of the key totals in passes of 1024 with a running carry
    int carry = 0;
    for (int c0 = 0; c0 < nk; c0 += 1024) {
        const int c = c0 + tid;
        const int v = c < nk ? a.tot[c] : 0;
        int x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wave_tot[w] = x;
        __syncthreads();
        int off = carry;
        for (int u = 0; u < w; ++u) off += wave_tot[u];
        if (c < nk) a.tot[c] = off + x - v;
        if (tid == 1023) carry_s = off + x;
        __syncthreads();
        carry = carry_s;
        __syncthreads();
    }
    for (int c = tid; c < nk; c += 1024) {
        const int base = a.tot[c];
        for (int b = 0; b < n_blk; ++b) a.table[(size_t)b * nk + c] += base;
    }
}

// each body to its slot: its block's first slot of its key plus the bodies of the same key
// before it in the block (caller order, so the sort is stable)
__global__ __launch_bounds__(CS_BLOCK) void class_scatter_kernel(ClassSortArgs a)
{
    __shared__ int keys[CS_BLOCK];
    const int i = blockIdx.x * CS_BLOCK + threadIdx.x;
    int key = -1;
    float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < a.n) {
        b = a.bodies[i];
        const uint32_t type = __float_as_uint(b.w);
        key = sort_key(a, type, i);
        int bad = type < (uint32_t)a.n_types ? 0 : GUARD_TYPE;
        const float dx = b.x - a.g.cx, dy = b.y - a.g.cy, dz = b.z - a.g.cz;
        if (!(dx * dx + dy * dy + dz * dz <= a.g.r2_max)) bad |= GUARD_RADIUS;
        if (bad) atomicOr(a.g.flag, bad);
    }
    keys[threadIdx.x] = key;
    __syncthreads();
    if (i < a.n) {
        int rank = 0;
        for (int t = 0; t < (int)threadIdx.x; ++t) rank += keys[t] == key ? 1 : 0;
        const int slot = a.table[(size_t)blockIdx.x * (a.n_class + 1) + key] + rank;
        a.cols[slot] = make_float4(b.x, b.y, b.z, __uint_as_float(key == a.n_class ? CLS_SKIP : (uint32_t)key));
        a.s2c[slot] = i;
    }
}

__global__ __launch_bounds__(1024) void q0_kernel(const float4* __restrict__ cols, int n,
                                                  const double* __restrict__ f0col, double* out)
{
    __shared__ double part[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const uint32_t c = __float_as_uint(cols[i].w);
        if (c != CLS_SKIP) s += f0col[c];
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += part[w];
        out[0] = t * t;
    }
}

__global__ __launch_bounds__(256) void copy_probe_kernel(float4* __restrict__ dst, const float4* __restrict__ src,
                                                         long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void fill_i32_kernel(int* p, int v, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// SASA: hashed cell grid + dot occlusion
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_coord(float x, float inv_cell)
{
    return (int)floorf(x * inv_cell);
}

// The occlusion predicate of rsasa_dot_spheres.cl:122-132 / rsasa_cpu_alg.cpp:62-68:
// (p + r_j)^2 - |dot - x_j|^2 > 1e-4, the squared length summed x, y, z in order with
// each product and sum rounded on its own.  Contraction is off inside this function
// (the build's default fuses a*b + c into one FMA, which rounds once and moves the
// near-ties of the 1e-4 test), so the integer counts equal an unfused FP32 evaluation.
__device__ __forceinline__ bool dot_occluded(float dx, float dy, float dz, float r2j)
{
#pragma clang fp contract(off)
    float dist2 = dx * dx;
    dist2 = dist2 + dy * dy;
    dist2 = dist2 + dz * dz;
    return (r2j - dist2) > 1e-4f;
}

// Bucket of cell (ix, iy, iz) among nb (a power of two): the low 10 bits of each coordinate
// packed into one word (distinct for cells within 1024 of each other per axis, so no two
// cells of a structure share the key), mixed by the 32-bit finalizer of MurmurHash3
// (A. Appleby, public domain) so that the low bits the mask keeps depend on every key bit.
// (A XOR of the coordinates times three large primes, taken modulo a power of two, gave
// 3117 distinct values for the 4805 cells of a 50,000-atom ellipsoid: 1.5 cells per bucket.)
// Atoms of other cells in a bucket only cost prefilter work, never a count.
__device__ __forceinline__ int cell_hash(int ix, int iy, int iz, int nb)
{
    uint32_t h = ((uint32_t)ix & 1023u) | (((uint32_t)iy & 1023u) << 10) | (((uint32_t)iz & 1023u) << 20);
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return (int)(h & (uint32_t)(nb - 1));
}

// First kernel of a local recount with a patch: the bucket counts cleared (as the fill kernel
// does otherwise) and the stage's moved atoms written into the atom array.
__global__ void sasa_patch_kernel(SasaArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= a.n_buckets) a.bucket_count[t] = 0;
    const LocalStage& st = *a.local_stage;
    if (t < st.len && t < a.patch_cap) {
        const float4* moved = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(a.local_stage) +
                                                              LOCAL_STAGE_HEADER);
        a.atoms_w[st.start + t] = moved[t];
    }
}

__global__ void sasa_bin_kernel(SasaArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float4 p = a.atoms[i];
    const int b = cell_hash(cell_coord(p.x, a.inv_cell), cell_coord(p.y, a.inv_cell), cell_coord(p.z, a.inv_cell),
                            a.n_buckets);
    a.bucket_of[i] = b;
    atomicAdd(&a.bucket_count[b], 1);
}

// Exclusive scan of the bucket counts in three steps: tiles of SCAN_TILE buckets
// (each thread one int4, coalesced) scanned locally with their totals set aside; the
// tile totals scanned by one workgroup; the tile offsets added back.  n_buckets is a
// power of two >= SCAN_TILE.
constexpr int SCAN_TILE = 1024;

__global__ __launch_bounds__(256) void sasa_scan_tiles_kernel(SasaArgs a)
{
    __shared__ int wave_tot[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int b = blockIdx.x * SCAN_TILE + tid * 4;
    const int4 c = *reinterpret_cast<const int4*>(a.bucket_count + b);
    const int sum = c.x + c.y + c.z + c.w;
    int x = sum;  // inclusive scan within the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wave_tot[w] = x;
    __syncthreads();
    int off = 0;
    for (int t = 0; t < w; ++t) off += wave_tot[t];
    const int e = off + x - sum;
    *reinterpret_cast<int4*>(a.bucket_start + b) = make_int4(e, e + c.x, e + c.x + c.y, e + c.x + c.y + c.z);
    if (tid == 255) a.tile_sum[blockIdx.x] = off + x;
}

// exclusive scan of the tile totals in place, in passes of 1024 with a running carry;
// bucket_start[n_buckets] = the total
__global__ __launch_bounds__(1024) void sasa_scan_top_kernel(SasaArgs a)
{
    __shared__ int wave_tot[16];
    __shared__ int carry_s;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n_tiles = a.n_buckets / SCAN_TILE;
    int carry = 0;
    for (int t0 = 0; t0 < n_tiles; t0 += 1024) {
        const int t = t0 + tid;
        const int v = t < n_tiles ? a.tile_sum[t] : 0;
        int x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wave_tot[w] = x;
        __syncthreads();
        int off = carry;
        for (int u = 0; u < w; ++u) off += wave_tot[u];
        if (t < n_tiles) a.tile_sum[t] = off + x - v;
        if (tid == 1023) carry_s = off + x;
        __syncthreads();
        carry = carry_s;
        __syncthreads();
    }
    if (tid == 0) a.bucket_start[a.n_buckets] = carry;
}

__global__ __launch_bounds__(256) void sasa_scan_add_kernel(SasaArgs a)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int v = a.bucket_start[b] + a.tile_sum[b / SCAN_TILE];
    a.bucket_start[b] = v;
    a.bucket_fill[b] = v;
    // both counter sets: this recount's, and the next one's (which the next recount, on the
    // slab, takes as cleared)
    if (a.local_hdr && b < LOCAL_HDR_TOTAL) a.local_hdr[b] = 0;
}

// Appends atom i to the local recount's list (set hdr) when its position p lies inside the
// box: one atomic per wave for the wave's listed atoms.
__device__ __forceinline__ void list_if_in_box(const SasaArgs& a, int* hdr, const LocalStage& st, int i, bool valid,
                                               const float4 p)
{
    const bool in = valid && p.x >= st.lo[0] && p.x <= st.hi[0] && p.y >= st.lo[1] && p.y <= st.hi[1] &&
                    p.z >= st.lo[2] && p.z <= st.hi[2];
    const uint64_t m = __ballot(in);
    if (m) {
        const int lane = threadIdx.x & 63, first = __builtin_ctzll(m);
        int base = 0;
        if (lane == first) base = atomicAdd(&hdr[0], __builtin_popcountll(m));
        base = __shfl(base, first, 64);
        const int pos = base + __builtin_popcountll(m & ((1ull << lane) - 1ull));
        if (in && pos < a.n) a.local_list[pos] = i;  // (each atom is listed once: pos < n)
    }
}

// Places atom i in its bucket; for a local recount also lists it when inside the box.
__global__ void sasa_place_kernel(SasaArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (a.local_hdr) {  // uniform
        const LocalStage& st = *a.local_stage;
        list_if_in_box(a, a.local_hdr + LOCAL_HDR_INTS * st.parity, st, i, i < a.n,
                       i < a.n ? a.atoms[i] : make_float4(0.f, 0.f, 0.f, 0.f));
    }
    if (i >= a.n) return;
    const int pos = atomicAdd(&a.bucket_fill[a.bucket_of[i]], 1);
    a.sorted[pos] = i;
    a.sorted_atoms[pos] = a.atoms[i];
}

// ---- slab cell list (local recounts) ----
__device__ __forceinline__ int slab_bucket(const SasaArgs& a, const float4 p)
{
    return cell_hash(cell_coord(p.x, a.inv_cell), cell_coord(p.y, a.inv_cell), cell_coord(p.z, a.inv_cell),
                     a.slab_buckets);
}

// Appends record p of atom i to bucket b; returns false (nothing written, slab_pos[i] = -1)
// when the bucket is full.
__device__ __forceinline__ bool slab_append(const SasaArgs& a, int i, int b, const float4 p)
{
    const int s = atomicAdd(&a.slab_count[b], 1);
    if (s >= SLAB_CAP) {
        a.slab_pos[i] = -1;
        return false;
    }
    a.slab[(size_t)b * SLAB_CAP + s] = p;
    a.slab_pos[i] = b * SLAB_CAP + s;
    return true;
}

__global__ void slab_clear_kernel(SasaArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.slab_buckets) a.slab_count[t] = 0;
    if (t < LOCAL_HDR_TOTAL) a.local_hdr[t] = 0;
}

__global__ void slab_build_kernel(SasaArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float4 p = a.atoms[i];
    if (!slab_append(a, i, slab_bucket(a, p), p)) {
        // flagged in both counter sets: whichever the next recount uses reports it
        atomicOr(&a.local_hdr[2], 1);
        atomicOr(&a.local_hdr[LOCAL_HDR_INTS + 2], 1);
    }
}

// The whole cell-list work of a local recount on the slab, one thread per atom: the stage's
// moved atoms written into the atom array and moved in the slab (in place when the bucket is
// unchanged, else a tombstone at the old record and an append to the new bucket), every atom
// inside the box listed, and the other counter set cleared for the next recount.  The moved
// atoms' old records are replaced only by their own threads, and appends take distinct
// records by the atomic, so no two threads write one record.
__global__ void slab_update_kernel(SasaArgs a)
{
    const LocalStage& st = *a.local_stage;
    int* hdr = a.local_hdr + LOCAL_HDR_INTS * st.parity;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < LOCAL_HDR_INTS) a.local_hdr[LOCAL_HDR_INTS * (1 - st.parity) + i] = 0;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < a.n) {
        const int m = i - st.start;
        if (m >= 0 && m < st.len && m < a.patch_cap) {
            const float4* moved = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(a.local_stage) +
                                                                  LOCAL_STAGE_HEADER);
            p = moved[m];
            a.atoms_w[i] = p;
            const int b = slab_bucket(a, p);
            int s_old = a.slab_pos[i];
            if (s_old >= a.slab_buckets * SLAB_CAP) s_old = -1;  // (never: slab_pos indexes the slab)
            if (s_old >= 0 && s_old / SLAB_CAP == b) {
                a.slab[s_old] = p;
            } else {
                if (s_old >= 0) a.slab[s_old] = make_float4(0.f, 0.f, 0.f, __uint_as_float(0xffffffffu));
                if (!slab_append(a, i, b, p)) atomicOr(&hdr[2], 1);
            }
        } else {
            p = a.atoms[i];
        }
    }
    list_if_in_box(a, hdr, st, i, i < a.n, p);
}

constexpr int SASA_MAX_DPL = 8;  // dots per lane (n_dots <= 512)

// Exposed-dot count of atom i, one wave.
//
// Slot t of lane l holds dot l + 64 t (DPL slots).  The live dots of slot t are a 64-bit
// lane mask, the same in every lane (SGPRs): a test is one compare whose result clears
// mask bits, and a slot whose dots are all occluded is skipped.
//
// Candidates: the 27 cells around the atom's cell (lanes 0..26 hold their bucket ranges)
// are taken 64 at a time, one per lane: each lane loads its candidate (positions in bucket
// order, copied there by the place kernel) and keeps it if it lies within r_i + r_j + 2 p
// (+0.01) of the atom.  The kept ones are staged in LDS (cand, 64 entries) and read back as
// broadcasts, two at a time, to test every live dot.  A dot once occluded stays so and the
// test per (dot, candidate) pair is the same expression in the same order whatever the
// visiting order, so the count does not depend on it; the wave stops once no dot is live.
// (Two or four atoms per wave for small spheres, all their cells scanned by the one wave,
// measured slower at 32 dots: 72 against 64 us for 20,000 atoms.)
template <int DPL>
__device__ __forceinline__ int sasa_count_atom(const SasaArgs& a, int i, float4* cand)
{
    const int lane = threadIdx.x;
    const float4 pi = a.atoms[i];
    const uint32_t ti = __float_as_uint(pi.w);
    if (ti >= (uint32_t)a.n_types) return 0;  // uniform: no dots
    const float rext_i = a.rext[ti];
    float px[DPL], py[DPL], pz[DPL];
    uint64_t lm[DPL];
#pragma unroll
    for (int t = 0; t < DPL; ++t) {
        const int d = lane + 64 * t;
        const bool ok = d < a.n_dots;
        const float4 tp = ok ? a.dots[(size_t)ti * a.n_dots + d] : make_float4(0.f, 0.f, 0.f, 0.f);
        px[t] = tp.x + pi.x;
        py[t] = tp.y + pi.y;
        pz[t] = tp.z + pi.z;
        lm[t] = __ballot(ok);
    }
    auto any_live = [&]() {
        uint64_t m = 0;
#pragma unroll
        for (int t = 0; t < DPL; ++t) m |= lm[t];
        return m != 0;
    };
    auto test = [&](const float4 cj) {
#pragma unroll
        for (int t = 0; t < DPL; ++t) {
            if (lm[t]) {
                const bool o = dot_occluded(px[t] - cj.x, py[t] - cj.y, pz[t] - cj.z, cj.w);
                lm[t] &= ~__ballot(o);
            }
        }
    };
    // lane c < 27: neighbour cell c's bucket range
    const int cx = cell_coord(pi.x, a.inv_cell), cy = cell_coord(pi.y, a.inv_cell), cz = cell_coord(pi.z, a.inv_cell);
    int cstart = 0, ccount = 0;
    const float4* __restrict__ src = a.slab ? a.slab : a.sorted_atoms;
    if (lane < 27) {
        const int oz = lane / 9 - 1, oy = (lane / 3) % 3 - 1, ox = lane % 3 - 1;
        if (a.slab) {  // uniform
            const int b = cell_hash(cx + ox, cy + oy, cz + oz, a.slab_buckets);
            cstart = b * SLAB_CAP;
            ccount = min(a.slab_count[b], SLAB_CAP);
        } else {
            const int b = cell_hash(cx + ox, cy + oy, cz + oz, a.n_buckets);
            cstart = a.bucket_start[b];
            ccount = a.bucket_start[b + 1] - cstart;
        }
    }
    int incl = ccount;  // inclusive prefix of the counts (lanes >= 27 add 0)
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
    }
    const int excl = incl - ccount;
    const int total = __builtin_amdgcn_readlane(incl, 26);
    // the candidate at list position q: its cell (the first c with incl[c] > q, by a binary
    // search over lanes 0..31: incl does not decrease and lanes 27..31 hold the total) and
    // its record, loaded unconditionally (slot 0 past the list, which `valid` marks) so
    // that nothing waits for the load before its use
    auto load = [&](int q, bool& valid) {
        int c = 0;
#pragma unroll
        for (int step = 16; step >= 1; step >>= 1)
            if (__shfl(incl, c + step - 1, 64) <= q) c += step;
        const int e0 = __shfl(excl, c, 64), s_c = __shfl(cstart, c, 64);
        valid = q < total;
        return src[valid ? s_c + (q - e0) : 0];
    };
    bool next_valid;
    float4 next = load(lane, next_valid);
    for (int base = 0; base < total; base += 64) {
        // this chunk's candidate per lane; the next chunk's lookup and load go out after the
        // prefilter's table loads and before this chunk's tests, which hide their latency
        const float4 pj = next;
        bool pass = false;
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t tj = next_valid ? __float_as_uint(pj.w) : 0xffffffffu;
        if (tj < (uint32_t)a.n_types) {
            const float ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
            const float lim = rext_i + a.rext[tj] + 0.01f;
            pass = ex * ex + ey * ey + ez * ez < lim * lim;
            rec = make_float4(pj.x, pj.y, pj.z, a.r2[tj]);
        }
        next = load(base + 64 + lane, next_valid);  // unconditional: one path for the wait counts
        const uint64_t m = __ballot(pass);
        const int np = __builtin_popcountll(m);
        if (pass) cand[__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = rec;
        __syncthreads();
        int k = 0;
        for (; k + 1 < np; k += 2) {
            const float4 c0 = cand[k], c1 = cand[k + 1];
            test(c0);
            test(c1);
        }
        if (k < np) test(cand[k]);
        __syncthreads();
        if (!any_live()) break;
    }
    int cnt = 0;
#pragma unroll
    for (int t = 0; t < DPL; ++t) cnt += __builtin_popcountll(lm[t]);
    return cnt;
}

// One wave per atom: every atom, the listed ones (subset), or for a local recount the
// listed ones over a fixed grid, each count compared with the stored one.
template <int DPL>
__global__ __launch_bounds__(64) void sasa_dots_kernel(SasaArgs a)
{
    __shared__ float4 cand[64];
    if (a.local_hdr) {
        int* hdr = a.local_hdr + LOCAL_HDR_INTS * a.local_stage->parity;
        const int listed = min(hdr[0], a.n);
        for (int k = blockIdx.x; k < listed; k += gridDim.x) {
            const int i = a.local_list[k];
            const uint32_t c = (uint32_t)sasa_count_atom<DPL>(a, i, cand);
            if (threadIdx.x == 0) {
                if (a.counts[i] != c) {
                    const int e = atomicAdd(&hdr[1], 1);
                    if (e < a.changed_cap) {
                        a.local_changed[2 * e] = i;
                        a.local_changed[2 * e + 1] = (int)c;
                    }
                }
                a.counts[i] = c;
            }
        }
        return;
    }
    const int i = a.subset ? a.subset[blockIdx.x] : (int)blockIdx.x;
    const int cnt = sasa_count_atom<DPL>(a, i, cand);
    if (threadIdx.x == 0) a.counts[i] = (uint32_t)cnt;
}

// ---------------------------------------------------------------------------
// adjoint: grad_i = sum_k w_k dI_k/dx_i
//   dI_k/dx_i = sum_{j != i} 2 F_i F_j [cos(q r) - sinc(q r)] (x_i - x_j) / r^2
// Over the triangle i < j (the pair jobs of the forward): per pair the scalar
//   G_ij = sum_k b_k [cos(q_k r) - sin(q_k r) / (q_k r)],   b_k = 2 w_k F_i,k F_j,k
// is formed once and both ends are updated: the row (lane) accumulates
// +G (x_i - x_j) / r^2, the column the negation, summed over the 64 lanes through
// LDS in tiles of ADJ_TC columns.  cos(q_k r) and sin(q_k r) / r advance by the
// same FP64 recurrence; b_k is formed once per column class run.  Each wave
// (job, q chunk) writes FP32 partial gradients for its 64 rows and its columns;
// adjoint_reduce_kernel sums them per atom in a fixed order.
// ---------------------------------------------------------------------------
constexpr int ADJ_TC = 8;  // columns per LDS reduction tile

// uniform tables read through the constant address space (scalar loads)
typedef const __attribute__((address_space(4))) float cfloat;

template <int K, int M>
__global__ __launch_bounds__(64, 2) void adjoint_tri_kernel(const AdjointArgs a)
{
    const int job_id = blockIdx.x / a.n_chunks;
    const int chunk = blockIdx.x - job_id * a.n_chunks;
    const PairJob jb = a.jobs[job_id];
    const int lane = threadIdx.x;
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const bool row_live = row_ok && cls_i < (uint32_t)a.n_class;
    const uint32_t ci = row_live ? cls_i : 0u;
    const float row_w = row_live ? 1.0f : 0.0f;
    const int k0 = chunk * K;
    const ChunkConst cc = a.chunks[chunk];
    const float* __restrict__ ftab = a.ftab;
    const int ldf = a.ldf;
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
    const double* __restrict__ wq = a.wq + k0;
    const double* __restrict__ invq = a.invq + k0;
    const double half_dq = a.half_dq;

    __shared__ float T[ADJ_TC][3][65];  // rows padded: the column sums read across banks
    double b[K];
#pragma unroll
    for (int k = 0; k < K; ++k) b[k] = 0.0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const ConstCols cols{(cv4f*)jb.cols};
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    const size_t part = (size_t)jb.slot * a.n_chunks + chunk;
    float* __restrict__ colpart = a.col_partial + part * (size_t)a.col_range * 3;
    uint32_t cur = CLS_SKIP;
    for (int t0 = 0; t0 < ncol; t0 += ADJ_TC) {
        const int tn = ncol - t0 < ADJ_TC ? ncol - t0 : ADJ_TC;
        for (int t = 0; t < tn; ++t) {
            const int j = t0 + t;
            const float4 c = cols[j];
            const uint32_t clr = __float_as_uint(c.w);
            const bool col_live = clr < (uint32_t)a.n_class;  // uniform
            const uint32_t cl = col_live ? clr : 0u;
            if (cl != cur) {
                cur = cl;
                const float* __restrict__ fr = ftab + __builtin_amdgcn_readfirstlane((int)cl * ldf + k0);
#pragma unroll
                for (int k = 0; k < K; ++k) b[k] = 2.0 * wq[k] * (double)fi[k] * (double)fr[k];
            }
            // weight: padding rows, records outside the class table and the lower half and
            // diagonal of the diagonal tile give 0
            const float w = col_live ? (j < jd ? (j > lane ? row_w : 0.0f) : row_w) : 0.0f;
            const float dx = ri.x - c.x, dy = ri.y - c.y, dz = ri.z - c.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const PairR pr = pair_r(r2);
            const float rinv = pr.rinv;
            const double rd = (double)pr.r;
            const double sh = sin_series<M>(rd * half_dq);
            const float sa = sin_rev(cc.a_rev * rd), ca = cos_rev(cc.a_rev * rd);
            const float sm = sin_rev(cc.mid_rev * rd), cm = cos_rev(cc.mid_rev * rd);
            const double c2 = fma(-4.0 * sh, sh, 2.0);
            // u_k = sin(q_k r) / r and v_k = cos(q_k r), both obey x_{k+1} = c2 x_k - x_{k-1}
            double up = (double)(sa * rinv), uc = fma(2.0 * (double)(cm * rinv), sh, up);
            double vp = (double)ca, vc = fma(-2.0 * (double)sm, sh, vp);
            double G = fma(b[0], fma(-up, invq[0], vp), 0.0);
            G = fma(b[1], fma(-uc, invq[1], vc), G);
#pragma unroll
            for (int k = 2; k < K; ++k) {
                const double un = fma(c2, uc, -up);
                const double vn = fma(c2, vc, -vp);
                G = fma(b[k], fma(-un, invq[k], vn), G);
                up = uc;
                uc = un;
                vp = vc;
                vc = vn;
            }
            const double gs = G * (double)(rinv * rinv * w);
            const double ex = gs * (double)dx, ey = gs * (double)dy, ez = gs * (double)dz;
            gx += ex;
            gy += ey;
            gz += ez;
            T[t][0][lane] = (float)-ex;
            T[t][1][lane] = (float)-ey;
            T[t][2][lane] = (float)-ez;
        }
        __syncthreads();
        // lane (t, comp) sums its column component over the 64 rows
        if (lane < 3 * tn) {
            const int t = lane / 3, comp = lane - 3 * (lane / 3);
            double s = 0.0;
#pragma unroll 16
            for (int l = 0; l < 64; ++l) s += (double)T[t][comp][l];
            colpart[(size_t)(t0 + t) * 3 + comp] = (float)s;
        }
        __syncthreads();
    }
    float* __restrict__ rowpart = a.row_partial + (part * 64 + lane) * 3;
    rowpart[0] = (float)gx;
    rowpart[1] = (float)gy;
    rowpart[2] = (float)gz;
}

// FP32 variant of the adjoint.  The difference-form recurrence of the forward's FP32
// kernel gives u_k = sin(q_k r) / r and the differences d_k = u_k - u_{k-1}
//   = 2 cos((q_k - dq/2) r) sin(dq r / 2) / r,
// so cos(q_k r) = (d_k + d_{k+1}) r / (2 sin(dq r)) comes with the sine chain:
//   G = sum_k b_k [cos(q_k r) - u_k / q_k] = R sum_{k=0..K} beta_k d_k - sum_k gamma_k u_k
// with R = r / (2 sin(dq r)), beta_k = b_k + b_{k-1}, gamma_k = b_k / q_k (formed once
// per column class run).  Four FP32 ops per (pair, q): the d and u updates and two
// weighted sums.  G and the per-row and per-column gradient sums are FP32 too; the
// fixed-order reduction over jobs runs in FP64.
template <int K, int M>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(K <= 67 ? 3 : 2))) void adjoint32_tri_kernel(
    const AdjointArgs a)
{
    const int job_id = blockIdx.x / a.n_chunks;
    const int chunk = blockIdx.x - job_id * a.n_chunks;
    const PairJob jb = a.jobs[job_id];
    const int lane = threadIdx.x;
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const bool row_live = row_ok && cls_i < (uint32_t)a.n_class;
    const uint32_t ci = row_live ? cls_i : 0u;
    const float row_w = row_live ? 1.0f : 0.0f;
    const int k0 = chunk * K;
    const ChunkConst cc = a.chunks[chunk];
    const float* __restrict__ ftab = a.ftab;
    const int ldf = a.ldf;
    const float h_dq = (float)a.half_dq;

    __shared__ float T[ADJ_TC][3][65];
    // beta, gamma per column class run, from b_k = 2 w_k F_i,k F_class,k
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
    float beta[K + 1], gamma[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        beta[k] = 0.0f;
        gamma[k] = 0.0f;
    }
    beta[K] = 0.0f;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;  // per-row sums of ~N terms: FP32 is ample for the 1e-4 bar
    const ConstCols cols{(cv4f*)jb.cols};
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    const size_t part = (size_t)jb.slot * a.n_chunks + chunk;
    float* __restrict__ colpart = a.col_partial + part * (size_t)a.col_range * 3;
    uint32_t cur = CLS_SKIP;
    // per-column seeds of the difference recurrence and the pair geometry
    struct Col {
        float u, d, delta, R, scale, dx, dy, dz;
    };
    auto seed = [&](const float4 c, float w) {
        Col o;
        o.dx = ri.x - c.x;
        o.dy = ri.y - c.y;
        o.dz = ri.z - c.z;
        const float r2 = fmaf(o.dz, o.dz, fmaf(o.dy, o.dy, o.dx * o.dx));
        const float nzf = r2 > 1e-30f ? 1.0f : 0.0f;
        const PairR pr = pair_r(r2);
        const float rinv = pr.rinv, r = pr.r;
        const double rd = (double)r;
        const float sh = sin_small_f32<M>(r * h_dq);                    // sin(dq r / 2)
        const float ch = __builtin_sqrtf(fmaxf(1.0f - sh * sh, 0.0f));  // cos(dq r / 2)
        const float fa = (float)__builtin_amdgcn_fract(cc.a_rev * rd);  // a r in revolutions
        const float sa = __builtin_amdgcn_sinf(fa);                     // sin(a r)
        const float clo = fmaf(__builtin_amdgcn_cosf(fa), ch, sa * sh);  // cos((a - dq/2) r)
        o.delta = 4.0f * sh * sh;
        // R = r / (2 sin(dq r)) = r / (4 sin(dq r/2) cos(dq r/2)); 0 for r == 0
        o.R = nzf * r * __builtin_amdgcn_rcpf(fmaxf(4.0f * sh * ch, 1e-30f));
        o.u = sa * rinv;
        o.d = (clo + clo) * sh * rinv;
        o.scale = rinv * rinv * w;
        return o;
    };
    for (int t0 = 0; t0 < ncol; t0 += ADJ_TC) {
        const int tn = ncol - t0 < ADJ_TC ? ncol - t0 : ADJ_TC;
        // columns in pairs of one class (two independent recurrence chains per step); a
        // column without a partner runs with a weight-0 copy of itself
#pragma unroll 1
        for (int t = 0; t < tn;) {
            const int j = t0 + t;
            const float4 c = cols[j];
            const uint32_t clr = __float_as_uint(c.w);
            const bool col_live = clr < (uint32_t)a.n_class;  // uniform; a dead record adds 0
            const uint32_t cl = col_live ? clr : 0u;
            if (cl != cur) {
                cur = cl;
                // b_k = F_i,k cb_k with cb_k = 2 w_k F_class,k and gamma_k = F_i,k cg_k,
                // cg_k = cb_k / q_k: the class tables come through scalar loads (uniform
                // address in the constant space) and the row factors in groups of 8, so
                // only beta and gamma stay in vector registers
                const int off = __builtin_amdgcn_readfirstlane((int)cl * ldf + k0);
                cfloat* cbp = (cfloat*)(a.cb + off);
                cfloat* cgp = (cfloat*)(a.cg + off);
                float prev = 0.0f;
                if (a.wframe) {  // uniform: the batch's per-frame weights scale the class tables
                    cfloat* wfp = (cfloat*)(a.wframe + (size_t)__builtin_amdgcn_readfirstlane(jb.pad) * ldf + k0);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        if (k % 8 == 0) asm volatile("" ::: "memory");
                        const float f = fi[k] * wfp[k];
                        const float bk = f * cbp[k];
                        beta[k] = bk + prev;
                        gamma[k] = f * cgp[k];
                        prev = bk;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        if (k % 8 == 0) asm volatile("" ::: "memory");
                        const float f = fi[k];
                        const float bk = f * cbp[k];
                        beta[k] = bk + prev;
                        gamma[k] = f * cgp[k];
                        prev = bk;
                    }
                }
                beta[K] = prev;
            }
            const bool two = t + 1 < tn && __float_as_uint(cols[j + 1].w) == clr;
            const float4 c1 = two ? cols[j + 1] : c;
            const float rw = col_live ? row_w : 0.0f;
            const float w0 = j < jd ? (j > lane ? rw : 0.0f) : rw;
            const float w1 = two ? (j + 1 < jd ? (j + 1 > lane ? rw : 0.0f) : rw) : 0.0f;
            const Col A = seed(c, w0), B = seed(c1, w1);
            float da = A.d, ua = A.u, db = B.d, ub = B.u;
            float Da = beta[0] * da, Ua = gamma[0] * ua, Db = beta[0] * db, Ub = gamma[0] * ub;
#pragma unroll
            for (int k = 1; k < K; ++k) {
                da = fmaf(-A.delta, ua, da);
                db = fmaf(-B.delta, ub, db);
                ua += da;
                ub += db;
                Da = fmaf(beta[k], da, Da);
                Db = fmaf(beta[k], db, Db);
                Ua = fmaf(gamma[k], ua, Ua);
                Ub = fmaf(gamma[k], ub, Ub);
            }
            da = fmaf(-A.delta, ua, da);
            db = fmaf(-B.delta, ub, db);
            Da = fmaf(beta[K], da, Da);
            Db = fmaf(beta[K], db, Db);
            const float gsa = fmaf(A.R, Da, -Ua) * A.scale, gsb = fmaf(B.R, Db, -Ub) * B.scale;
            const float eax = gsa * A.dx, eay = gsa * A.dy, eaz = gsa * A.dz;
            const float ebx = gsb * B.dx, eby = gsb * B.dy, ebz = gsb * B.dz;
            gx += eax + ebx;
            gy += eay + eby;
            gz += eaz + ebz;
            T[t][0][lane] = -eax;
            T[t][1][lane] = -eay;
            T[t][2][lane] = -eaz;
            if (two) {  // uniform
                T[t + 1][0][lane] = -ebx;
                T[t + 1][1][lane] = -eby;
                T[t + 1][2][lane] = -ebz;
            }
            t += two ? 2 : 1;
        }
        __syncthreads();
        if (lane < 3 * tn) {
            const int t = lane / 3, comp = lane - 3 * (lane / 3);
            float s = 0.0f;
#pragma unroll 16
            for (int l = 0; l < 64; ++l) s += T[t][comp][l];
            colpart[(size_t)(t0 + t) * 3 + comp] = s;
        }
        __syncthreads();
    }
    float* __restrict__ rowpart = a.row_partial + (part * 64 + lane) * 3;
    rowpart[0] = gx;
    rowpart[1] = gy;
    rowpart[2] = gz;
}

// The FP32 adjoint with class-pair coefficients (plain layouts).  Atoms are class-sorted, so
// the 64 rows of a job nearly always share one class; then beta_k and gamma_k are the same
// in every lane and come from the (row class, column class, chunk) entry of a.coef through
// scalar loads, as the scalar operand of the vector FMAs.  No table sits in vector registers:
// K reaches 128 (one chunk, one seed per pair at 128 q) and more waves fit a SIMD than in
// adjoint32_tri_kernel.  A job whose rows hold several classes runs each column pair once
// per row class present, and every lane keeps the pass of its own class.  The recurrence,
// the coefficients (built with the same rounding by adjoint_pair_coef_kernel) and the sums
// are those of adjoint32_tri_kernel, so both give the same bits at the same K.
template <int K, int M>
__global__ __launch_bounds__(64) void adjoint32p_tri_kernel(const AdjointArgs a)
{
    constexpr int CS = adjoint_pair_stride(K);
    const int job_id = blockIdx.x / a.n_chunks;
    const int chunk = blockIdx.x - job_id * a.n_chunks;
    const PairJob jb = a.jobs[job_id];
    const int lane = threadIdx.x;
    const bool row_ok = lane < jb.nrow;
    const float4 ri = row_ok ? jb.rows[lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const uint32_t ncl = (uint32_t)a.n_class;
    const bool row_live = row_ok && cls_i < ncl;
    const float row_w = row_live ? 1.0f : 0.0f;
    const uint64_t rows_live = __builtin_amdgcn_ballot_w64(row_live);
    const ChunkConst cc = a.chunks[chunk];
    const float h_dq = (float)a.half_dq;
    const int group = a.coef_grouped ? __builtin_amdgcn_readfirstlane(jb.pad) : 0;
    // entry (group, rc, cl, chunk) = coef + (((group ncl + rc) ncl + cl) n_chunks + chunk) CS
    const size_t group_base = (size_t)group * ncl * ncl;

    __shared__ float T[ADJ_TC][3][65];
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    const ConstCols cols{(cv4f*)jb.cols};
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    const size_t part = (size_t)jb.slot * a.n_chunks + chunk;
    float* __restrict__ colpart = a.col_partial + part * (size_t)a.col_range * 3;
    struct Col {
        float u, d, delta, R, scale, dx, dy, dz;
    };
    auto seed = [&](const float4 c, float w) {
        Col o;
        o.dx = ri.x - c.x;
        o.dy = ri.y - c.y;
        o.dz = ri.z - c.z;
        const float r2 = fmaf(o.dz, o.dz, fmaf(o.dy, o.dy, o.dx * o.dx));
        const float nzf = r2 > 1e-30f ? 1.0f : 0.0f;
        const PairR pr = pair_r(r2);
        const float rinv = pr.rinv, r = pr.r;
        const double rd = (double)r;
        const float sh = sin_small_f32<M>(r * h_dq);                    // sin(dq r / 2)
        const float ch = __builtin_sqrtf(fmaxf(1.0f - sh * sh, 0.0f));  // cos(dq r / 2)
        const float fa = (float)__builtin_amdgcn_fract(cc.a_rev * rd);  // a r in revolutions
        const float sa = __builtin_amdgcn_sinf(fa);                     // sin(a r)
        const float clo = fmaf(__builtin_amdgcn_cosf(fa), ch, sa * sh);  // cos((a - dq/2) r)
        o.delta = 4.0f * sh * sh;
        o.R = nzf * r * __builtin_amdgcn_rcpf(fmaxf(4.0f * sh * ch, 1e-30f));
        o.u = sa * rinv;
        o.d = (clo + clo) * sh * rinv;
        o.scale = rinv * rinv * w;
        return o;
    };
    for (int t0 = 0; t0 < ncol; t0 += ADJ_TC) {
        const int tn = ncol - t0 < ADJ_TC ? ncol - t0 : ADJ_TC;
#pragma unroll 1
        for (int t = 0; t < tn;) {
            const int j = t0 + t;
            const float4 c = cols[j];
            const uint32_t clr = __float_as_uint(c.w);
            const bool col_live = clr < ncl;  // uniform; a dead record adds 0
            const uint32_t cl = col_live ? clr : 0u;
            const bool two = t + 1 < tn && __float_as_uint(cols[j + 1].w) == clr;
            const float4 c1 = two ? cols[j + 1] : c;
            const float rw = col_live ? row_w : 0.0f;
            const float w0 = j < jd ? (j > lane ? rw : 0.0f) : rw;
            const float w1 = two ? (j + 1 < jd ? (j + 1 > lane ? rw : 0.0f) : rw) : 0.0f;
            const Col A = seed(c, w0), B = seed(c1, w1);
            float gsa = 0.0f, gsb = 0.0f;
            // one pass per row class present (uniform loop; nearly always one)
            uint64_t left = rows_live;
#pragma unroll 1
            while (left) {
                const int leader = (int)__builtin_ctzll(left);
                const uint32_t rc = (uint32_t)__builtin_amdgcn_readlane((int)cls_i, leader);
                const bool mine = cls_i == rc;
                left &= ~__builtin_amdgcn_ballot_w64(mine);
                const size_t entry = ((group_base + (size_t)rc * ncl + cl) * a.n_chunks + chunk) * CS;
                cfloat* cp = (cfloat*)(a.coef + entry);  // beta_k at 2 k, gamma_k at 2 k + 1
                float da = A.d, ua = A.u, db = B.d, ub = B.u;
                // the recurrence runs inside each pass: held over from one pass to the next
                // it would take 4 K registers
                asm volatile("" : "+v"(da), "+v"(ua), "+v"(db), "+v"(ub));
                float Da = cp[0] * da, Ua = cp[1] * ua, Db = cp[0] * db, Ub = cp[1] * ub;
#pragma unroll
                for (int k = 1; k < K; ++k) {
                    da = fmaf(-A.delta, ua, da);
                    db = fmaf(-B.delta, ub, db);
                    ua += da;
                    ub += db;
                    Da = fmaf(cp[2 * k], da, Da);
                    Db = fmaf(cp[2 * k], db, Db);
                    Ua = fmaf(cp[2 * k + 1], ua, Ua);
                    Ub = fmaf(cp[2 * k + 1], ub, Ub);
                }
                da = fmaf(-A.delta, ua, da);
                db = fmaf(-B.delta, ub, db);
                Da = fmaf(cp[2 * K], da, Da);
                Db = fmaf(cp[2 * K], db, Db);
                const float pa = fmaf(A.R, Da, -Ua) * A.scale, pb = fmaf(B.R, Db, -Ub) * B.scale;
                gsa = mine ? pa : gsa;
                gsb = mine ? pb : gsb;
            }
            const float eax = gsa * A.dx, eay = gsa * A.dy, eaz = gsa * A.dz;
            const float ebx = gsb * B.dx, eby = gsb * B.dy, ebz = gsb * B.dz;
            gx += eax + ebx;
            gy += eay + eby;
            gz += eaz + ebz;
            T[t][0][lane] = -eax;
            T[t][1][lane] = -eay;
            T[t][2][lane] = -eaz;
            if (two) {  // uniform
                T[t + 1][0][lane] = -ebx;
                T[t + 1][1][lane] = -eby;
                T[t + 1][2][lane] = -ebz;
            }
            t += two ? 2 : 1;
        }
        __syncthreads();
        if (lane < 3 * tn) {
            const int t = lane / 3, comp = lane - 3 * (lane / 3);
            float s = 0.0f;
#pragma unroll 16
            for (int l = 0; l < 64; ++l) s += T[t][comp][l];
            colpart[(size_t)(t0 + t) * 3 + comp] = s;
        }
        __syncthreads();
    }
    float* __restrict__ rowpart = a.row_partial + (part * 64 + lane) * 3;
    rowpart[0] = gx;
    rowpart[1] = gy;
    rowpart[2] = gz;
}

// the class-pair table of adjoint32p_tri_kernel (AdjointPairCoefArgs): one thread per float
__global__ void adjoint_pair_coef_kernel(const AdjointPairCoefArgs a)
{
    const int K = a.K, cs = adjoint_pair_stride(K);
    const size_t total = (size_t)a.groups * a.n_class * a.n_class * a.n_chunks * cs;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int s = (int)(t % cs);
    size_t rest = t / cs;
    const int chunk = (int)(rest % a.n_chunks);
    rest /= a.n_chunks;
    const int cl = (int)(rest % a.n_class);
    rest /= a.n_class;
    const int ci = (int)(rest % a.n_class);
    const int g = (int)(rest / a.n_class);
    const int k0 = chunk * K;
    auto fk = [&](int k) {
        const float f = a.ftab[(size_t)ci * a.ldf + k0 + k];
        return a.wframe ? f * a.wframe[(size_t)g * a.ldf + k0 + k] : f;
    };
    auto bk = [&](int k) { return k >= 0 && k < K ? fk(k) * a.cb[(size_t)cl * a.ldf + k0 + k] : 0.0f; };
    float v = 0.0f;  // beta_k at 2 k (k <= K), gamma_k at 2 k + 1 (k < K), 0 in the padding
    const int k = s >> 1;
    if ((s & 1) == 0 && k <= K)
        v = bk(k) + bk(k - 1);
    else if ((s & 1) == 1 && k < K)
        v = fk(k) * a.cg[(size_t)cl * a.ldf + k0 + k];
    a.coef[t] = v;
}

// per-class adjoint coefficients of the FP32 kernel: cb[c][k] = 2 w_k F_c,k and
// cg[c][k] = cb[c][k] / q_k, formed in FP64 and rounded once
__global__ void adjoint_coef_kernel(const float* __restrict__ ftab, const double* __restrict__ wq,
                                    const double* __restrict__ invq, int n, float* cb, float* cg, int ldf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = t % ldf;
    const double b = 2.0 * wq[k] * (double)ftab[t];
    cb[t] = (float)b;
    cg[t] = (float)(b * invq[k]);
}

// grad of slot s = its row partials (jobs of its row block) + its column partials
// (one job per row block at or before its own), over every q chunk.  One wave per atom:
// lanes take the (job, chunk) items in turn and a fixed shuffle tree adds the lanes, so
// the order is fixed run to run.
__global__ __launch_bounds__(256) void adjoint_reduce_kernel(AdjointReduceArgs a)
{
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= a.n) return;  // whole waves
    const int I = s >> 6, ln = s & 63;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    const int nm = (a.n - 64 * I + a.col_range - 1) / a.col_range;
    const int n_row_items = nm * a.n_chunks;
    for (int t = lane; t < n_row_items; t += 64) {
        const int m = t / a.n_chunks, c = t - m * a.n_chunks;
        const int slot = a.tri_slot[I * a.tri_stride + m];
        const float* v = a.row_partial + (((size_t)slot * a.n_chunks + c) * 64 + ln) * 3;
        g0 += v[0];
        g1 += v[1];
        g2 += v[2];
    }
    const int n_col_items = (I + 1) * a.n_chunks;
    for (int t = lane; t < n_col_items; t += 64) {
        const int J = t / a.n_chunks, c = t - J * a.n_chunks;
        const int rel = s - 64 * J;
        const int m = rel / a.col_range;
        const int off = rel - m * a.col_range;
        const int slot = a.tri_slot[J * a.tri_stride + m];
        const float* v = a.col_partial + (((size_t)slot * a.n_chunks + c) * a.col_range + off) * 3;
        g0 += v[0];
        g1 += v[1];
        g2 += v[2];
    }
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    g2 = wave_sum(g2);
    if (lane == 0) {
        const int o = a.slot_to_caller[s];
        a.grad[3 * o] = (float)g0;
        a.grad[3 * o + 1] = (float)g1;
        a.grad[3 * o + 2] = (float)g2;
    }
}

// adjoint_reduce_kernel over a batch of frames (AdjointBatchReduceArgs)
__global__ __launch_bounds__(256) void adjoint_batch_reduce_kernel(AdjointBatchReduceArgs a)
{
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= a.n) return;  // whole waves
    const int f = a.atom_frame[s];
    const int base = a.frame_off[f], m = a.frame_off[f + 1] - base;
    const int ls = s - base;
    const int stride = (m + a.col_range - 1) / a.col_range;
    const int* __restrict__ tri = a.tri_slot + a.tri_off[f];
    const int I = ls >> 6, ln = ls & 63;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    const int nm = (m - 64 * I + a.col_range - 1) / a.col_range;
    const int n_row_items = nm * a.n_chunks;
    for (int t = lane; t < n_row_items; t += 64) {
        const int mm = t / a.n_chunks, c = t - mm * a.n_chunks;
        const int slot = tri[I * stride + mm];
        const float* v = a.row_partial + (((size_t)slot * a.n_chunks + c) * 64 + ln) * 3;
        g0 += v[0];
        g1 += v[1];
        g2 += v[2];
    }
    const int n_col_items = (I + 1) * a.n_chunks;
    for (int t = lane; t < n_col_items; t += 64) {
        const int J = t / a.n_chunks, c = t - J * a.n_chunks;
        const int rel = ls - 64 * J;
        const int mm = rel / a.col_range;
        const int off = rel - mm * a.col_range;
        const int slot = tri[J * stride + mm];
        const float* v = a.col_partial + (((size_t)slot * a.n_chunks + c) * a.col_range + off) * 3;
        g0 += v[0];
        g1 += v[1];
        g2 += v[2];
    }
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    g2 = wave_sum(g2);
    if (lane == 0) {
        const int o = a.slot_to_caller[s];
        a.grad[3 * o] = (float)g0;
        a.grad[3 * o + 1] = (float)g1;
        a.grad[3 * o + 2] = (float)g2;
    }
}

template <int... Ks>
struct KList {};

constexpr int UNIFORM_KS[] = {8, 16, 24, 32, 40, 48, 50, 56, 64, 67};
// besides powers of two: K that split common grids into equal chunks with little
// padding (256 = 3 x 86 - 2, 200 = 2 x 100, 101 <= 2 x 51, 200 <= 3 x 67)
constexpr int REC32_KS[] = {16, 32, 43, 48, 51, 64, 67, 80, 86, 96, 100};
// run-flush form: only the run accumulators live in registers, so K reaches 128 (256 = 2 x 128)
constexpr int REC32R_KS[] = {64, 86, 96, 100, 128};
constexpr int DIRECT_KS[] = {8, 16, 32, 48, 64};
constexpr int ADJOINT_KS[] = {8, 16, 32, 48, 64};
constexpr int ADJOINT32_KS[] = {16, 32, 40, 50, 64, 67, 100};
constexpr int ADJOINT32P_KS[] = {64, 100, 128, 200};

}  // namespace

int debye_k_count(int path)
{
    return path == PATH_DIRECT    ? (int)(sizeof(DIRECT_KS) / sizeof(int))
           : path == PATH_REC32   ? (int)(sizeof(REC32_KS) / sizeof(int))
           : path == PATH_REC32R  ? (int)(sizeof(REC32R_KS) / sizeof(int))
                                  : (int)(sizeof(UNIFORM_KS) / sizeof(int));
}

int debye_k_value(int path, int i)
{
    return path == PATH_DIRECT    ? DIRECT_KS[i]
           : path == PATH_REC32   ? REC32_KS[i]
           : path == PATH_REC32R  ? REC32R_KS[i]
                                  : UNIFORM_KS[i];
}

int adjoint_k_count(bool fp32)
{
    return fp32 ? (int)(sizeof(ADJOINT32_KS) / sizeof(int)) : (int)(sizeof(ADJOINT_KS) / sizeof(int));
}

int adjoint_k_value(bool fp32, int i)
{
    return fp32 ? ADJOINT32_KS[i] : ADJOINT_KS[i];
}

int adjoint_pair_k_count()
{
    return (int)(sizeof(ADJOINT32P_KS) / sizeof(int));
}

int adjoint_pair_k_value(int i)
{
    return ADJOINT32P_KS[i];
}

namespace {

// series length instantiated for the recurrence kernels: the short one covers
// layouts with dq R up to ~0.2 (most proteins at q <= 0.6), the long one the
// rest of the recurrence range
constexpr int SIN_SHORT = 5;

// the shared-geometry kernel is instantiated for K <= 64 and K % 8 == 0
template <int K>
constexpr bool has_shared_kernel()
{
    return K <= 64 && K % 8 == 0;
}

template <int K, int M, bool SHARED>
struct SharedLaunch {
    static void go(const DebyeArgs& a, int n_jobs, int nw, int ng, hipStream_t s)
    {
        hipLaunchKernelGGL((debye_shared_kernel<K, M>), dim3((unsigned)(n_jobs * ng)), dim3((unsigned)(64 * nw)), 0, s,
                           a);
    }
};

template <int K, int M>
struct SharedLaunch<K, M, false> {
    static void go(const DebyeArgs& a, int n_jobs, int, int, hipStream_t s)
    {
        hipLaunchKernelGGL((debye_uniform_kernel<K, M>), dim3((unsigned)(n_jobs * a.n_chunks)), dim3(64), 0, s, a);
    }
};

template <int K, int M>
void launch_recurrence(const DebyeArgs& a, int n_jobs, hipStream_t s)
{
    if (a.variant == VARIANT_SHARED && a.n_chunks > 1) {
        const int nw = a.n_chunks < SH_MAXW ? a.n_chunks : SH_MAXW;
        const int ng = (a.n_chunks + nw - 1) / nw;
        SharedLaunch<K, M, has_shared_kernel<K>()>::go(a, n_jobs, nw, ng, s);
    } else {
        hipLaunchKernelGGL((debye_uniform_kernel<K, M>), dim3((unsigned)(n_jobs * a.n_chunks)), dim3(64), 0, s, a);
    }
}

template <int K>
void launch_recurrence_k(const DebyeArgs& a, int n_jobs, hipStream_t s)
{
    if (a.sin_terms <= SIN_SHORT)
        launch_recurrence<K, SIN_SHORT>(a, n_jobs, s);
    else
        launch_recurrence<K, SIN_MAX_TERMS>(a, n_jobs, s);
}

template <int K>
void launch_direct_k(const DebyeArgs& a, int n_jobs, hipStream_t s)
{
    const dim3 grid((unsigned)(n_jobs * a.n_chunks)), block(64);
    if (a.phase32 && a.qsplit)
        hipLaunchKernelGGL((debye_direct_kernel<K, true>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((debye_direct_kernel<K, false>), grid, block, 0, s, a);
}

template <int K>
void launch_rec32_k(const DebyeArgs& a, int n_jobs, hipStream_t s)
{
    const int n_items = n_jobs * a.n_chunks;
    if (a.queue && a.persistent_waves > 0 && a.persistent_waves < n_items) {
        hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, s, a.queue, 0, 1LL);
        const dim3 grid((unsigned)a.persistent_waves), block(64);
        if (a.sin_terms <= SIN_SHORT)
            hipLaunchKernelGGL((debye_rec32_persistent<K, SIN_SHORT>), grid, block, 0, s, a, n_items);
        else
            hipLaunchKernelGGL((debye_rec32_persistent<K, SIN_MAX_TERMS>), grid, block, 0, s, a, n_items);
        return;
    }
    const dim3 grid((unsigned)n_items), block(64);
    if (a.sin_terms <= SIN_SHORT)
        hipLaunchKernelGGL((debye_rec32_kernel<K, SIN_SHORT>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((debye_rec32_kernel<K, SIN_MAX_TERMS>), grid, block, 0, s, a);
}

template <int K>
void launch_rec32r_k(const DebyeArgs& a, int n_jobs, hipStream_t s)
{
    const int n_items = n_jobs * a.n_chunks;
    if (a.queue && a.persistent_waves > 0 && a.persistent_waves < n_items) {
        hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, s, a.queue, 0, 1LL);
        const dim3 grid((unsigned)a.persistent_waves), block(64);
        if (a.sin_terms <= SIN_SHORT)
            hipLaunchKernelGGL((debye_rec32r_persistent<K, SIN_SHORT>), grid, block, 0, s, a, n_items);
        else
            hipLaunchKernelGGL((debye_rec32r_persistent<K, SIN_MAX_TERMS>), grid, block, 0, s, a, n_items);
        return;
    }
    const dim3 grid((unsigned)n_items), block(64);
    if (a.sin_terms <= SIN_SHORT)
        hipLaunchKernelGGL((debye_rec32r_kernel<K, SIN_SHORT>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((debye_rec32r_kernel<K, SIN_MAX_TERMS>), grid, block, 0, s, a);
}

}  // namespace

namespace {

template <int K, bool RUN_FLUSH>
int rec32_blocks_per_cu(int sin_terms, bool persistent)
{
    int nb = 0;
    hipError_t e;
    const bool sh = sin_terms <= SIN_SHORT;
    if constexpr (RUN_FLUSH) {
        if (persistent)
            e = sh ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32r_persistent<K, SIN_SHORT>, 64, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32r_persistent<K, SIN_MAX_TERMS>, 64, 0);
        else
            e = sh ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32r_kernel<K, SIN_SHORT>, 64, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32r_kernel<K, SIN_MAX_TERMS>, 64, 0);
    } else {
        if (persistent)
            e = sh ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32_persistent<K, SIN_SHORT>, 64, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32_persistent<K, SIN_MAX_TERMS>, 64, 0);
        else
            e = sh ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32_kernel<K, SIN_SHORT>, 64, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, debye_rec32_kernel<K, SIN_MAX_TERMS>, 64, 0);
    }
    return e == hipSuccess ? nb : 0;
}

}  // namespace

int rec32_blocks_per_cu(int variant, int K, int sin_terms, bool persistent)
{
    if (variant == VARIANT_FP32R) {
        switch (K) {
        case 64: return rec32_blocks_per_cu<64, true>(sin_terms, persistent);
        case 86: return rec32_blocks_per_cu<86, true>(sin_terms, persistent);
        case 96: return rec32_blocks_per_cu<96, true>(sin_terms, persistent);
        case 100: return rec32_blocks_per_cu<100, true>(sin_terms, persistent);
        case 128: return rec32_blocks_per_cu<128, true>(sin_terms, persistent);
        default: return 0;
        }
    }
    if (variant != VARIANT_FP32) return 0;
    switch (K) {
    case 16: return rec32_blocks_per_cu<16, false>(sin_terms, persistent);
    case 32: return rec32_blocks_per_cu<32, false>(sin_terms, persistent);
    case 43: return rec32_blocks_per_cu<43, false>(sin_terms, persistent);
    case 48: return rec32_blocks_per_cu<48, false>(sin_terms, persistent);
    case 51: return rec32_blocks_per_cu<51, false>(sin_terms, persistent);
    case 64: return rec32_blocks_per_cu<64, false>(sin_terms, persistent);
    case 67: return rec32_blocks_per_cu<67, false>(sin_terms, persistent);
    case 80: return rec32_blocks_per_cu<80, false>(sin_terms, persistent);
    case 86: return rec32_blocks_per_cu<86, false>(sin_terms, persistent);
    case 96: return rec32_blocks_per_cu<96, false>(sin_terms, persistent);
    case 100: return rec32_blocks_per_cu<100, false>(sin_terms, persistent);
    default: return 0;
    }
}

hipError_t launch_debye(const DebyeArgs& a, int n_jobs, int K, bool uniform, hipStream_t s)
{
    if (n_jobs <= 0) return hipSuccess;
    if (uniform && a.variant == VARIANT_FP32) {
        switch (K) {
        case 16: launch_rec32_k<16>(a, n_jobs, s); break;
        case 32: launch_rec32_k<32>(a, n_jobs, s); break;
        case 48: launch_rec32_k<48>(a, n_jobs, s); break;
        case 64: launch_rec32_k<64>(a, n_jobs, s); break;
        case 43: launch_rec32_k<43>(a, n_jobs, s); break;
        case 51: launch_rec32_k<51>(a, n_jobs, s); break;
        case 67: launch_rec32_k<67>(a, n_jobs, s); break;
        case 80: launch_rec32_k<80>(a, n_jobs, s); break;
        case 86: launch_rec32_k<86>(a, n_jobs, s); break;
        case 96: launch_rec32_k<96>(a, n_jobs, s); break;
        case 100: launch_rec32_k<100>(a, n_jobs, s); break;
        default: return hipErrorInvalidValue;
        }
    } else if (uniform && a.variant == VARIANT_FP32R) {
        switch (K) {
        case 64: launch_rec32r_k<64>(a, n_jobs, s); break;
        case 86: launch_rec32r_k<86>(a, n_jobs, s); break;
        case 96: launch_rec32r_k<96>(a, n_jobs, s); break;
        case 100: launch_rec32r_k<100>(a, n_jobs, s); break;
        case 128: launch_rec32r_k<128>(a, n_jobs, s); break;
        default: return hipErrorInvalidValue;
        }
    } else if (uniform) {
        switch (K) {
        case 8: launch_recurrence_k<8>(a, n_jobs, s); break;
        case 16: launch_recurrence_k<16>(a, n_jobs, s); break;
        case 24: launch_recurrence_k<24>(a, n_jobs, s); break;
        case 32: launch_recurrence_k<32>(a, n_jobs, s); break;
        case 40: launch_recurrence_k<40>(a, n_jobs, s); break;
        case 48: launch_recurrence_k<48>(a, n_jobs, s); break;
        case 50: launch_recurrence_k<50>(a, n_jobs, s); break;
        case 56: launch_recurrence_k<56>(a, n_jobs, s); break;
        case 64: launch_recurrence_k<64>(a, n_jobs, s); break;
        case 67: launch_recurrence_k<67>(a, n_jobs, s); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (K) {
        case 8: launch_direct_k<8>(a, n_jobs, s); break;
        case 16: launch_direct_k<16>(a, n_jobs, s); break;
        case 32: launch_direct_k<32>(a, n_jobs, s); break;
        case 48: launch_direct_k<48>(a, n_jobs, s); break;
        case 64: launch_direct_k<64>(a, n_jobs, s); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

namespace {

// slot chunks per output: enough blocks to spread the reads over the CUs, at least
// 256 slots per chunk
int reduce_splits(int n_q, int n_out, int max_slots)
{
    const int base = ((n_q + 63) / 64) * (n_out > 0 ? n_out : 1);
    const int by_blocks = (1024 + base - 1) / base;
    const int by_slots = (max_slots + 255) / 256;
    const int n = by_blocks < by_slots ? by_blocks : by_slots;
    return n > 1 ? n : 1;
}

}  // namespace

size_t reduce_scratch_doubles(int n_q, int n_out, int max_slots)
{
    const int ns = reduce_splits(n_q, n_out, max_slots);
    return ns > 1 ? (size_t)ns * (size_t)n_out * (size_t)n_q : 0;
}

hipError_t launch_reduce(const double* partials, int ldf, int n_q, const int* slot_begin, int n_out, double* out,
                         int out_ld, int accumulate, const double* q0_val, int q0_index, int max_slots,
                         double* scratch, hipStream_t s)
{
    if (n_out <= 0 || n_q <= 0) return hipSuccess;
    const int ns = reduce_splits(n_q, n_out, max_slots);
    if (ns > 1 && scratch) {
        const int chunk = (max_slots + ns - 1) / ns;
        hipLaunchKernelGGL(reduce_part_kernel, dim3((unsigned)((n_q + 63) / 64), (unsigned)n_out, (unsigned)ns),
                           dim3(64 * RED_GROUPS), 0, s, partials, ldf, n_q, slot_begin, chunk, ns, scratch);
        hipLaunchKernelGGL(reduce_final_kernel, dim3((unsigned)((n_q + 63) / 64), (unsigned)n_out), dim3(64), 0, s,
                           scratch, ns, n_q, out, out_ld, accumulate, q0_val, q0_index);
        return hipGetLastError();
    }
    const dim3 grid((unsigned)((n_q + 63) / 64), (unsigned)n_out), block(64 * RED_GROUPS);
    hipLaunchKernelGGL(reduce_kernel, grid, block, 0, s, partials, ldf, n_q, slot_begin, out, out_ld, accumulate,
                       q0_val, q0_index);
    return hipGetLastError();
}

hipError_t launch_frame_mean(const double* in, int n_frames, int n_q, double* out, hipStream_t s)
{
    if (n_q <= 0) return hipSuccess;
    hipLaunchKernelGGL(frame_mean_kernel, dim3((unsigned)((n_q + 255) / 256)), dim3(256), 0, s, in, n_frames, n_q,
                       out);
    return hipGetLastError();
}

hipError_t launch_scatter_slots(float4* cols, int* s2c, const int* slot, const float4* vals, const int* callers, int m,
                                hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_slots_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, cols, s2c, slot, vals,
                       callers, m);
    return hipGetLastError();
}

hipError_t launch_scatter_value(float4* cols, const int* slot, float4 v, int m, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_value_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, cols, slot, v, m);
    return hipGetLastError();
}

hipError_t launch_gather_cols(float4* cols, const float4* bodies, const int* src, const uint32_t* counts, int d1,
                              int n, const GatherGuard& g, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_cols_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cols, bodies, src,
                       counts, d1, n, g);
    return hipGetLastError();
}

hipError_t launch_adjoint_weights(const double* w, int n_q, int q0_index, int ldfa, double* wq, hipStream_t s)
{
    if (ldfa <= 0) return hipSuccess;
    hipLaunchKernelGGL(adjoint_weights_kernel, dim3((unsigned)((ldfa + 255) / 256)), dim3(256), 0, s, w, n_q, q0_index,
                       ldfa, wq);
    return hipGetLastError();
}

hipError_t launch_chi2_weights(const double* iq, const float* i_exp, const float* sigma, int n_q, int q0_index,
                               int ldfa, double* out, double* wq, hipStream_t s)
{
    hipLaunchKernelGGL(chi2_weights_kernel, dim3(1), dim3(1024), 0, s, iq, i_exp, sigma, n_q, q0_index, ldfa, out, wq);
    return hipGetLastError();
}

hipError_t launch_class_sort(const ClassSortArgs& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    if (a.n_class < 1 || a.n_class + 1 > CLASS_SORT_MAX_KEYS) return hipErrorInvalidValue;
    const int n_blk = (a.n + CS_BLOCK - 1) / CS_BLOCK;
    hipLaunchKernelGGL(class_hist_kernel, dim3((unsigned)n_blk), dim3(CS_BLOCK), (size_t)(a.n_class + 1) * sizeof(int),
                       s, a);
    hipLaunchKernelGGL(class_scan_kernel, dim3(1), dim3(1024), 0, s, a, n_blk);
    hipLaunchKernelGGL(class_scatter_kernel, dim3((unsigned)n_blk), dim3(CS_BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_q0(const float4* cols, int n, const double* f0col, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(q0_kernel, dim3(1), dim3(1024), 0, s, cols, n, f0col, out);
    return hipGetLastError();
}

hipError_t launch_copy(float4* dst, const float4* src, long long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_probe_kernel, dim3(4096), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

namespace {

// One workgroup of 64 lanes per (row tile I, column tile J >= I) of the triangle: lane =
// row atom, the 64 columns of tile J staged in LDS; on the diagonal tile only j > i.
// Global atomics without return (fire and forget into L2).
__global__ __launch_bounds__(64) void hist_pairs_kernel(HistArgs a, int n_tiles)
{
    // blockIdx.x -> (I, J) with I <= J, row-major over the upper triangle
    const long long t = blockIdx.x;
    int I = (int)((2.0 * n_tiles + 1.0 - sqrt((2.0 * n_tiles + 1.0) * (2.0 * n_tiles + 1.0) - 8.0 * (double)t)) / 2.0);
    long long start = (long long)I * n_tiles - (long long)I * (I - 1) / 2;
    while (I > 0 && start > t) {
        --I;
        start = (long long)I * n_tiles - (long long)I * (I - 1) / 2;
    }
    while (start + (n_tiles - I) <= t) {
        start += n_tiles - I;
        ++I;
    }
    const int J = I + (int)(t - start);
    const int lane = threadIdx.x;
    __shared__ float4 cj[64];
    __shared__ unsigned int cntj[64];
    const int j0 = J * 64;
    const int ncol = a.n - j0 < 64 ? a.n - j0 : 64;
    if (lane < ncol) {
        cj[lane] = a.atoms[j0 + lane];
        cntj[lane] = a.counts ? a.counts[j0 + lane] : 0u;
    }
    __syncthreads();
    const int i = I * 64 + lane;
    if (i >= a.n) return;
    const float4 pi = a.atoms[i];
    const int ti = (int)__float_as_uint(pi.w);
    const unsigned int ci = a.counts ? a.counts[i] : 0u;
    const int jstart = I == J ? lane + 1 : 0;
    for (int jj = jstart; jj < ncol; ++jj) {
        const float4 pj = cj[jj];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r = __builtin_sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        int b = (int)(r * a.inv_dr);
        b = b < a.n_bins - 1 ? b : a.n_bins - 1;
        const int tj = (int)__float_as_uint(pj.w);
        atomicAdd(&a.h0[(size_t)a.pair_index[ti * a.n_types + tj] * a.n_bins + b], 1u);
        if (a.counts) {
            const unsigned int cj_ = cntj[jj];
            atomicAdd(&a.h1[(size_t)tj * a.n_bins + b], (unsigned long long)ci);
            atomicAdd(&a.h1[(size_t)ti * a.n_bins + b], (unsigned long long)cj_);
            atomicAdd(&a.h2[b], (unsigned long long)ci * cj_);
        }
    }
}

// Plain histograms through LDS: one workgroup of HIST_ROWS lanes per (tile, bin window),
// one row per lane, the tile's columns read in turn by every lane (uniform loads); each
// pair's bin (the same expression as hist_pairs_kernel, so the counts are identical)
// is counted into the workgroup's LDS window with an LDS atomic, and the window's
// nonzero bins are added to the global histogram at the end.
__global__ __launch_bounds__(HIST_ROWS) void hist_tiles_kernel(const float4* __restrict__ atoms,
                                                               const HistTile* __restrict__ tiles, float inv_dr,
                                                               int n_bins, unsigned int* __restrict__ h0)
{
    __shared__ unsigned int win[HIST_WINDOW];
    const HistTile tl = tiles[blockIdx.x];
    const int b0 = blockIdx.y * HIST_WINDOW;
    const int tid = threadIdx.x;
    for (int b = tid; b < HIST_WINDOW; b += HIST_ROWS) win[b] = 0u;
    __syncthreads();
    const bool row_ok = tid < tl.nr;
    const int i = tl.r0 + tid;
    const float4 pi = row_ok ? atoms[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int jmin = tl.diag ? i + 1 : 0;  // same segment: only j > i
    const int c1 = tl.c0 + tl.nc;
    for (int j = tl.c0; j < c1; ++j) {
        const float4 pj = atoms[j];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r = __builtin_sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        int b = (int)(r * inv_dr);
        b = b < n_bins - 1 ? b : n_bins - 1;
        const int bw = b - b0;
        if (row_ok && j >= jmin && bw >= 0 && bw < HIST_WINDOW) atomicAdd(&win[bw], 1u);
    }
    __syncthreads();
    unsigned int* __restrict__ dst = h0 + (size_t)tl.pair * n_bins + b0;
    const int nb = n_bins - b0 < HIST_WINDOW ? n_bins - b0 : HIST_WINDOW;
    for (int b = tid; b < nb; b += HIST_ROWS) {
        const unsigned int v = win[b];
        if (v) atomicAdd(&dst[b], v);
    }
}

// Transform of the histograms over bin ranges of HIST_RANGE bins: grid (q blocks,
// histograms, ranges), lane = q index.  hw holds hist[x][b] / r_b (r_b the bin centre);
// partial[z][x][k] = sum over the range of hw[x][b] sin(q_k r_b), the sine advanced
// along b by the FP64 three-term recurrence (angle q dr per bin) from seeds taken
// directly at the range start; at q = 0 the sum is of hw[x][b] r_b = hist[x][b].
constexpr int HIST_RANGE = 512;

__global__ __launch_bounds__(64) void hist_transform_kernel(const double* __restrict__ hw, int n_bins, double dr,
                                                            const double* __restrict__ q, int n_q, int n_hist,
                                                            double* __restrict__ partial)
{
    const int x = blockIdx.y, z = blockIdx.z;
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= n_q) return;
    const int b0 = z * HIST_RANGE;
    const int b1 = b0 + HIST_RANGE < n_bins ? b0 + HIST_RANGE : n_bins;
    const double* __restrict__ h = hw + (size_t)x * n_bins;
    const double qk = q[k];
    double acc = 0.0;
    if (qk == 0.0) {
        for (int b = b0; b < b1; ++b) acc = fma(h[b], (b + 0.5) * dr, acc);
    } else {
        const double th = qk * dr;
        const double c2 = 2.0 * cos(th);
        double sp = sin(qk * (b0 - 0.5) * dr), sc = sin(qk * (b0 + 0.5) * dr);  // sin(q r_{b0-1}), sin(q r_b0)
        for (int b = b0; b < b1; ++b) {
            acc = fma(h[b], sc, acc);
            const double sn = fma(c2, sc, -sp);
            sp = sc;
            sc = sn;
        }
    }
    partial[((size_t)z * n_hist + x) * n_q + k] = acc;
}

// out[x][k] = sum over ranges of partial (fixed order) / q_k (1 / 1 at q = 0)
__global__ void hist_range_sum_kernel(const double* __restrict__ partial, int n_ranges, int n_hist,
                                      const double* __restrict__ q, int n_q, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_hist * n_q) return;
    const int k = t % n_q;
    double s = 0.0;
    for (int z = 0; z < n_ranges; ++z) s += partial[(size_t)z * n_hist * n_q + t];
    out[t] = q[k] == 0.0 ? s : s / q[k];
}

// dst[i] = src[i] / r_b with r_b = (b + 0.5) dr the centre of bin b = i mod n_bins
__global__ void widen_u32_kernel(const unsigned int* src, double* dst, long long n, int n_bins, double dr)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i] / (((double)(i % n_bins) + 0.5) * dr);
}

__global__ void widen_u64_kernel(const unsigned long long* src, double* dst, long long n, int n_bins, double dr)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i] / (((double)(i % n_bins) + 0.5) * dr);
}

}  // namespace

hipError_t launch_hist_pairs(const HistArgs& a, hipStream_t s)
{
    if (a.n <= 1) return hipSuccess;
    const int T = (a.n + 63) / 64;
    const long long n_blocks = (long long)T * (T + 1) / 2;
    hipLaunchKernelGGL(hist_pairs_kernel, dim3((unsigned)n_blocks), dim3(64), 0, s, a, T);
    return hipGetLastError();
}

hipError_t launch_hist_tiles(const float4* sorted_atoms, const HistTile* tiles, int n_tiles, float inv_dr,
                             int n_bins, unsigned int* h0, hipStream_t s)
{
    if (n_tiles <= 0 || n_bins <= 0) return hipSuccess;
    const unsigned n_win = (unsigned)((n_bins + HIST_WINDOW - 1) / HIST_WINDOW);
    hipLaunchKernelGGL(hist_tiles_kernel, dim3((unsigned)n_tiles, n_win), dim3(HIST_ROWS), 0, s, sorted_atoms, tiles,
                       inv_dr, n_bins, h0);
    return hipGetLastError();
}

size_t hist_transform_scratch(int n_hist, int n_bins, int n_q)
{
    return (size_t)((n_bins + HIST_RANGE - 1) / HIST_RANGE) * n_hist * n_q;
}

hipError_t launch_hist_transform(const double* hw, int n_hist, int n_bins, double dr, const double* q, int n_q,
                                 double* partial, double* out, hipStream_t s)
{
    if (n_hist <= 0 || n_q <= 0 || n_bins <= 0) return hipSuccess;
    const int n_ranges = (n_bins + HIST_RANGE - 1) / HIST_RANGE;
    hipLaunchKernelGGL(hist_transform_kernel, dim3((unsigned)((n_q + 63) / 64), (unsigned)n_hist, (unsigned)n_ranges),
                       dim3(64), 0, s, hw, n_bins, dr, q, n_q, n_hist, partial);
    const int nt = n_hist * n_q;
    hipLaunchKernelGGL(hist_range_sum_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, partial, n_ranges,
                       n_hist, q, n_q, out);
    return hipGetLastError();
}

hipError_t launch_widen_u32(const unsigned int* src, double* dst, long long n, int n_bins, double dr, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(widen_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n, n_bins, dr);
    return hipGetLastError();
}

hipError_t launch_widen_u64(const unsigned long long* src, double* dst, long long n, int n_bins, double dr,
                            hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(widen_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n, n_bins, dr);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int* p, int value, long long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_slab_build(const SasaArgs& a, hipStream_t s)
{
    if (!a.slab || !a.local_hdr || a.slab_buckets <= 0 || (a.slab_buckets & (a.slab_buckets - 1)))
        return hipErrorInvalidValue;
    const int nc = a.slab_buckets > LOCAL_HDR_TOTAL ? a.slab_buckets : LOCAL_HDR_TOTAL;
    hipLaunchKernelGGL(slab_clear_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, a);
    if (a.n > 0) hipLaunchKernelGGL(slab_build_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sasa(const SasaArgs& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    if (a.n_dots <= 0 || a.n_dots > 64 * SASA_MAX_DPL) return hipErrorInvalidValue;
    if (a.local_hdr && !a.local_stage) return hipErrorInvalidValue;
    if (a.slab && (!a.local_hdr || a.slab_buckets <= 0 || (a.slab_buckets & (a.slab_buckets - 1))))
        return hipErrorInvalidValue;
    if (a.slab) {
        // local recount on the slab: one update kernel (patch, slab, list), then the dots
        const int nt = a.n > LOCAL_HDR_INTS ? a.n : LOCAL_HDR_INTS;
        hipLaunchKernelGGL(slab_update_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a);
    } else if (a.patch_cap > 0) {
        const long long nt = (long long)a.n_buckets + 1 > a.patch_cap ? (long long)a.n_buckets + 1 : a.patch_cap;
        hipLaunchKernelGGL(sasa_patch_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a);
    } else {
        hipError_t e = launch_fill_i32(a.bucket_count, 0, (long long)a.n_buckets + 1, s);
        if (e != hipSuccess) return e;
    }
    if (!a.slab) {
        const dim3 g((unsigned)((a.n + 255) / 256)), b(256);
        if (a.n_buckets < SCAN_TILE || (a.n_buckets & (a.n_buckets - 1))) return hipErrorInvalidValue;
        hipLaunchKernelGGL(sasa_bin_kernel, g, b, 0, s, a);
        hipLaunchKernelGGL(sasa_scan_tiles_kernel, dim3((unsigned)(a.n_buckets / SCAN_TILE)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(sasa_scan_top_kernel, dim3(1), dim3(1024), 0, s, a);
        hipLaunchKernelGGL(sasa_scan_add_kernel, dim3((unsigned)(a.n_buckets / 256)), dim3(256), 0, s, a);
        hipLaunchKernelGGL(sasa_place_kernel, g, b, 0, s, a);
    }
    // a local recount loops over its list with a fixed grid (the list's length is on the device)
    const int n_waves = a.local_hdr ? (a.n < 8192 ? a.n : 8192) : a.subset ? a.n_subset : a.n;
    if (n_waves <= 0) return hipGetLastError();
    const dim3 gd((unsigned)n_waves), bd(64);
    const int dpl = (a.n_dots + 63) / 64;
    if (dpl <= 1)
        hipLaunchKernelGGL(sasa_dots_kernel<1>, gd, bd, 0, s, a);
    else if (dpl <= 2)
        hipLaunchKernelGGL(sasa_dots_kernel<2>, gd, bd, 0, s, a);
    else if (dpl <= 4)
        hipLaunchKernelGGL(sasa_dots_kernel<4>, gd, bd, 0, s, a);
    else
        hipLaunchKernelGGL(sasa_dots_kernel<8>, gd, bd, 0, s, a);
    return hipGetLastError();
}

namespace {

template <int K>
void launch_adjoint_k(const AdjointArgs& a, int n_jobs, hipStream_t s)
{
    const dim3 grid((unsigned)(n_jobs * a.n_chunks)), block(64);
    if (a.sin_terms <= SIN_SHORT)
        hipLaunchKernelGGL((adjoint_tri_kernel<K, SIN_SHORT>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((adjoint_tri_kernel<K, SIN_MAX_TERMS>), grid, block, 0, s, a);
}

template <int K>
void launch_adjoint32_k(const AdjointArgs& a, int n_jobs, hipStream_t s)
{
    const dim3 grid((unsigned)(n_jobs * a.n_chunks)), block(64);
    if (a.sin_terms <= SIN_SHORT)
        hipLaunchKernelGGL((adjoint32_tri_kernel<K, SIN_SHORT>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((adjoint32_tri_kernel<K, SIN_MAX_TERMS>), grid, block, 0, s, a);
}

template <int K>
void launch_adjoint32p_k(const AdjointArgs& a, int n_jobs, hipStream_t s)
{
    const dim3 grid((unsigned)(n_jobs * a.n_chunks)), block(64);
    if (a.sin_terms <= SIN_SHORT)
        hipLaunchKernelGGL((adjoint32p_tri_kernel<K, SIN_SHORT>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((adjoint32p_tri_kernel<K, SIN_MAX_TERMS>), grid, block, 0, s, a);
}

}  // namespace

hipError_t launch_adjoint_pair_coef(const AdjointPairCoefArgs& a, hipStream_t s)
{
    if (a.K <= 0 || a.n_chunks <= 0 || a.ldf < a.n_chunks * a.K) return hipErrorInvalidValue;
    const size_t total = (size_t)a.groups * a.n_class * a.n_class * a.n_chunks * adjoint_pair_stride(a.K);
    if (total == 0) return hipSuccess;
    if ((total + 255) / 256 > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adjoint_pair_coef_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adjoint_coef(const float* ftab, const double* wq, const double* invq, int n_class, int ldf, float* cb,
                               float* cg, hipStream_t s)
{
    const int n = n_class * ldf;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(adjoint_coef_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ftab, wq, invq, n, cb, cg,
                       ldf);
    return hipGetLastError();
}

hipError_t launch_adjoint(const AdjointArgs& a, int n_jobs, int K, hipStream_t s)
{
    if (n_jobs <= 0) return hipSuccess;
    if (a.n_class <= 0) return hipErrorInvalidValue;         // every record would count as outside the table
    if (a.wframe && !a.fp32) return hipErrorInvalidValue;  // per-frame weights: FP32 kernel only
    if (a.coef) {  // class-pair coefficients: FP32 only
        if (!a.fp32) return hipErrorInvalidValue;
        switch (K) {
        case 64: launch_adjoint32p_k<64>(a, n_jobs, s); break;
        case 100: launch_adjoint32p_k<100>(a, n_jobs, s); break;
        case 128: launch_adjoint32p_k<128>(a, n_jobs, s); break;
        case 200: launch_adjoint32p_k<200>(a, n_jobs, s); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (a.fp32) {
        switch (K) {
        case 16: launch_adjoint32_k<16>(a, n_jobs, s); break;
        case 32: launch_adjoint32_k<32>(a, n_jobs, s); break;
        case 40: launch_adjoint32_k<40>(a, n_jobs, s); break;
        case 50: launch_adjoint32_k<50>(a, n_jobs, s); break;
        case 64: launch_adjoint32_k<64>(a, n_jobs, s); break;
        case 67: launch_adjoint32_k<67>(a, n_jobs, s); break;
        case 100: launch_adjoint32_k<100>(a, n_jobs, s); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (K) {
    case 8: launch_adjoint_k<8>(a, n_jobs, s); break;
    case 16: launch_adjoint_k<16>(a, n_jobs, s); break;
    case 32: launch_adjoint_k<32>(a, n_jobs, s); break;
    case 48: launch_adjoint_k<48>(a, n_jobs, s); break;
    case 64: launch_adjoint_k<64>(a, n_jobs, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adjoint_reduce(const AdjointReduceArgs& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)((a.n + 3) / 4)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adjoint_batch_reduce(const AdjointBatchReduceArgs& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(adjoint_batch_reduce_kernel, dim3((unsigned)((a.n + 3) / 4)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace rsx
