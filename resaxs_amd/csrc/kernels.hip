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
// grid; the predicate is evaluated in FP32 without contraction so the integer
// counts equal the CPU restatement bit for bit.

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

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ float wave_sum_f(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v)
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

// sin(x) for 0 <= x <= 1.5 in FP64 (odd Taylor series to x^17; truncation
// below 2e-14 relative).  The recurrence's frequency 2 cos(dq r) = 2 - 4 sin^2(dq r / 2)
// must be far more accurate than v_sin's ~4e-7 absolute error: a relative
// error e in sin(dq r / 2) shifts every later term's phase by ~ e q r.
__device__ __forceinline__ double sin_poly(double x)
{
    const double x2 = x * x;
    double p = 2.8114572543455206e-15;
    p = fma(p, x2, -7.647163731819816e-13);
    p = fma(p, x2, 1.6059043836821613e-10);
    p = fma(p, x2, -2.505210838544172e-08);
    p = fma(p, x2, 2.7557319223985893e-06);
    p = fma(p, x2, -1.984126984126984e-04);
    p = fma(p, x2, 8.333333333333333e-03);
    p = fma(p, x2, -1.6666666666666666e-01);
    p = fma(p, x2, 1.0);
    return x * p;
}

template <bool WEIGHTED>
__device__ __forceinline__ Seeds make_seeds(float xi, float yi, float zi, const float4 c,
                                            const ChunkConst& cc, double half_dq, float dq, float w)
{
    // branch-free: every value stays finite and r == 0 is blended in with a 0/1 mask,
    // so the seeds of the next column pair can be scheduled into the recurrence block
    const float dx = xi - c.x, dy = yi - c.y, dz = zi - c.z;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float nzf = r2 > 1e-30f ? 1.0f : 0.0f;
    const float rinv = __builtin_amdgcn_rsqf(fmaxf(r2, 1e-30f)) * nzf;
    const float r = r2 * rinv;
    const double rd = (double)r;
    const double sh = sin_poly(rd * half_dq);  // sin(dq r / 2)
    const float sa = sin_rev(cc.a_rev * rd);   // sin(a r)
    const float cm = cos_rev(cc.mid_rev * rd); // cos((a + dq/2) r)
    const float wr = WEIGHTED ? w * rinv : rinv;
    const float wz = (WEIGHTED ? w : 1.0f) * (1.0f - nzf);
    // r == 0 (the i == j diagonal or coincident atoms): u_k -> q_k, linear in k.
    const float f0 = fmaf(sa, wr, wz * cc.a);
    const double fd = fma(2.0 * (double)(cm * wr), sh, (double)(wz * dq));
    Seeds s;
    s.c2 = fma(-4.0 * sh, sh, 2.0);
    s.u0 = (double)f0;
    s.u1 = s.u0 + fd;
    return s;
}

template <int K>
__device__ __forceinline__ void run1(double (&acc)[K], const Seeds& s)
{
    acc[0] += s.u0;
    acc[1] += s.u1;
    double up = s.u0, uc = s.u1;
#pragma unroll
    for (int k = 2; k < K; ++k) {
        const double un = fma(s.c2, uc, -up);
        acc[k] += un;
        up = uc;
        uc = un;
    }
}

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

template <int K>
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
        const float* __restrict__ fr = ftab + (size_t)cls * ldf + k0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            X[k] = fmaf(fr[k], (float)acc[k], X[k]);
            acc[k] = 0.0;
        }
    };

    const float4* __restrict__ cols = jb.cols;
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    // one weighted pair step per iteration: the second column is the next one
    // when it has the same class, otherwise a zero-weight copy.  Outside the
    // diagonal tile the weights are wave-uniform (0 for skipped records).
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
            // diagonal tile: weights 1 (j > i), 1/2 (j == i), 0 (j < i)
            w0 *= j > lane ? 1.0f : (j == lane ? 0.5f : 0.0f);
            w1 *= j + 1 > lane ? 1.0f : (j + 1 == lane ? 0.5f : 0.0f);
        }
        run2<K>(acc, make_seeds<true>(ri.x, ri.y, ri.z, c0, cc, h_rev, dq, w0),
                make_seeds<true>(ri.x, ri.y, ri.z, c1, cc, h_rev, dq, w1));
        j += pair ? 2 : 1;
    }
    if (cur != CLS_SKIP) flush(cur);

    // epilogue: v_ik = F_i(q_k) X_ik (the all-zero class row serves padding
    // lanes), transposed through LDS so lane k sums column k over the 64 rows in
    // FP64; then 1/q_k and the job scale.  Rows of T are padded to 65 words so
    // both the row writes and the column reads are free of bank conflicts.
    __shared__ float T[K][65];
    const uint32_t ci = cls_i == CLS_SKIP ? (uint32_t)a.zero_class : cls_i;
    const float* __restrict__ fi = ftab + (size_t)ci * ldf + k0;
#pragma unroll
    for (int k = 0; k < K; ++k) T[k][lane] = fi[k] * X[k];
    __syncthreads();
    if (lane < K) {
        double sum = 0.0;
#pragma unroll 16
        for (int t = 0; t < 64; ++t) sum += (double)T[lane][t];
        a.partials[(size_t)jb.slot * ldf + k0 + lane] = sum * a.invq[k0 + lane] * (double)jb.scale;
    }
}

// ---------------------------------------------------------------------------
// uniform-grid path, workgroup variant: the waves of one workgroup are the
// q chunks of one job (up to 4).  The pair geometry that does not depend on q
// (r, 1/r and the FP64 sin(dq r / 2)) is computed once per (row, column) by one
// of the waves and shared through LDS in tiles of SH_TC columns, so each wave
// only forms its chunk's seeds and runs the recurrence.
// ---------------------------------------------------------------------------
constexpr int SH_TC = 32;   // columns per shared tile
constexpr int SH_MAXW = 4;  // waves (q chunks) per workgroup

__device__ __forceinline__ void pair_geometry(float xi, float yi, float zi, const float4 c, double half_dq,
                                              double& rd, double& sh, float& rinv)
{
    const float dx = xi - c.x, dy = yi - c.y, dz = zi - c.z;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float nzf = r2 > 1e-30f ? 1.0f : 0.0f;
    rinv = __builtin_amdgcn_rsqf(fmaxf(r2, 1e-30f)) * nzf;
    rd = (double)(r2 * rinv);
    sh = sin_poly(rd * half_dq);
}

template <bool WEIGHTED>
__device__ __forceinline__ Seeds chunk_seeds(double rd, double sh, float rinv, const ChunkConst& cc, float dq,
                                             float w)
{
    const float sa = sin_rev(cc.a_rev * rd);
    const float cm = cos_rev(cc.mid_rev * rd);
    const float nzf = rinv > 0.0f ? 1.0f : 0.0f;
    const float wr = WEIGHTED ? w * rinv : rinv;
    const float wz = (WEIGHTED ? w : 1.0f) * (1.0f - nzf);
    const float f0 = fmaf(sa, wr, wz * cc.a);
    const double fd = fma(2.0 * (double)(cm * wr), sh, (double)(wz * dq));
    Seeds s;
    s.c2 = fma(-4.0 * sh, sh, 2.0);
    s.u0 = (double)f0;
    s.u1 = s.u0 + fd;
    return s;
}

template <int K>
__global__ __launch_bounds__(64 * SH_MAXW, 2) void debye_shared_kernel(const DebyeArgs a)
{
    union Shared {
        struct {
            double rd[SH_TC][64];
            double sh[SH_TC][64];
            float ri[SH_TC][64];
        } p;
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
        const float* __restrict__ fr = ftab + (size_t)cls * ldf + k0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            X[k] = fmaf(fr[k], (float)acc[k], X[k]);
            acc[k] = 0.0;
        }
    };

    const float4* __restrict__ cols = jb.cols;
    const int ncol = jb.ncol;
    const int jd = jb.diag ? (ncol < 64 ? ncol : 64) : 0;
    uint32_t cur = CLS_SKIP;
    for (int t0 = 0; t0 < ncol; t0 += SH_TC) {
        const int tn = ncol - t0 < SH_TC ? ncol - t0 : SH_TC;
        for (int t = wave; t < tn; t += nw) {
            double rd, sh;
            float rinv;
            pair_geometry(ri.x, ri.y, ri.z, cols[t0 + t], half_dq, rd, sh, rinv);
            sm.p.rd[t][lane] = rd;
            sm.p.sh[t][lane] = sh;
            sm.p.ri[t][lane] = rinv;
        }
        __syncthreads();
        if (active) {
            // one weighted pair step per iteration: the second column is the next
            // one when it is in this tile and of the same class, otherwise a
            // zero-weight copy; skipped records and the lower half of the
            // diagonal tile also carry weight 0
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
                    w0 *= j > lane ? 1.0f : (j == lane ? 0.5f : 0.0f);
                    w1 *= j + 1 > lane ? 1.0f : (j + 1 == lane ? 0.5f : 0.0f);
                }
                run2<K>(acc, chunk_seeds<true>(sm.p.rd[t][lane], sm.p.sh[t][lane], sm.p.ri[t][lane], cc, dq, w0),
                        chunk_seeds<true>(sm.p.rd[t1][lane], sm.p.sh[t1][lane], sm.p.ri[t1][lane], cc, dq, w1));
                t += pair ? 2 : 1;
            }
        }
        __syncthreads();
    }
    if (active && cur != CLS_SKIP) flush(cur);

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
// direct policy (any q grid): FP32 accumulators, one v_sin per term
// ---------------------------------------------------------------------------
template <int K, bool WEIGHTED>
__device__ __forceinline__ void direct_step(float (&acc)[K], float& zsum, float xi, float yi, float zi,
                                            const float4 c, const double* __restrict__ qrev, float w)
{
    const float dx = xi - c.x, dy = yi - c.y, dz = zi - c.z;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const bool nz = r2 > 1e-30f;
    const float rinv = nz ? __builtin_amdgcn_rsqf(r2) : 0.0f;
    const float r = r2 * rinv;
    const float wr = WEIGHTED ? w * rinv : rinv;
    // r == 0 pairs contribute w * q_k; their weights are summed and applied at the flush
    zsum += nz ? 0.0f : (WEIGHTED ? w : 1.0f);
    const double rd = (double)r;
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = fmaf(sin_rev(qrev[k] * rd), wr, acc[k]);
}

template <int K>
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
        const float* __restrict__ fr = ftab + (size_t)cls * ldf + k0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            X[k] = fma((double)fr[k], (double)fmaf(zsum, qv[k], acc[k]), X[k]);
            acc[k] = 0.0f;
        }
        zsum = 0.0f;
    };

    const float4* __restrict__ cols = jb.cols;
    const int ncol = jb.ncol;
    uint32_t cur = CLS_SKIP;
    int j = 0;
    if (jb.diag) {
        const int jd = ncol < 64 ? ncol : 64;
        for (; j < jd; ++j) {
            const float4 c = cols[j];
            const uint32_t cl = __float_as_uint(c.w);
            if (cl != cur) {
                if (cur != CLS_SKIP) flush(cur);
                cur = cl;
            }
            if (cl != CLS_SKIP) {
                const float w = j > lane ? 1.0f : (j == lane ? 0.5f : 0.0f);
                direct_step<K, true>(acc, zsum, ri.x, ri.y, ri.z, c, qrev, w);
            }
        }
    }
    for (; j < ncol; ++j) {
        const float4 c = cols[j];
        const uint32_t cl = __float_as_uint(c.w);
        if (cl != cur) {
            if (cur != CLS_SKIP) flush(cur);
            cur = cl;
        }
        if (cl != CLS_SKIP) direct_step<K, false>(acc, zsum, ri.x, ri.y, ri.z, c, qrev, 1.0f);
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

__global__ void frame_mean_kernel(const double* __restrict__ in, int n_frames, int n_q, double* out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_q) return;
    double s = 0.0;
    for (int f = 0; f < n_frames; ++f) s += in[(size_t)f * n_q + k];
    out[k] = s / (double)n_frames;
}

__global__ void scatter_kernel(float4* cols, const int* __restrict__ slot, const float4* __restrict__ vals, int m)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) cols[slot[t]] = vals[t];
}

__global__ void gather_cols_kernel(float4* cols, const float4* __restrict__ bodies, const int* __restrict__ src,
                                   const uint32_t* __restrict__ counts, int d1, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int c = src[s];
    const float4 b = bodies[c];
    const uint32_t type = __float_as_uint(b.w);
    const uint32_t cls = type * (uint32_t)d1 + (counts ? counts[c] : 0u);
    cols[s] = make_float4(b.x, b.y, b.z, __uint_as_float(cls));
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

__device__ __forceinline__ int cell_hash(int ix, int iy, int iz, int nb)
{
    const uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
    return (int)(h & (uint32_t)(nb - 1));
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

// exclusive scan of bucket_count into bucket_start (one workgroup of 1024)
__global__ __launch_bounds__(1024) void sasa_scan_kernel(SasaArgs a)
{
    __shared__ int warp_tot[16];
    __shared__ int carry;
    const int tid = threadIdx.x;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < a.n_buckets; base += 1024) {
        const int idx = base + tid;
        const int v = idx < a.n_buckets ? a.bucket_count[idx] : 0;
        // inclusive scan within the wave
        int x = v;
        const int lane = tid & 63, w = tid >> 6;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) warp_tot[w] = x;
        __syncthreads();
        int off = 0;
        for (int t = 0; t < w; ++t) off += warp_tot[t];
        const int excl = carry + off + x - v;
        if (idx < a.n_buckets) {
            a.bucket_start[idx] = excl;
            a.bucket_fill[idx] = excl;
        }
        __syncthreads();
        if (tid == 1023) carry = excl + v;
        __syncthreads();
    }
    if (tid == 0) a.bucket_start[a.n_buckets] = carry;
}

__global__ void sasa_place_kernel(SasaArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int pos = atomicAdd(&a.bucket_fill[a.bucket_of[i]], 1);
    a.sorted[pos] = i;
}

constexpr int SASA_MAX_DPL = 8;  // dots per lane (n_dots <= 512)

#pragma clang fp contract(off)
__global__ __launch_bounds__(64) void sasa_dots_kernel(SasaArgs a)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const float4 pi = a.atoms[i];
    const uint32_t ti = __float_as_uint(pi.w);
    const float rext_i = a.rext[ti];
    const int dpl = (a.n_dots + 63) / 64;
    float px[SASA_MAX_DPL], py[SASA_MAX_DPL], pz[SASA_MAX_DPL];
    bool live[SASA_MAX_DPL];
#pragma unroll
    for (int t = 0; t < SASA_MAX_DPL; ++t) {
        const int d = lane + 64 * t;
        live[t] = t < dpl && d < a.n_dots;
        const float4 tp = live[t] ? a.dots[(size_t)ti * a.n_dots + d] : make_float4(0.f, 0.f, 0.f, 0.f);
        px[t] = tp.x + pi.x;
        py[t] = tp.y + pi.y;
        pz[t] = tp.z + pi.z;
    }
    const int cx = cell_coord(pi.x, a.inv_cell), cy = cell_coord(pi.y, a.inv_cell), cz = cell_coord(pi.z, a.inv_cell);
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int b = cell_hash(cx + ox, cy + oy, cz + oz, a.n_buckets);
                const int e = a.bucket_start[b + 1];
                for (int p = a.bucket_start[b]; p < e; ++p) {
                    const int jdx = a.sorted[p];
                    const float4 pj = a.atoms[jdx];
                    const uint32_t tj = __float_as_uint(pj.w);
                    const float ex = pj.x - pi.x, ey = pj.y - pi.y, ez = pj.z - pi.z;
                    const float lim = rext_i + a.rext[tj] + 0.01f;
                    if (ex * ex + ey * ey + ez * ez >= lim * lim) continue;
                    const float r2j = a.r2[tj];
#pragma unroll
                    for (int t = 0; t < SASA_MAX_DPL; ++t) {
                        const float dx = px[t] - pj.x, dy = py[t] - pj.y, dz = pz[t] - pj.z;
                        float dist2 = dx * dx;
                        dist2 = dist2 + dy * dy;
                        dist2 = dist2 + dz * dz;
                        live[t] = live[t] && !((r2j - dist2) > 1e-4f);
                    }
                }
            }
    int cnt = 0;
#pragma unroll
    for (int t = 0; t < SASA_MAX_DPL; ++t) cnt += live[t] ? 1 : 0;
    cnt = wave_sum_i(cnt);
    if (lane == 0) a.counts[i] = (uint32_t)cnt;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// adjoint: grad_i = sum_k w_k dI_k/dx_i
//   dI_k/dx_i = sum_{j != i} 2 F_i F_j [cos(q r) - sinc(q r)] (x_i - x_j) / r^2
// Full row sweep (each lane owns row i and visits every column).  For a run of
// one column class the per-lane weights b_k = 2 w_k F_i,k F_class,k are formed
// once; per (pair, q) the cos and sin/r sequences advance by the same FP64
// recurrence and G += b_k (c_k - u_k / q_k).
// ---------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(64, 2) void adjoint_kernel(const AdjointArgs a)
{
    const int n_rt = (a.n + 63) / 64;
    int bid = blockIdx.x;
    const int chunk = bid % a.n_chunks;
    bid /= a.n_chunks;
    const int cb = bid % a.n_colblocks;
    const int rt = bid / a.n_colblocks;
    if (rt >= n_rt) return;
    const int lane = threadIdx.x;
    const int i = rt * 64 + lane;
    const bool row_ok = i < a.n;
    const float4 ri = row_ok ? a.rows[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t cls_i = row_ok ? __float_as_uint(ri.w) : CLS_SKIP;
    const uint32_t ci = cls_i == CLS_SKIP ? 0u : cls_i;
    const int k0 = chunk * K;
    const ChunkConst cc = a.chunks[chunk];
    const float* __restrict__ ftab = a.ftab;
    const float* __restrict__ fi = ftab + (size_t)ci * a.ldf + k0;
    const double* __restrict__ wq = a.wq + k0;
    const double* __restrict__ invq = a.invq + k0;

    double b[K];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const int j0 = cb * a.col_range;
    const int j1 = min(a.n, j0 + a.col_range);
    uint32_t cur = CLS_SKIP;
    for (int j = j0; j < j1; ++j) {
        const float4 c = a.cols[j];
        const uint32_t cl = __float_as_uint(c.w);
        if (cl == CLS_SKIP) continue;
        if (cl != cur) {
            cur = cl;
            const float* __restrict__ fr = ftab + (size_t)cl * a.ldf + k0;
#pragma unroll
            for (int k = 0; k < K; ++k)
                b[k] = cls_i == CLS_SKIP ? 0.0 : 2.0 * wq[k] * (double)fi[k] * (double)fr[k];
        }
        const float dx = ri.x - c.x, dy = ri.y - c.y, dz = ri.z - c.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const bool nz = r2 > 1e-30f;
        const float rinv = nz ? __builtin_amdgcn_rsqf(r2) : 0.0f;
        const float r = r2 * rinv;
        const double rd = (double)r;
        const double sh = sin_poly(rd * a.half_dq);
        const float sa = sin_rev(cc.a_rev * rd), ca = cos_rev(cc.a_rev * rd);
        const float sm = sin_rev(cc.mid_rev * rd), cm = cos_rev(cc.mid_rev * rd);
        const double c2 = fma(-4.0 * sh, sh, 2.0);
        // u_k = sin(q_k r) / r and v_k = cos(q_k r), both obey x_{k+1} = c2 x_k - x_{k-1}
        double up = (double)(sa * rinv), uc = up + 2.0 * (double)(cm * rinv) * sh;
        double vp = (double)ca, vc = vp - 2.0 * (double)sm * sh;
        double G = 0.0;
        G = fma(b[0], fma(-up, invq[0], vp), G);
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
        const double gs = nz ? G * (double)(rinv * rinv) : 0.0;
        gx = fma(gs, (double)dx, gx);
        gy = fma(gs, (double)dy, gy);
        gz = fma(gs, (double)dz, gz);
    }
    if (row_ok) {
        const size_t part = (size_t)cb * a.n_chunks + chunk;
        double* dst = a.grad_partial + (part * a.n + i) * 4;
        dst[0] = gx;
        dst[1] = gy;
        dst[2] = gz;
        dst[3] = 0.0;
    }
}

__global__ void grad_reduce_kernel(const double* __restrict__ gp, int n_parts, int n, const int* __restrict__ s2c,
                                   float* out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    double x = 0, y = 0, z = 0;
    for (int p = 0; p < n_parts; ++p) {
        const double* v = gp + ((size_t)p * n + s) * 4;
        x += v[0];
        y += v[1];
        z += v[2];
    }
    const int c = s2c[s];
    out[3 * c] = (float)x;
    out[3 * c + 1] = (float)y;
    out[3 * c + 2] = (float)z;
}

template <int... Ks>
struct KList {};

constexpr int UNIFORM_KS[] = {8, 16, 24, 32, 40, 48, 56, 64};
constexpr int DIRECT_KS[] = {8, 16, 32, 48, 64};
constexpr int ADJOINT_KS[] = {8, 16, 32, 48, 64};

}  // namespace

int debye_k_count(bool uniform)
{
    return uniform ? (int)(sizeof(UNIFORM_KS) / sizeof(int)) : (int)(sizeof(DIRECT_KS) / sizeof(int));
}

int debye_k_value(bool uniform, int i)
{
    return uniform ? UNIFORM_KS[i] : DIRECT_KS[i];
}

int adjoint_k_count()
{
    return (int)(sizeof(ADJOINT_KS) / sizeof(int));
}

int adjoint_k_value(int i)
{
    return ADJOINT_KS[i];
}

hipError_t launch_debye(const DebyeArgs& a, int n_jobs, int K, bool uniform, hipStream_t s)
{
    if (n_jobs <= 0) return hipSuccess;
    const dim3 grid((unsigned)(n_jobs * a.n_chunks)), block(64);
    if (uniform && a.shared_geometry && a.n_chunks > 1) {
        const int nw = a.n_chunks < SH_MAXW ? a.n_chunks : SH_MAXW;
        const int ng = (a.n_chunks + nw - 1) / nw;
        const dim3 g2((unsigned)(n_jobs * ng)), b2((unsigned)(64 * nw));
        switch (K) {
        case 8: hipLaunchKernelGGL(debye_shared_kernel<8>, g2, b2, 0, s, a); break;
        case 16: hipLaunchKernelGGL(debye_shared_kernel<16>, g2, b2, 0, s, a); break;
        case 24: hipLaunchKernelGGL(debye_shared_kernel<24>, g2, b2, 0, s, a); break;
        case 32: hipLaunchKernelGGL(debye_shared_kernel<32>, g2, b2, 0, s, a); break;
        case 40: hipLaunchKernelGGL(debye_shared_kernel<40>, g2, b2, 0, s, a); break;
        case 48: hipLaunchKernelGGL(debye_shared_kernel<48>, g2, b2, 0, s, a); break;
        case 56: hipLaunchKernelGGL(debye_shared_kernel<56>, g2, b2, 0, s, a); break;
        case 64: hipLaunchKernelGGL(debye_shared_kernel<64>, g2, b2, 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (uniform) {
        switch (K) {
        case 8: hipLaunchKernelGGL(debye_uniform_kernel<8>, grid, block, 0, s, a); break;
        case 16: hipLaunchKernelGGL(debye_uniform_kernel<16>, grid, block, 0, s, a); break;
        case 24: hipLaunchKernelGGL(debye_uniform_kernel<24>, grid, block, 0, s, a); break;
        case 32: hipLaunchKernelGGL(debye_uniform_kernel<32>, grid, block, 0, s, a); break;
        case 40: hipLaunchKernelGGL(debye_uniform_kernel<40>, grid, block, 0, s, a); break;
        case 48: hipLaunchKernelGGL(debye_uniform_kernel<48>, grid, block, 0, s, a); break;
        case 56: hipLaunchKernelGGL(debye_uniform_kernel<56>, grid, block, 0, s, a); break;
        case 64: hipLaunchKernelGGL(debye_uniform_kernel<64>, grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (K) {
        case 8: hipLaunchKernelGGL(debye_direct_kernel<8>, grid, block, 0, s, a); break;
        case 16: hipLaunchKernelGGL(debye_direct_kernel<16>, grid, block, 0, s, a); break;
        case 32: hipLaunchKernelGGL(debye_direct_kernel<32>, grid, block, 0, s, a); break;
        case 48: hipLaunchKernelGGL(debye_direct_kernel<48>, grid, block, 0, s, a); break;
        case 64: hipLaunchKernelGGL(debye_direct_kernel<64>, grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_reduce(const double* partials, int ldf, int n_q, const int* slot_begin, int n_out, double* out,
                         int out_ld, int accumulate, const double* q0_val, int q0_index, hipStream_t s)
{
    if (n_out <= 0 || n_q <= 0) return hipSuccess;
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

hipError_t launch_scatter(float4* cols, const int* slot, const float4* vals, int m, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, cols, slot, vals, m);
    return hipGetLastError();
}

hipError_t launch_gather_cols(float4* cols, const float4* bodies, const int* src, const uint32_t* counts, int d1,
                              int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_cols_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, cols, bodies, src,
                       counts, d1, n);
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

hipError_t launch_fill_i32(int* p, int value, long long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_sasa(const SasaArgs& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    if (a.n_dots <= 0 || a.n_dots > 64 * SASA_MAX_DPL) return hipErrorInvalidValue;
    hipError_t e = launch_fill_i32(a.bucket_count, 0, (long long)a.n_buckets + 1, s);
    if (e != hipSuccess) return e;
    const dim3 g((unsigned)((a.n + 255) / 256)), b(256);
    hipLaunchKernelGGL(sasa_bin_kernel, g, b, 0, s, a);
    hipLaunchKernelGGL(sasa_scan_kernel, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(sasa_place_kernel, g, b, 0, s, a);
    hipLaunchKernelGGL(sasa_dots_kernel, dim3((unsigned)a.n), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_adjoint(const AdjointArgs& a, int K, hipStream_t s)
{
    const int n_rt = (a.n + 63) / 64;
    const dim3 grid((unsigned)(n_rt * a.n_colblocks * a.n_chunks)), block(64);
    switch (K) {
    case 8: hipLaunchKernelGGL(adjoint_kernel<8>, grid, block, 0, s, a); break;
    case 16: hipLaunchKernelGGL(adjoint_kernel<16>, grid, block, 0, s, a); break;
    case 32: hipLaunchKernelGGL(adjoint_kernel<32>, grid, block, 0, s, a); break;
    case 48: hipLaunchKernelGGL(adjoint_kernel<48>, grid, block, 0, s, a); break;
    case 64: hipLaunchKernelGGL(adjoint_kernel<64>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_grad_reduce(const double* grad_partial, int n_parts, int n, const int* slot_to_caller,
                              float* grad_out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, grad_partial, n_parts,
                       n, slot_to_caller, grad_out);
    return hipGetLastError();
}

}  // namespace rsx
