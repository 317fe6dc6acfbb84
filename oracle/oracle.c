/*
 * oracle.c -- CPU restatement of the reference's Debye-SAXS arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; it is loaded by the test-suite, by smoke() and by the bench's
 * cpu_baseline leg as the checker.  Every entry point is a plain C function
 * over plain pointers so that Python (ctypes) can drive it.
 *
 * Reference behaviour restated here (paths relative to the reference tree):
 *   orc_debye_f64 / orc_debye_rows   resaxslib/src/host_debye.cpp:54-109
 *        (FP64 accumulation of FP32 inputs, lower triangle x2 + diagonal),
 *        with the sinc(0)=1 rule of resaxslib/src/gpu_saxs_pt_wf.cl:44-56
 *        (q == 0 or r == 0 -> 1).  Distances are streamed instead of the
 *        reference's N x N matrix; sin is glibc's (vectorised through
 *        libmvec in the *_f64 variants, scalar in orc_debye_ref_rows).
 *   orc_debye_ref_rows               same loop order as host_debye.cpp:54-109
 *        (distance table for the sampled rows, q outer, scalar libm sin);
 *        this is the timed CPU baseline.
 *   orc_body_factors                 F_i(q) = f_type(q) + s_i * w * f_water(q),
 *        resaxslib/src/gpu_saxs_pt_wf.cl:130-138 (evaluated in FP64).
 *   orc_fibonacci_dots               resaxslib/src/rsasa_algorithm.cpp:96-116
 *   orc_dodeca_icosa_dots            resaxslib/src/rsasa_algorithm.cpp:157-208
 *   orc_sasa_counts[_cells]          occlusion predicate of
 *        resaxslib/src/rsasa_dot_spheres.cl:108-135 (FP32, no contraction,
 *        (p+r_j)^2 - |dot - x_j|^2 > 1e-4f over every atom j incl. i).
 *   orc_atomic_ff / orc_atomic_radii resaxslib/include/ff_coef.hpp:78-211 with
 *        the constants of resaxslib/include/constants.hpp:28-81 (FP32, as the
 *        reference's float instantiation).
 *   orc_scale_chi2                   resaxslib/include/saxs_profile.hpp:137-166
 *        (least-squares scale; chi^2 = sum (I_exp - c I)^2 / sigma^2).
 *   orc_chi2_grad_f64 / orc_chi2_fd  no reference counterpart: analytic FP64
 *        gradient of chi^2 at the optimal scale and its central finite
 *        differences (h in Angstrom), the checks for the adjoint kernel.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* glibc's libmvec vector sin/cos (4 x double, AVX2) for the FP64 sums. */
extern double sin(double) __attribute__((__simd__("notinbranch")));
extern double cos(double) __attribute__((__simd__("notinbranch")));

#define ORC_PI 3.14159265358979323846

static int clamp_threads(int t)
{
    if (t <= 0) t = omp_get_max_threads();
    if (t < 1) t = 1;
    return t;
}

/* ------------------------------------------------------------------------ */
/* Debye sums                                                               */
/* ------------------------------------------------------------------------ */

/* sinc with the reference GPU rule: q == 0 or r == 0 gives 1. */
static inline double sinc_rule(double arg)
{
    return arg != 0.0 ? sin(arg) / arg : 1.0;
}

/*
 * Full I(q) = sum_i sum_j F_i F_j sinc(q r_ij), FP64.
 *   q[nq], xyz[n*3] (FP64 copies of FP32 inputs), F[n*nq] (atom-major).
 * Lower triangle doubled plus the diagonal, as host_debye.cpp:83-106.
 */
void orc_debye_f64(const double *q, int nq, const double *xyz, int n,
                   const double *F, double *out, int nthreads)
{
    nthreads = clamp_threads(nthreads);
    double *acc = (double *)calloc((size_t)nthreads * nq, sizeof(double));
    double *auto_acc = (double *)calloc((size_t)nthreads * nq, sizeof(double));

#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
        double *Iq = acc + (size_t)t * nq;
        double *Au = auto_acc + (size_t)t * nq;
        double *inner = (double *)malloc(sizeof(double) * nq);
#pragma omp for schedule(dynamic, 16)
        for (int y = 0; y < n; ++y) {
            const double *Fy = F + (size_t)y * nq;
            for (int k = 0; k < nq; ++k) {
                Au[k] += Fy[k] * Fy[k];
                inner[k] = 0.0;
            }
            const double xy = xyz[3 * y], yy = xyz[3 * y + 1], zy = xyz[3 * y + 2];
            for (int x = 0; x < y; ++x) {
                const double dx = xyz[3 * x] - xy, dy = xyz[3 * x + 1] - yy,
                             dz = xyz[3 * x + 2] - zy;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                const double *Fx = F + (size_t)x * nq;
#pragma omp simd
                for (int k = 0; k < nq; ++k) {
                    const double arg = q[k] * r;
                    const double sa = sin(arg);
                    const double s = arg != 0.0 ? sa / arg : 1.0;
                    inner[k] += Fx[k] * s;
                }
            }
            for (int k = 0; k < nq; ++k) Iq[k] += inner[k] * Fy[k];
        }
        free(inner);
    }
    for (int k = 0; k < nq; ++k) {
        double I = 0.0, A = 0.0;
        for (int t = 0; t < nthreads; ++t) {
            I += acc[(size_t)t * nq + k];
            A += auto_acc[(size_t)t * nq + k];
        }
        out[k] = 2.0 * I + A;
    }
    free(acc);
    free(auto_acc);
}

/*
 * Row partial sums: out[k] = sum_{i in rows} sum_{j=0..n-1} F_i F_j sinc(q r_ij)
 * (every j, including j == i).  This is the quantity a row sample of the GPU
 * pair sum reports, and it also yields exact incremental differences.
 */
void orc_debye_rows(const double *q, int nq, const double *xyz, int n,
                    const double *F, const int *rows, int nrows, double *out,
                    int nthreads)
{
    nthreads = clamp_threads(nthreads);
    double *acc = (double *)calloc((size_t)nthreads * nq, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
        double *Iq = acc + (size_t)t * nq;
        double *inner = (double *)malloc(sizeof(double) * nq);
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < nrows; ++a) {
            const int i = rows[a];
            const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
            for (int k = 0; k < nq; ++k) inner[k] = 0.0;
            for (int j = 0; j < n; ++j) {
                const double dx = xyz[3 * j] - xi, dy = xyz[3 * j + 1] - yi,
                             dz = xyz[3 * j + 2] - zi;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                const double *Fj = F + (size_t)j * nq;
#pragma omp simd
                for (int k = 0; k < nq; ++k) {
                    const double arg = q[k] * r;
                    const double sa = sin(arg);
                    const double s = arg != 0.0 ? sa / arg : 1.0;
                    inner[k] += Fj[k] * s;
                }
            }
            const double *Fi = F + (size_t)i * nq;
            for (int k = 0; k < nq; ++k) Iq[k] += Fi[k] * inner[k];
        }
        free(inner);
    }
    for (int k = 0; k < nq; ++k) {
        double s = 0.0;
        for (int t = 0; t < nthreads; ++t) s += acc[(size_t)t * nq + k];
        out[k] = s;
    }
    free(acc);
}

/*
 * The reference CPU path's own loop order (host_debye.cpp:54-109) over the
 * row range y in [y0, y1): distances for those rows are tabulated first, then
 * q is the outer loop and libm sin is called once per (pair, q).  Inputs are
 * the reference's: FP32 q, FP32 bodies (x,y,z,type bits), FP32 factor table
 * [nq][n_types].  Output: sum over the rows of 2 F_y sum_{x<y} F_x sinc + F_y^2
 * (the full I(q) when y0 = 0 and y1 = n).  With nthreads > 1 the q loop is
 * split across threads.
 */
void orc_debye_ref_rows(const float *q, int nq, const float *bodies, int n,
                        const float *table, int n_types, int y0, int y1,
                        double *out, int nthreads)
{
    nthreads = clamp_threads(nthreads);
    if (y0 < 0) y0 = 0;
    if (y1 > n) y1 = n;
    const int nr = y1 > y0 ? y1 - y0 : 0;
    /* distance rows y0..y1-1, each of length y (lower triangle) */
    size_t *off = (size_t *)malloc(sizeof(size_t) * (nr + 1));
    off[0] = 0;
    for (int a = 0; a < nr; ++a) off[a + 1] = off[a] + (size_t)(y0 + a);
    double *r = (double *)malloc(sizeof(double) * (off[nr] ? off[nr] : 1));
    uint32_t *type = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (int i = 0; i < n; ++i) memcpy(&type[i], &bodies[4 * i + 3], 4);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 8)
    for (int a = 0; a < nr; ++a) {
        const int y = y0 + a;
        for (int x = 0; x < y; ++x) {
            const double dx = (double)bodies[4 * y] - bodies[4 * x];
            const double dy = (double)bodies[4 * y + 1] - bodies[4 * x + 1];
            const double dz = (double)bodies[4 * y + 2] - bodies[4 * x + 2];
            r[off[a] + x] = sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int k = 0; k < nq; ++k) {
        const double qk = q[k];
        const float *Fq = table + (size_t)k * n_types;
        double Iq = 0.0, auto_Iq = 0.0;
        for (int a = 0; a < nr; ++a) {
            const int y = y0 + a;
            const double Fy = Fq[type[y]];
            auto_Iq += Fy * Fy;
            double inner = 0.0;
            const double *ry = r + off[a];
            for (int x = 0; x < y; ++x) {
                double v = 1.0;
                if (qk != 0.0) v = sinc_rule(qk * ry[x]);
                inner += Fq[type[x]] * v;
            }
            Iq += inner * Fy;
        }
        out[k] = 2.0 * Iq + auto_Iq;
    }
    free(off);
    free(r);
    free(type);
}

/*
 * Per-atom factors F[i*nq + k] = table[k][type_i] + s_i * (w * table[k][water])
 * in FP64 (gpu_saxs_pt_wf.cl:130-138).  water < 0 or w == 0 gives the plain
 * factors.  type ids are read from the bit pattern of bodies[4i+3].
 */
void orc_body_factors(const float *bodies, int n, const float *table, int nq,
                      int n_types, int water, float w, const float *s,
                      double *F)
{
    for (int i = 0; i < n; ++i) {
        uint32_t t;
        memcpy(&t, &bodies[4 * i + 3], 4);
        for (int k = 0; k < nq; ++k) {
            double f = table[(size_t)k * n_types + t];
            if (water >= 0 && w != 0.0f && s)
                f += (double)s[i] * ((double)w * (double)table[(size_t)k * n_types + water]);
            F[(size_t)i * nq + k] = f;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* SASA                                                                     */
/* ------------------------------------------------------------------------ */

/*
 * Fibonacci dot sphere, rsasa_algorithm.cpp:96-116, in float arithmetic as
 * the reference's float instantiation: for i in [-(n-1), n]:
 *   theta = 2*pi*i/phi, sphi = i/n, cphi = sqrt((n+i)(n-i))/n (double sqrt),
 *   (x, y, z) = (cphi sin(theta), cphi cos(theta), sphi) * (radius + probe).
 * out: 2n float4 {x, y, z, radius}.
 */
void orc_fibonacci_dots(int nhalf, float radius, float probe, float *out)
{
    const float phi = (1.0f + sqrtf(5.0f)) / 2.0f;
    const float ext = radius + probe;
    int d = 0;
    for (int i = -(nhalf - 1); i <= nhalf; ++i, ++d) {
        const float theta = 2.0f * (float)ORC_PI * (float)i / phi;
        const float sphi = (float)i / (float)nhalf;
        const float cphi = (float)(sqrt((double)((nhalf + i) * (nhalf - i))) / nhalf);
        out[4 * d + 0] = cphi * sinf(theta) * ext;
        out[4 * d + 1] = cphi * cosf(theta) * ext;
        out[4 * d + 2] = sphi * ext;
        out[4 * d + 3] = radius;
    }
}

/*
 * 20 dodecahedron + 12 icosahedron vertices on a sphere of radius
 * radius+probe, rsasa_algorithm.cpp:157-208 (float instantiation; the
 * icosahedron scale is evaluated in double as there).  out: 32 float4.
 */
void orc_dodeca_icosa_dots(float radius, float probe, float *out)
{
    const float ext = radius + probe;
    const float phi = (1.0f + sqrtf(5.0f)) / 2.0f;
    const float r_phi = 1.0f / phi;
    const float r_sqrt3 = 1.0f / sqrtf(3.0f);
    const float v1 = r_sqrt3 * ext, v2 = r_phi * v1, v3 = phi * v1;
    const double cc = (double)phi * cos(3.0 * ORC_PI / 10.0);
    const float ico = (float)((double)((sqrtf(5.0f) + 5.0f) / 10.0f) /
                              sqrt(3.0 - 1.0 / pow(cc, 2.0)) * (double)ext);
    const float a = ico, b = phi * ico;
    const float pts[32][3] = {
        {v1, v1, v1}, {v1, v1, -v1}, {v1, -v1, v1}, {v1, -v1, -v1},
        {-v1, v1, v1}, {-v1, v1, -v1}, {-v1, -v1, v1}, {-v1, -v1, -v1},
        {0, v2, v3}, {0, v2, -v3}, {0, -v2, v3}, {0, -v2, -v3},
        {v2, v3, 0}, {v2, -v3, 0}, {-v2, v3, 0}, {-v2, -v3, 0},
        {v3, 0, v2}, {v3, 0, -v2}, {-v3, 0, v2}, {-v3, 0, -v2},
        {a, 0, b}, {a, 0, -b}, {-a, 0, b}, {-a, 0, -b},
        {0, b, a}, {0, b, -a}, {0, -b, a}, {0, -b, -a},
        {b, a, 0}, {b, -a, 0}, {-b, a, 0}, {-b, -a, 0}};
    for (int d = 0; d < 32; ++d) {
        out[4 * d + 0] = pts[d][0];
        out[4 * d + 1] = pts[d][1];
        out[4 * d + 2] = pts[d][2];
        out[4 * d + 3] = radius;
    }
}

/* The occlusion predicate, FP32, fixed operation order, no contraction. */
static inline int dot_occluded(float px, float py, float pz, float xj, float yj,
                               float zj, float R2j)
{
    const float dx = px - xj, dy = py - yj, dz = pz - zj;
    float dist2 = dx * dx;
    dist2 = dist2 + dy * dy;
    dist2 = dist2 + dz * dz;
    return (R2j - dist2) > 1e-4f;
}

/*
 * Exposed-dot count per atom by brute force over every atom j (including i),
 * rsasa_dot_spheres.cl:114-135.
 *   bodies[n*4] (x,y,z,type bits), R2[type] = (probe + r_type)^2 (FP32),
 *   dots[n_types][D][4] templates (relative to the atom centre).
 */
void orc_sasa_counts(const float *bodies, int n, const float *R2,
                     const float *dots, int D, uint32_t *counts, int nthreads)
{
    nthreads = clamp_threads(nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        uint32_t ti;
        memcpy(&ti, &bodies[4 * i + 3], 4);
        uint32_t cnt = 0;
        for (int d = 0; d < D; ++d) {
            const float *t = dots + ((size_t)ti * D + d) * 4;
            const float px = t[0] + bodies[4 * i], py = t[1] + bodies[4 * i + 1],
                        pz = t[2] + bodies[4 * i + 2];
            int occ = 0;
            for (int j = 0; j < n && !occ; ++j) {
                uint32_t tj;
                memcpy(&tj, &bodies[4 * j + 3], 4);
                occ = dot_occluded(px, py, pz, bodies[4 * j], bodies[4 * j + 1],
                                   bodies[4 * j + 2], R2[tj]);
            }
            cnt += occ ? 0u : 1u;
        }
        counts[i] = cnt;
    }
}

/*
 * Same counts with a uniform cell grid.  A candidate j is skipped only when
 * |x_i - x_j| >= Rext_i + Rext_j + 0.01, where Rext = sqrt(R2): every dot of
 * i then lies farther than sqrt(R2_j) from x_j and the predicate is false, so
 * the result equals orc_sasa_counts bit for bit.
 */
void orc_sasa_counts_cells(const float *bodies, int n, const float *R2,
                           int n_types, const float *dots, int D,
                           uint32_t *counts, int nthreads)
{
    nthreads = clamp_threads(nthreads);
    if (n <= 0) return;
    double rmax = 0.0;
    for (int t = 0; t < n_types; ++t)
        if (sqrt((double)R2[t]) > rmax) rmax = sqrt((double)R2[t]);
    const double cs = 2.0 * rmax + 0.02;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            double v = bodies[4 * i + c];
            if (v < lo[c]) lo[c] = v;
            if (v > hi[c]) hi[c] = v;
        }
    int dim[3];
    for (int c = 0; c < 3; ++c) dim[c] = (int)floor((hi[c] - lo[c]) / cs) + 1;
    const size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
    int *cell_of = (int *)malloc(sizeof(int) * n);
    int *start = (int *)calloc(ncell + 1, sizeof(int));
    int *order = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        int ix = (int)floor((bodies[4 * i] - lo[0]) / cs);
        int iy = (int)floor((bodies[4 * i + 1] - lo[1]) / cs);
        int iz = (int)floor((bodies[4 * i + 2] - lo[2]) / cs);
        cell_of[i] = (iz * dim[1] + iy) * dim[0] + ix;
        start[cell_of[i] + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    int *fill = (int *)malloc(sizeof(int) * (ncell ? ncell : 1));
    memcpy(fill, start, sizeof(int) * ncell);
    for (int i = 0; i < n; ++i) order[fill[cell_of[i]]++] = i;
    free(fill);

#pragma omp parallel num_threads(nthreads)
    {
        int *cand = (int *)malloc(sizeof(int) * n);
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; ++i) {
            uint32_t ti;
            memcpy(&ti, &bodies[4 * i + 3], 4);
            const double Ri = sqrt((double)R2[ti]);
            const int c0 = cell_of[i];
            const int ix = c0 % dim[0], iy = (c0 / dim[0]) % dim[1], iz = c0 / (dim[0] * dim[1]);
            int nc = 0;
            for (int z = iz - 1; z <= iz + 1; ++z)
                for (int y = iy - 1; y <= iy + 1; ++y)
                    for (int x = ix - 1; x <= ix + 1; ++x) {
                        if (x < 0 || y < 0 || z < 0 || x >= dim[0] || y >= dim[1] || z >= dim[2])
                            continue;
                        const int c = (z * dim[1] + y) * dim[0] + x;
                        for (int p = start[c]; p < start[c + 1]; ++p) {
                            const int j = order[p];
                            uint32_t tj;
                            memcpy(&tj, &bodies[4 * j + 3], 4);
                            const double dx = (double)bodies[4 * j] - bodies[4 * i];
                            const double dy = (double)bodies[4 * j + 1] - bodies[4 * i + 1];
                            const double dz = (double)bodies[4 * j + 2] - bodies[4 * i + 2];
                            const double lim = Ri + sqrt((double)R2[tj]) + 0.01;
                            if (dx * dx + dy * dy + dz * dz < lim * lim) cand[nc++] = j;
                        }
                    }
            uint32_t cnt = 0;
            for (int d = 0; d < D; ++d) {
                const float *t = dots + ((size_t)ti * D + d) * 4;
                const float px = t[0] + bodies[4 * i], py = t[1] + bodies[4 * i + 1],
                            pz = t[2] + bodies[4 * i + 2];
                int occ = 0;
                for (int a = 0; a < nc && !occ; ++a) {
                    const int j = cand[a];
                    uint32_t tj;
                    memcpy(&tj, &bodies[4 * j + 3], 4);
                    occ = dot_occluded(px, py, pz, bodies[4 * j], bodies[4 * j + 1],
                                       bodies[4 * j + 2], R2[tj]);
                }
                cnt += occ ? 0u : 1u;
            }
            counts[i] = cnt;
        }
        free(cand);
    }
    free(cell_of);
    free(start);
    free(order);
}

/* ------------------------------------------------------------------------ */
/* Default atomic form factors (ff_coef.hpp:78-211, constants.hpp:28-81)    */
/* ------------------------------------------------------------------------ */

/* Cromer-Mann a1..a4, c, b1..b4 for H, C, N, O, P, S, Zn (ff_coef.hpp:39-48). */
static const float CM[7][9] = {
    {0.4930020f, 0.3229120f, 0.1401910f, 4.0810000E-02f, 3.0380001E-03f, 10.51090f, 26.12570f, 3.142360f, 57.79970f},
    {2.310000f, 1.020000f, 1.588600f, 0.8650000f, 0.2156000f, 20.84390f, 10.20750f, 0.5687000f, 51.65120f},
    {12.21260f, 3.132200f, 2.012500f, 1.166300f, -11.52900f, 5.7000001E-03f, 9.893300f, 28.99750f, 0.5826000f},
    {3.048500f, 2.286800f, 1.546300f, 0.8670000f, 0.2508000f, 13.27710f, 5.701100f, 0.3239000f, 32.90890f},
    {6.434500f, 4.179100f, 1.780000f, 1.490800f, 1.114900f, 1.906700f, 27.15700f, 0.5260000f, 68.16450f},
    {6.905300f, 5.203400f, 1.437900f, 1.586300f, 0.8669000f, 1.467900f, 22.21510f, 0.2536000f, 56.17200f},
    {14.07430f, 7.031800f, 5.165200f, 2.410000f, 1.304100f, 3.265500f, 0.2333000f, 10.31630f, 58.70970f}};
/* excluded volumes (constants.hpp:61-69) and complex types (:71-81) */
static const float EXV[7] = {5.15f, 16.44f, 2.49f, 9.13f, 5.73f, 19.86f, 9.85f};
static const int CPLX[9][2] = {{1, 1}, {1, 2}, {1, 3}, {2, 1}, {2, 2}, {2, 3}, {3, 1}, {3, 2}, {5, 1}};
#define ORC_RHO0 0.334f

/* The reference evaluates exp and pow through the C library's double
 * functions (float arguments promoted), rounding once to float at the end of
 * each expression; restated in that order so the table agrees bit for bit. */
static float excl_ff(int t, float k_sq)
{
    const double e = exp((double)(-(float)ORC_PI) * pow((double)EXV[t], (double)(2.0f / 3.0f)) * (double)k_sq);
    return (float)((double)(ORC_RHO0 * EXV[t]) * e);
}

/* radii of the 16 atom types from the excluded-volume factor at q = 0. */
void orc_atomic_radii(float *out16)
{
    float fe[16];
    for (int t = 0; t < 7; ++t) fe[t] = excl_ff(t, 0.0f);
    for (int t = 7; t < 16; ++t) fe[t] = fe[CPLX[t - 7][0]] + (float)CPLX[t - 7][1] * fe[0];
    const float coef = (float)(3.0 / (4.0 * ORC_PI * 0.334f));
    for (int t = 0; t < 16; ++t) out16[t] = powf(coef * fe[t], (float)(1.0 / 3));
}

/*
 * Table [nq][16] of f_vac - f_excl * E(k^2) for H..Zn then CH..SH
 * (base + n_H * f_H).  E = 1 when c == 1, otherwise
 * c^3 exp(-pi (4pi/3)^(2/3) k^2 rm^2 (c^2 - 1)).
 */
void orc_atomic_ff(const float *q, int nq, float avg_radius, float c, float *out)
{
    const float k43 = (float)pow((double)(4.0f * (float)ORC_PI / 3.0f), (double)(2.0f / 3.0f));
    for (int i = 0; i < nq; ++i) {
        float *page = out + (size_t)i * 16;
        const float k = q[i] / (4.0f * (float)ORC_PI);
        const float k_sq = (float)(k * k);
        float E = 1.0f;
        if (c != 1.0f) {
            const float c_sq = c * c;
            const float arg = -(float)ORC_PI * k43 * k_sq * avg_radius * avg_radius * (c_sq - 1.0f);
            E = (float)((double)(c_sq * c) * exp((double)arg));
        }
        for (int t = 0; t < 7; ++t) {
            float f = CM[t][4];
            for (int a = 0; a < 4; ++a) f = (float)((double)f + (double)CM[t][a] * exp((double)(-CM[t][5 + a] * k_sq)));
            page[t] = f - excl_ff(t, k_sq) * E;
        }
        for (int t = 7; t < 16; ++t)
            page[t] = page[CPLX[t - 7][0]] + (float)CPLX[t - 7][1] * page[0];
    }
}

/* ------------------------------------------------------------------------ */
/* chi^2 and its gradient                                                   */
/* ------------------------------------------------------------------------ */

/* c* = sum(Iexp I / s^2) / sum(I^2 / s^2); chi2 = sum((Iexp - c* I)^2 / s^2). */
void orc_scale_chi2(const double *Iexp, const double *sigma, const double *I,
                    int nq, double *scale, double *chi2)
{
    double num = 0.0, den = 0.0;
    for (int k = 0; k < nq; ++k) {
        const double p2 = 1.0 / (sigma[k] * sigma[k]);
        num += Iexp[k] * I[k] * p2;
        den += I[k] * I[k] * p2;
    }
    const double c = num / den;
    double x2 = 0.0;
    for (int k = 0; k < nq; ++k) {
        const double d = Iexp[k] - c * I[k];
        x2 += d * d / (sigma[k] * sigma[k]);
    }
    *scale = c;
    *chi2 = x2;
}

/*
 * Analytic dchi2/dx_i at the optimal scale (held fixed: d chi2/d c = 0):
 *   w_k = -2 c (Iexp_k - c I_k) / s_k^2,
 *   g_i = sum_j sum_k w_k 2 F_ik F_jk [cos(q r) - sinc(q r)] (x_i - x_j) / r^2.
 * Also returns I, c and chi2.  grad[n*3].
 */
void orc_chi2_grad_f64(const double *q, int nq, const double *xyz, int n,
                       const double *F, const double *Iexp, const double *sigma,
                       double *grad, double *I_out, double *scale, double *chi2,
                       int nthreads)
{
    nthreads = clamp_threads(nthreads);
    double *I = (double *)malloc(sizeof(double) * nq);
    orc_debye_f64(q, nq, xyz, n, F, I, nthreads);
    double c, x2;
    orc_scale_chi2(Iexp, sigma, I, nq, &c, &x2);
    double *w = (double *)malloc(sizeof(double) * nq);
    for (int k = 0; k < nq; ++k) w[k] = -2.0 * c * (Iexp[k] - c * I[k]) / (sigma[k] * sigma[k]);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        const double *Fi = F + (size_t)i * nq;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                         dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            const double r = sqrt(r2);
            const double *Fj = F + (size_t)j * nq;
            double G = 0.0;
#pragma omp simd reduction(+ : G)
            for (int k = 0; k < nq; ++k) {
                const double u = q[k] * r;
                const double su = sin(u), cu = cos(u);
                const double t = u != 0.0 ? cu - su / u : 0.0;
                G += w[k] * 2.0 * Fi[k] * Fj[k] * t;
            }
            G /= r2;
            gx += G * dx;
            gy += G * dy;
            gz += G * dz;
        }
        grad[3 * i] = gx;
        grad[3 * i + 1] = gy;
        grad[3 * i + 2] = gz;
    }
    if (I_out) memcpy(I_out, I, sizeof(double) * nq);
    *scale = c;
    *chi2 = x2;
    free(I);
    free(w);
}

/*
 * Central finite differences of chi2 (optimal scale re-solved at every
 * point) for coordinates comp[a] (= 3*atom + axis), a < m, step h:
 *   fd[a] = (chi2(x + h e) - chi2(x - h e)) / (2h).
 * Moving one atom changes only its row of the pair sum, so each displaced
 * I(q) is formed exactly as I + 2 F_i sum_{j != i} F_j (sinc' - sinc).
 */
void orc_chi2_fd(const double *q, int nq, const double *xyz, int n,
                 const double *F, const double *Iexp, const double *sigma,
                 const int *comp, int m, double h, double *fd, int nthreads)
{
    nthreads = clamp_threads(nthreads);
    double *I = (double *)malloc(sizeof(double) * nq);
    orc_debye_f64(q, nq, xyz, n, F, I, nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        double *Ip = (double *)malloc(sizeof(double) * nq);
        double *Im = (double *)malloc(sizeof(double) * nq);
        double *dp = (double *)malloc(sizeof(double) * nq);
        double *dm = (double *)malloc(sizeof(double) * nq);
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < m; ++a) {
            const int i = comp[a] / 3, ax = comp[a] % 3;
            const double *Fi = F + (size_t)i * nq;
            for (int k = 0; k < nq; ++k) dp[k] = dm[k] = 0.0;
            double pp[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
            double pm[3] = {pp[0], pp[1], pp[2]};
            pp[ax] += h;
            pm[ax] -= h;
            for (int j = 0; j < n; ++j) {
                if (j == i) continue;
                const double *xj = xyz + 3 * j;
                const double r0 = sqrt((xj[0] - xyz[3 * i]) * (xj[0] - xyz[3 * i]) +
                                       (xj[1] - xyz[3 * i + 1]) * (xj[1] - xyz[3 * i + 1]) +
                                       (xj[2] - xyz[3 * i + 2]) * (xj[2] - xyz[3 * i + 2]));
                const double rp = sqrt((xj[0] - pp[0]) * (xj[0] - pp[0]) + (xj[1] - pp[1]) * (xj[1] - pp[1]) +
                                       (xj[2] - pp[2]) * (xj[2] - pp[2]));
                const double rm = sqrt((xj[0] - pm[0]) * (xj[0] - pm[0]) + (xj[1] - pm[1]) * (xj[1] - pm[1]) +
                                       (xj[2] - pm[2]) * (xj[2] - pm[2]));
                const double *Fj = F + (size_t)j * nq;
#pragma omp simd
                for (int k = 0; k < nq; ++k) {
                    const double a0 = q[k] * r0, ap = q[k] * rp, am = q[k] * rm;
                    const double s0 = sin(a0), sp = sin(ap), sm = sin(am);
                    const double c0 = a0 != 0.0 ? s0 / a0 : 1.0;
                    dp[k] += Fj[k] * ((ap != 0.0 ? sp / ap : 1.0) - c0);
                    dm[k] += Fj[k] * ((am != 0.0 ? sm / am : 1.0) - c0);
                }
            }
            for (int k = 0; k < nq; ++k) {
                Ip[k] = I[k] + 2.0 * Fi[k] * dp[k];
                Im[k] = I[k] + 2.0 * Fi[k] * dm[k];
            }
            double cp, xp, cm, xm;
            orc_scale_chi2(Iexp, sigma, Ip, nq, &cp, &xp);
            orc_scale_chi2(Iexp, sigma, Im, nq, &cm, &xm);
            fd[a] = (xp - xm) / (2.0 * h);
        }
        free(Ip);
        free(Im);
        free(dp);
        free(dm);
    }
    free(I);
}
