// Exposed-dot counts that allocate nothing: rsx_calc_rsasa_noalloc.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, serving loops,
// harnesses that forbid allocation).  The SASA kernels' working memory (hashed cell list:
// bucket ids, counts, starts and fill pointers, scan tile totals, atoms sorted by bucket;
// the per-type radius tables) lives in one file-scope device array, `scratch`, of
// SCRATCH_BYTES per GPU.  Host code never uses the array's name as an address: the
// first call on a device takes its address with hipGetSymbolAddress, with that device
// current, and keeps it per device; kernels receive that pointer.
//
// Every call on one GPU shares that GPU's array, which keeps the last call's contents
// (the bucket counts are cleared by a fill kernel at the start of every call).  Calls that
// overlap on one GPU -- two streams, or two host threads -- race on it.

#include "kernels.h"
#include "../../include/rsx_capi.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

namespace {

constexpr int MAX_ATOMS = RSX_SASA_NOALLOC_MAX_ATOMS;
constexpr int MAX_TYPES = 64;
constexpr int MAX_BUCKETS = 2 * MAX_ATOMS;  // a power of two >= 2 n (n <= MAX_ATOMS)
constexpr int MAX_DEVICES = 64;

constexpr size_t align256(size_t b)
{
    return (b + 255) & ~(size_t)255;
}

// scratch layout for n atoms and nb buckets (each part 256-byte aligned)
struct Layout {
    size_t bucket_of, bucket_count, bucket_start, bucket_fill, tile_sum, sorted, sorted_atoms, r2, rext, end;
    Layout(int n, int nb)
    {
        size_t o = 0;
        sorted_atoms = o;
        o += align256((size_t)n * sizeof(float4));
        bucket_of = o;
        o += align256((size_t)n * sizeof(int));
        sorted = o;
        o += align256((size_t)n * sizeof(int));
        bucket_count = o;
        o += align256(((size_t)nb + 1) * sizeof(int));
        bucket_start = o;
        o += align256(((size_t)nb + 1) * sizeof(int));
        bucket_fill = o;
        o += align256((size_t)nb * sizeof(int));
        tile_sum = o;
        o += align256((size_t)(nb / 1024) * sizeof(int));
        r2 = o;
        o += align256(MAX_TYPES * sizeof(float));
        rext = o;
        o += align256(MAX_TYPES * sizeof(float));
        end = o;
    }
};

constexpr size_t scratch_bytes()
{
    // the Layout above at n = MAX_ATOMS, nb = MAX_BUCKETS, computed without the constructor
    return align256((size_t)MAX_ATOMS * 16) + 2 * align256((size_t)MAX_ATOMS * 4) +
           2 * align256(((size_t)MAX_BUCKETS + 1) * 4) + align256((size_t)MAX_BUCKETS * 4) +
           align256((size_t)(MAX_BUCKETS / 1024) * 4) + 2 * align256(MAX_TYPES * 4);
}
constexpr size_t SCRATCH_BYTES = scratch_bytes();
static_assert(SCRATCH_BYTES == RSX_SASA_NOALLOC_SCRATCH_BYTES, "rsx_capi.h states the scratch size");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// per-type tables from the radii passed by value: r2 = (probe + r)^2 in FP32, rext =
// sqrt(r2) (rounded once from FP64), as the engine's host set-up forms them
struct Radii {
    float r[MAX_TYPES];
    int n_types;
    float probe;
};

__global__ void sasa_tables_kernel(Radii in, float* r2, float* rext)
{
    const int t = threadIdx.x;
    if (t < in.n_types) {
        const float R = in.probe + in.r[t];
        const float s = R * R;
        r2[t] = s;
        rext[t] = (float)sqrt((double)s);
    }
}

std::mutex g_mu;
unsigned char* g_scratch[MAX_DEVICES] = {};

// this device's scratch address, taken once per device with the device current
int scratch_of(int device, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RSX_HIP_ERROR;
        g_scratch[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[device];
    return RSX_OK;
}

}  // namespace

extern "C" int rsx_calc_rsasa_noalloc(const void* bodies_dev, int n, const float* radii, int n_types, float probe,
                                      const void* dots_dev, int n_dots, unsigned* counts_dev, void* stream)
{
    if (n < 0 || n > MAX_ATOMS) return RSX_INVALID_ARG;
    if (n_types < 1 || n_types > MAX_TYPES || !radii || !(probe >= 0.0f) || !std::isfinite(probe))
        return RSX_INVALID_ARG;
    if (n_dots < 1 || n_dots > 512) return RSX_INVALID_ARG;
    if (n == 0) return RSX_OK;
    if (!bodies_dev || !dots_dev || !counts_dev) return RSX_INVALID_ARG;
    Radii rt;
    for (int t = 0; t < MAX_TYPES; ++t) rt.r[t] = t < n_types ? radii[t] : 0.0f;
    float rmax = 0.0f;
    for (int t = 0; t < n_types; ++t) {
        if (!(radii[t] >= 0.0f) || !std::isfinite(radii[t])) return RSX_INVALID_ARG;
        const float R = probe + radii[t];
        rmax = std::fmax(rmax, (float)std::sqrt((double)(R * R)));
    }
    rt.n_types = n_types;
    rt.probe = probe;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= MAX_DEVICES) return RSX_INVALID_DEVICE;
    unsigned char* base = nullptr;
    const int rc = scratch_of(device, &base);
    if (rc != RSX_OK) return rc;
    int nb = 1024;
    while (nb < 2 * n) nb <<= 1;
    const Layout L(n, nb);
    if (L.end > SCRATCH_BYTES) return RSX_INVALID_ARG;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    rsx::SasaArgs a;
    a.atoms = static_cast<const float4*>(bodies_dev);
    a.n = n;
    a.n_types = n_types;
    a.r2 = reinterpret_cast<const float*>(base + L.r2);
    a.rext = reinterpret_cast<const float*>(base + L.rext);
    a.dots = static_cast<const float4*>(dots_dev);
    a.n_dots = n_dots;
    a.cell = 2.0f * rmax + 0.02f;
    a.inv_cell = 1.0f / a.cell;
    a.n_buckets = nb;
    a.bucket_of = reinterpret_cast<int*>(base + L.bucket_of);
    a.bucket_count = reinterpret_cast<int*>(base + L.bucket_count);
    a.bucket_start = reinterpret_cast<int*>(base + L.bucket_start);
    a.bucket_fill = reinterpret_cast<int*>(base + L.bucket_fill);
    a.tile_sum = reinterpret_cast<int*>(base + L.tile_sum);
    a.sorted = reinterpret_cast<int*>(base + L.sorted);
    a.sorted_atoms = reinterpret_cast<float4*>(base + L.sorted_atoms);
    a.counts = counts_dev;
    a.subset = nullptr;
    a.n_subset = 0;
    hipLaunchKernelGGL(sasa_tables_kernel, dim3(1), dim3(MAX_TYPES), 0, s, rt,
                       reinterpret_cast<float*>(base + L.r2), reinterpret_cast<float*>(base + L.rext));
    if (hipGetLastError() != hipSuccess) return RSX_HIP_ERROR;
    return rsx::launch_sasa(a, s) == hipSuccess ? RSX_OK : RSX_HIP_ERROR;
}
