// af_ic.hip -- seeded band-limited initial velocity, generated on the GPU.
//
// No reference counterpart: the reference seeds fields by writing Mesh.Data
// (engine_test.go:302-335).  Served jobs need a fresh random field per
// admission, which the host cannot afford to build, so the device builds it:
//
//   u_c(x) = sum over integer k, 1 <= |k| <= kmax, of
//            |k|^slope * Re( exp(i phi_{c,k}) exp(i 2 pi k.x / L) ),
//
// evaluated at each component's own interior face positions (u at
// ((i-g)H, (j-g+1/2)H, (k-g+1/2)H), and likewise), then scaled so that the
// mean square over all interior faces of U, V and W is rms^2.  The phases
// are 2 pi times SplitMix64 uniforms -- component c, mode (kx, ky, kz) in
// C order over [-kmax, kmax]^3 takes output c*(2kmax+1)^3 + index + 1 of the
// generator started at `seed` -- the same stream the host generator
// aeroflux_amd.ic.band_limited draws, so both build the same field up to
// summation order.  The sum is separable: three passes of (2kmax+1) complex
// multiply-adds per point (over kz, then ky, then kx), O((2kmax+1) N^3).
#include <hip/hip_runtime.h>

#include <cmath>

#include "af_bodies.h"
#include "af_ic.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int RB = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// t1[kx][ky][z] = sum_kz coef[kx][ky][kz] * ez[kz][z]
__global__ __launch_bounds__(RB) void k_ic_t1(const double2 *coef, const double2 *ez, int nk,
                                              int nz, double2 *t1)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= (int64_t)nk * nk * nz) return;
    const int z = (int)(i % nz);
    const int kxy = (int)(i / nz);
    double2 s = make_double2(0.0, 0.0);
    for (int kz = 0; kz < nk; kz++) {
        const double2 p = cmul(coef[kxy * nk + kz], ez[kz * nz + z]);
        s.x += p.x;
        s.y += p.y;
    }
    t1[i] = s;
}

// t2[kx][z][y] = sum_ky t1[kx][ky][z] * ey[ky][y]
__global__ __launch_bounds__(RB) void k_ic_t2(const double2 *t1, const double2 *ey, int nk, int nz,
                                              int ny, double2 *t2)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= (int64_t)nk * nz * ny) return;
    const int y = (int)(i % ny);
    const int z = (int)((i / ny) % nz);
    const int kx = (int)(i / ((int64_t)ny * nz));
    double2 s = make_double2(0.0, 0.0);
    for (int ky = 0; ky < nk; ky++) {
        const double2 p = cmul(t1[((int64_t)kx * nk + ky) * nz + z], ey[ky * ny + y]);
        s.x += p.x;
        s.y += p.y;
    }
    t2[i] = s;
}

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    return sh[0];
}

// f(x0+x, y0+y, z0+z) = Re sum_kx t2[kx][z][y] * ex[kx][x]; one partial sum
// of squares per block.
__global__ __launch_bounds__(RB) void k_ic_val(const double2 *t2, const double2 *ex, int nk,
                                               int nx, int ny, int nz, DField f, int x0, int y0,
                                               int z0, double *partials)
{
    __shared__ double sh[RB];
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    double sq = 0;
    if (i < (int64_t)nx * ny * nz) {
        const int x = (int)(i % nx);
        const int y = (int)((i / nx) % ny);
        const int z = (int)(i / ((int64_t)nx * ny));
        double s = 0;
        for (int kx = 0; kx < nk; kx++) {
            const double2 a = t2[((int64_t)kx * nz + z) * ny + y];
            const double2 b = ex[kx * nx + x];
            s += a.x * b.x - a.y * b.y;
        }
        f.d[(x0 + x) + (y0 + y) * f.sy + (z0 + z) * f.sz] = s;
        sq = s * s;
    }
    const double b = block_sum(sq, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// scale = rms / sqrt(sum of squares / count), folded in a fixed order.
__global__ __launch_bounds__(RB) void k_ic_scale(const double *partials, int n, double count,
                                                 double rms, double *scale)
{
    __shared__ double sh[RB];
    double v = 0;
    for (int k = threadIdx.x; k < n; k += RB) v += partials[k];
    const double s = block_sum(v, sh);
    if (threadIdx.x == 0) *scale = rms / sqrt(s / count);
}

__global__ __launch_bounds__(RB) void k_ic_apply(DField f, int x0, int y0, int z0, int nx, int ny,
                                                 int nz, const double *scale)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= (int64_t)nx * ny * nz) return;
    const int x = (int)(i % nx);
    const int y = (int)((i / nx) % ny);
    const int z = (int)(i / ((int64_t)nx * ny));
    double &v = f.d[(x0 + x) + (y0 + y) * f.sy + (z0 + z) * f.sz];
    v = v * *scale;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + RB - 1) / RB); }

// SplitMix64 output `k` (1-based) of the generator started at `seed`.
__device__ __forceinline__ uint64_t splitmix64(uint64_t seed, uint64_t k)
{
    uint64_t z = seed + k * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Mode coefficients |k|^slope exp(i phi) of the three components, modes in
// C order over (kx, ky, kz) in [-kmax, kmax]^3.
__global__ __launch_bounds__(RB) void k_ic_coef(uint64_t seed, int kmax, double slope,
                                                double2 *coef)
{
    const int nk = 2 * kmax + 1;
    const int64_t nmode = (int64_t)nk * nk * nk;
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= 3 * nmode) return;
    const int64_t m = i % nmode;
    const int kx = (int)(m / (nk * nk)) - kmax, ky = (int)((m / nk) % nk) - kmax,
              kz = (int)(m % nk) - kmax;
    const double kk = sqrt((double)(kx * kx + ky * ky + kz * kz));
    const double amp = (kk >= 1 && kk <= kmax) ? pow(kk, slope) : 0.0;
    const uint64_t r = splitmix64(seed, (uint64_t)(i + 1));
    const double phase = (2.0 * M_PI) * ((double)(r >> 11) * 0x1.0p-53);
    double sn, cs;
    sincos(phase, &sn, &cs);
    coef[i] = make_double2(amp * cs, amp * sn);
}

// Phase table exp(i 2 pi k pos / L) of one axis: entry [k][j] for
// pos = (j + off) * h, k in [-kmax, kmax].
__global__ __launch_bounds__(RB) void k_ic_tab(int kmax, int n, double off, double h, double L,
                                               double2 *tab)
{
    const int nk = 2 * kmax + 1;
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= (int64_t)nk * n) return;
    const int k = (int)(i / n) - kmax, j = (int)(i % n);
    const double pos = ((double)j + off) * h;
    double sn, cs;
    sincos((2.0 * M_PI) * (double)k * pos / L, &sn, &cs);
    tab[i] = make_double2(cs, sn);
}

struct Comp {
    int x0, y0, z0, nx, ny, nz;  // interior face range
    double ox, oy, oz;           // face offsets in cells
};

Comp comp_range(const Geom &G, int c)
{
    const int g = G.g;
    Comp r{g, g, g, G.NX + (c == 0), G.NY + (c == 1), G.NZ + (c == 2), 0.5, 0.5, 0.5};
    if (c == 0) r.ox = 0.0;
    if (c == 1) r.oy = 0.0;
    if (c == 2) r.oz = 0.0;
    return r;
}

}  // namespace

int64_t ic_scratch_bytes(Geom G, int kmax)
{
    const int64_t nk = 2 * kmax + 1;
    const int64_t n1 = G.NX + 1, n2 = G.NY + 1, n3 = G.NZ + 1;
    const int64_t coef = 3 * nk * nk * nk, tabs = 3 * nk * (n1 + n2 + n3);
    const int64_t t1 = nk * nk * n3, t2 = nk * n3 * n2;
    const int64_t parts = (n1 * n2 * n3 + RB - 1) / RB * 3 + 1;
    return 16 * (coef + tabs + t1 + t2) + 8 * parts + 256;
}

hipError_t ICScratch::reserve(int64_t bytes)
{
    if (bytes <= dev_bytes) return hipSuccess;
    hipError_t err;
    if (dev) {
        if ((err = hipFree(dev)) != hipSuccess) return err;
        dev = nullptr;
        dev_bytes = 0;
    }
    if ((err = hipMalloc(&dev, (size_t)bytes)) != hipSuccess) return err;
    dev_bytes = bytes;
    return hipSuccess;
}

void ICScratch::release()
{
    if (dev) (void)hipFree(dev);
    *this = ICScratch();
}

hipError_t launch_band_limited(DField U, DField V, DField W, Geom G, double h, uint64_t seed,
                               int kmax, double slope, double rms, ICScratch &sc, hipStream_t s)
{
    const int nk = 2 * kmax + 1;
    const int64_t nmode = (int64_t)nk * nk * nk;
    hipError_t err = sc.reserve(ic_scratch_bytes(G, kmax));
    if (err != hipSuccess) return err;
    double2 *dc = static_cast<double2 *>(sc.dev);
    double2 *dt = dc + 3 * nmode;
    hipLaunchKernelGGL(k_ic_coef, dim3(blocks(3 * nmode)), dim3(RB), 0, s, seed, kmax, slope, dc);
    // Phase tables: ex, ey, ez of each component.
    const double L[3] = {G.NX * h, G.NY * h, G.NZ * h};
    int64_t toff[3][3];
    int64_t at = 0;
    for (int c = 0; c < 3; c++) {
        const Comp r = comp_range(G, c);
        const int n[3] = {r.nx, r.ny, r.nz};
        const double o[3] = {r.ox, r.oy, r.oz};
        for (int a = 0; a < 3; a++) {
            toff[c][a] = at;
            hipLaunchKernelGGL(k_ic_tab, dim3(blocks((int64_t)nk * n[a])), dim3(RB), 0, s, kmax,
                               n[a], o[a], h, L[a], dt + at);
            at += (int64_t)nk * n[a];
        }
    }
    double2 *t1 = dt + at;
    double2 *t2 = t1 + (int64_t)nk * nk * (G.NZ + 1);
    double *partials = reinterpret_cast<double *>(t2 + (int64_t)nk * (G.NZ + 1) * (G.NY + 1));
    const DField F[3] = {U, V, W};
    int64_t np = 0, count = 0;
    for (int c = 0; c < 3; c++) {
        const Comp r = comp_range(G, c);
        const double2 *ex = dt + toff[c][0], *ey = dt + toff[c][1], *ez = dt + toff[c][2];
        hipLaunchKernelGGL(k_ic_t1, dim3(blocks((int64_t)nk * nk * r.nz)), dim3(RB), 0, s,
                           dc + c * nmode, ez, nk, r.nz, t1);
        hipLaunchKernelGGL(k_ic_t2, dim3(blocks((int64_t)nk * r.nz * r.ny)), dim3(RB), 0, s, t1,
                           ey, nk, r.nz, r.ny, t2);
        const int64_t n = (int64_t)r.nx * r.ny * r.nz;
        hipLaunchKernelGGL(k_ic_val, dim3(blocks(n)), dim3(RB), 0, s, t2, ex, nk, r.nx, r.ny,
                           r.nz, F[c], r.x0, r.y0, r.z0, partials + np);
        np += blocks(n);
        count += n;
    }
    double *scale = partials + np;
    hipLaunchKernelGGL(k_ic_scale, dim3(1), dim3(RB), 0, s, partials, (int)np, (double)count, rms,
                       scale);
    for (int c = 0; c < 3; c++) {
        const Comp r = comp_range(G, c);
        hipLaunchKernelGGL(k_ic_apply, dim3(blocks((int64_t)r.nx * r.ny * r.nz)), dim3(RB), 0, s,
                           F[c], r.x0, r.y0, r.z0, r.nx, r.ny, r.nz, scale);
    }
    return hipGetLastError();
}

}  // namespace af
