// af_direct.hip -- direct pressure solve (no counterpart in the Linux
// reference build; the Darwin build's EnableSparseSolver, sparse_darwin.go:19-172).
//
// The reference assembles the 7-point Laplacian of the interior cells with
// homogeneous Neumann boundaries, L = (neighbour count on the diagonal, -1
// off it) / H^2, pins the first cell by adding 1/H^2 to its diagonal, and
// solves (L + e0 e0^T / H^2) p = -Div with a sparse Cholesky factorisation.
// L is diagonalised by the orthonormal DCT-II along each axis (eigenvalues
// (2 - 2 cos(pi k / N)) / H^2 summed over the axes), so the same system is
// solved exactly, up to rounding, by
//   S = sum(b),  q = L^+ (b - S e0)  (zero mean; the right-hand side has sum 0),
//   p = q + (H^2 S - q_0) 1,
// which satisfies L p = b - S e0 and p_0 = H^2 S, i.e. the pinned equations.
// The three 1-D transforms and their inverses are dense FP64 GEMMs against
// N x N DCT matrices (rocBLAS), the spectral division and the gathers are
// small kernels here.  Sums use a fixed order.
#include "af_direct.h"

#include <cmath>
#include <vector>

#include <algorithm>

#include "af_bodies.h"
#include "af_dct.h"
#include "af_xcd.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int kSumBlocks = 1024;

__global__ __launch_bounds__(256) void k_gather_rhs(DField D, Geom G, double *b)
{
    const int64_t n = (int64_t)G.NX * G.NY * G.NZ;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const int x = (int)(i % G.NX);
        const int64_t r = i / G.NX;
        const int y = (int)(r % G.NY), z = (int)(r / G.NY);
        b[i] = -body::gp(D.d)[body::idx(D, x + G.g, y + G.g, z + G.g)];
    }
}

// Block b sums elements [b*chunk, (b+1)*chunk) in a fixed order.
__global__ __launch_bounds__(256) void k_sum_partial(const double *v, int64_t n, double *partials)
{
    __shared__ double sh[256];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    double acc = 0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) acc += v[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// S = sum of the partials (fixed tree); b[0] -= S.
__global__ __launch_bounds__(256) void k_sum_final(const double *partials, int np, double *sum,
                                                   double *b)
{
    __shared__ double sh[256];
    double acc = 0;
    for (int i = threadIdx.x; i < np; i += 256) acc += partials[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *sum = sh[0];
        b[0] = b[0] - sh[0];
    }
}

// Spectral division; the constant mode (eigenvalue 0) becomes 0.
__global__ __launch_bounds__(256) void k_divide(double *v, const double *lx, const double *ly,
                                                const double *lz, int nx, int ny, int nz)
{
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const int x = (int)(i % nx);
        const int64_t r = i / nx;
        const int y = (int)(r % ny), z = (int)(r / ny);
        const double lam = (lx[x] + ly[y]) + lz[z];
        v[i] = i == 0 ? 0.0 : v[i] / lam;
    }
}

// P interior = q + (H^2 S - q_0).
__global__ __launch_bounds__(256) void k_scatter(const double *q, const double *sum, double h2,
                                                 DField P, Geom G)
{
    const int64_t n = (int64_t)G.NX * G.NY * G.NZ;
    const double shift = h2 * *sum - q[0];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
        const int x = (int)(i % G.NX);
        const int64_t r = i / G.NX;
        const int y = (int)(r % G.NY), z = (int)(r / G.NY);
        body::gp(P.d)[body::idx(P, x + G.g, y + G.g, z + G.g)] = q[i] + shift;
    }
}

// Orthonormal DCT-II matrix, column-major: C(k, i) = s_k cos(pi k (2i+1) / (2N)),
// s_0 = sqrt(1/N), s_k = sqrt(2/N).  The angle is reduced exactly in
// integers first.
std::vector<double> dct_matrix(int n)
{
    std::vector<double> c((size_t)n * n);
    const double s0 = std::sqrt(1.0 / n), s1 = std::sqrt(2.0 / n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            const int64_t m = ((int64_t)k * (2 * i + 1)) % (4LL * n);
            c[(size_t)k + (size_t)n * i] = (k == 0 ? s0 : s1) * std::cos(M_PI * (double)m / (2.0 * n));
        }
    return c;
}

// (2 - 2 cos(pi k / N)) / H^2 = 4 sin^2(pi k / (2N)) / H^2.
std::vector<double> eigenvalues(int n, double h2)
{
    std::vector<double> l((size_t)n);
    for (int k = 0; k < n; k++) {
        const double s = std::sin(M_PI * (double)k / (2.0 * n));
        l[(size_t)k] = 4.0 * s * s / h2;
    }
    return l;
}

unsigned grid_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

// Forward (inverse = false) or inverse 3-D transform from `in` to `out`
// (`tmp` used between axes): x, then y (one GEMM per z-plane, batched), then
// z; the inverse runs z, y, x with transposed matrices.
rocblas_status transform(DirectState &d, const double *in, double *tmp, double *out, bool inverse)
{
    const double one = 1.0, zero = 0.0;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int nxy = nx * ny;
    rocblas_status st;
    if (!inverse) {
        // x: out1 = Cx * M (nx x ny*nz)
        st = rocblas_dgemm(d.blas, rocblas_operation_none, rocblas_operation_none, nx, ny * nz, nx,
                           &one, d.cx, nx, in, nx, &zero, tmp, nx);
        if (st != rocblas_status_success) return st;
        // y: per z-plane, (nx x ny) * Cy^T
        st = rocblas_dgemm_strided_batched(d.blas, rocblas_operation_none,
                                           rocblas_operation_transpose, nx, ny, ny, &one, tmp, nx,
                                           (rocblas_stride)nxy, d.cy, ny, 0, &zero, out, nx,
                                           (rocblas_stride)nxy, nz);
        if (st != rocblas_status_success) return st;
        // z: (nx*ny x nz) * Cz^T
        return rocblas_dgemm(d.blas, rocblas_operation_none, rocblas_operation_transpose, nxy, nz,
                             nz, &one, out, nxy, d.cz, nz, &zero, tmp, nxy);
    }
    st = rocblas_dgemm(d.blas, rocblas_operation_none, rocblas_operation_none, nxy, nz, nz, &one,
                       in, nxy, d.cz, nz, &zero, tmp, nxy);
    if (st != rocblas_status_success) return st;
    st = rocblas_dgemm_strided_batched(d.blas, rocblas_operation_none, rocblas_operation_none, nx,
                                       ny, ny, &one, tmp, nx, (rocblas_stride)nxy, d.cy, ny, 0,
                                       &zero, out, nx, (rocblas_stride)nxy, nz);
    if (st != rocblas_status_success) return st;
    return rocblas_dgemm(d.blas, rocblas_operation_transpose, rocblas_operation_none, nx, ny * nz,
                         nx, &one, d.cx, nx, out, nx, &zero, tmp, nx);
}

}  // namespace

bool direct_fft_supported(const Geom &G)
{
    return dct_fft_supported(G.NX) && dct_fft_supported(G.NY) && dct_fft_supported(G.NZ);
}

hipError_t init_direct(DirectState &d, const Geom &G, double H, int method, hipStream_t s,
                       int *blas_status)
{
    *blas_status = 0;
    d.nx = G.NX; d.ny = G.NY; d.nz = G.NZ;
    d.h2 = H * H;
    const int64_t n = (int64_t)G.NX * G.NY * G.NZ;
    hipError_t err;
    auto up = [&](double **dst, const std::vector<double> &v) -> hipError_t {
        hipError_t e = hipMalloc(dst, v.size() * sizeof(double));
        if (e == hipSuccess)
            e = hipMemcpy(*dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice);
        return e;
    };
    if (method == DIRECT_AUTO) method = direct_fft_supported(G) ? DIRECT_FFT : DIRECT_DENSE;
    if (method == DIRECT_FFT) {
        const int ext[3] = {G.NX, G.NY, G.NZ};
        for (int ax = 0; ax < 3; ax++) {
            std::vector<double> tw, pw, wc, iq, c0((size_t)ext[ax]);
            dct_tables(ext[ax], &tw, &pw, &wc, &iq);
            for (int k = 0; k < ext[ax]; k++) c0[(size_t)k] = pw[2 * (size_t)k];  // s_k cos(pi k / 2N)
            if ((err = up(&d.tw[ax], tw)) != hipSuccess) return err;
            if ((err = up(&d.pw[ax], pw)) != hipSuccess) return err;
            if ((err = up(&d.wc[ax], wc)) != hipSuccess) return err;
            if ((err = up(&d.iq[ax], iq)) != hipSuccess) return err;
            if ((err = up(&d.c0[ax], c0)) != hipSuccess) return err;
        }
        if ((err = up(&d.lx, eigenvalues(G.NX, d.h2))) != hipSuccess) return err;
        if ((err = up(&d.ly, eigenvalues(G.NY, d.h2))) != hipSuccess) return err;
        if ((err = up(&d.lz, eigenvalues(G.NZ, d.h2))) != hipSuccess) return err;
        d.npartials = std::max(dct_blocks(G.NX, G.NY, G.NZ), dct_blocks(G.NZ, G.NX, G.NY));
        if ((err = hipMalloc(&d.w0, (size_t)n * sizeof(double))) != hipSuccess) return err;
        if ((err = hipMalloc(&d.w1, (size_t)n * sizeof(double))) != hipSuccess) return err;
        if ((err = hipMalloc(&d.partials, (size_t)d.npartials * sizeof(double))) != hipSuccess)
            return err;
        if ((err = hipMalloc(&d.sum, sizeof(double))) != hipSuccess) return err;
        if ((err = hipMalloc(&d.shift, sizeof(double))) != hipSuccess) return err;
        d.method = DIRECT_FFT;
        return hipSuccess;
    }
    if ((err = up(&d.cx, dct_matrix(G.NX))) != hipSuccess) return err;
    if ((err = up(&d.cy, dct_matrix(G.NY))) != hipSuccess) return err;
    if ((err = up(&d.cz, dct_matrix(G.NZ))) != hipSuccess) return err;
    if ((err = up(&d.lx, eigenvalues(G.NX, d.h2))) != hipSuccess) return err;
    if ((err = up(&d.ly, eigenvalues(G.NY, d.h2))) != hipSuccess) return err;
    if ((err = up(&d.lz, eigenvalues(G.NZ, d.h2))) != hipSuccess) return err;
    if ((err = hipMalloc(&d.w0, (size_t)n * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMalloc(&d.w1, (size_t)n * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMalloc(&d.partials, kSumBlocks * sizeof(double))) != hipSuccess) return err;
    if ((err = hipMalloc(&d.sum, sizeof(double))) != hipSuccess) return err;
    if ((err = hipMemset(d.w0, 0, (size_t)n * sizeof(double))) != hipSuccess) return err;
    rocblas_status st = rocblas_create_handle(&d.blas);
    if (st == rocblas_status_success) st = rocblas_set_stream(d.blas, s);
    if (st == rocblas_status_success) st = rocblas_set_pointer_mode(d.blas, rocblas_pointer_mode_host);
    // Run every transform once outside any capture (loads the GEMM kernels).
    if (st == rocblas_status_success) st = transform(d, d.w0, d.w1, d.w0, false);
    if (st == rocblas_status_success) st = transform(d, d.w0, d.w1, d.w0, true);
    if (st != rocblas_status_success) {
        *blas_status = (int)st;
        return hipErrorUnknown;
    }
    d.npartials = kSumBlocks;
    d.method = DIRECT_DENSE;
    return hipStreamSynchronize(s);
}

void release_direct(DirectState &d)
{
    if (d.blas) (void)rocblas_destroy_handle(d.blas);
    double **bufs[] = {&d.cx, &d.cy, &d.cz, &d.lx, &d.ly, &d.lz, &d.w0, &d.w1, &d.partials, &d.sum,
                       &d.shift};
    for (double **b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    for (int ax = 0; ax < 3; ax++) {
        double **t[] = {&d.tw[ax], &d.pw[ax], &d.wc[ax], &d.iq[ax], &d.c0[ax]};
        for (double **b : t) {
            if (*b) (void)hipFree(*b);
            *b = nullptr;
        }
    }
    d.blas = nullptr;
    d.method = 0;
}

namespace {

// The FFT form: five transform passes and two fixed-order reductions.  With
// `vel` the first pass computes the divergence from the velocity itself.
struct VelocitySource {
    DField U, V, W;
    double scale;
};

hipError_t enqueue_fft_solve(DirectState &d, const DField &P, const DField &Div, const Geom &G,
                             hipStream_t s, const VelocitySource *vel = nullptr)
{
    const int64_t NX = G.NX, NY = G.NY, NXY = NX * NY;
    hipError_t err;
    auto tables = [&](DctPass &p, int ax) {
        p.tw = d.tw[ax];
        p.pw = d.pw[ax];
        p.wc = d.wc[ax];
        p.iq = d.iq[ax];
    };
    // x forward: -Div (arena interior) -> w0, with block sums of the rhs.
    DctPass x;
    // Field.Index (grid.go:24-26) of the first interior cell.
    const int64_t g = G.g;
    x.in = Div.d + (g + g * Div.sy + g * (int64_t)Div.sz);
    x.in_sn = 1; x.in_sa = Div.sy; x.in_sb = Div.sz;
    x.out = d.w0;
    x.out_sn = 1; x.out_sa = NX; x.out_sb = NXY;
    x.A = G.NY; x.B = G.NZ;
    x.sign = -1.0;
    x.partials = d.partials;
    tables(x, 0);
    if (vel) {
        x.in = nullptr;
        x.U = vel->U; x.V = vel->V; x.W = vel->W; x.D = Div;
        x.G = G;
        x.dscale = vel->scale;
    }
    if ((err = launch_dct_pass(vel ? 5 : 1, G.NX, x, s)) != hipSuccess) return err;
    if ((err = launch_reduce_partials(d.partials, dct_blocks(G.NX, G.NY, G.NZ), nullptr, 0.0, 0,
                                      d.sum, s)) != hipSuccess)
        return err;
    // y forward: w0 -> w1.
    DctPass y;
    y.in = d.w0; y.out = d.w1;
    y.in_sn = y.out_sn = NX; y.in_sa = y.out_sa = 1; y.in_sb = y.out_sb = NXY;
    y.A = G.NX; y.B = G.NZ;
    tables(y, 1);
    if ((err = launch_dct_pass(0, G.NY, y, s)) != hipSuccess) return err;
    // z forward, division, z inverse, in place in w1.
    DctPass z;
    z.in = d.w1; z.out = d.w1;
    z.in_sn = z.out_sn = NXY; z.in_sa = z.out_sa = 1; z.in_sb = z.out_sb = NX;
    z.A = G.NX; z.B = G.NY;
    z.lx = d.lx; z.ly = d.ly; z.lz = d.lz;
    z.c0x = d.c0[0]; z.c0y = d.c0[1]; z.c0z = d.c0[2];
    z.sum = d.sum;
    z.partials = d.partials;
    tables(z, 2);
    if ((err = launch_dct_pass(2, G.NZ, z, s)) != hipSuccess) return err;
    if ((err = launch_reduce_partials(d.partials, dct_blocks(G.NZ, G.NX, G.NY), d.sum, d.h2, 1,
                                      d.shift, s)) != hipSuccess)
        return err;
    // y inverse: w1 -> w0.
    y.in = d.w1; y.out = d.w0;
    if ((err = launch_dct_pass(3, G.NY, y, s)) != hipSuccess) return err;
    // x inverse: w0 -> P's interior, plus the shift.
    DctPass xi;
    xi.in = d.w0;
    xi.in_sn = 1; xi.in_sa = NX; xi.in_sb = NXY;
    xi.out = P.d + (g + g * P.sy + g * (int64_t)P.sz);
    xi.out_sn = 1; xi.out_sa = P.sy; xi.out_sb = P.sz;
    xi.A = G.NY; xi.B = G.NZ;
    xi.shift = d.shift;
    tables(xi, 0);
    if ((err = launch_dct_pass(4, G.NX, xi, s)) != hipSuccess) return err;
    return launch_pressure_bc(P, G, false, s);
}

}  // namespace

hipError_t enqueue_direct_project_solve(DirectState &d, const DField &P, const DField &Div,
                                        const DField &U, const DField &V, const DField &W,
                                        double scale, const Geom &G, hipStream_t s)
{
    if (d.method != DIRECT_FFT) return hipErrorInvalidValue;
    const VelocitySource vel{U, V, W, scale};
    return enqueue_fft_solve(d, P, Div, G, s, &vel);
}

hipError_t enqueue_direct_solve(DirectState &d, const DField &P, const DField &Div, const Geom &G,
                                hipStream_t s, int *blas_status)
{
    *blas_status = 0;
    if (d.method == DIRECT_FFT) return enqueue_fft_solve(d, P, Div, G, s);
    const int64_t n = (int64_t)G.NX * G.NY * G.NZ;
    const unsigned gb = grid_for(n);
    const int np = (int)(n < kSumBlocks ? n : kSumBlocks);
    hipLaunchKernelGGL(k_gather_rhs, dim3(gb), dim3(256), 0, s, Div, G, d.w0);
    hipLaunchKernelGGL(k_sum_partial, dim3(np), dim3(256), 0, s, d.w0, n, d.partials);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, d.partials, np, d.sum, d.w0);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    // w0 -> (w1 as scratch) -> w1: forward transform; the result lands in w1.
    rocblas_status st = transform(d, d.w0, d.w1, d.w0, false);
    // transform() leaves its result in its `tmp` argument (w1).
    if (st != rocblas_status_success) {
        *blas_status = (int)st;
        return hipErrorUnknown;
    }
    hipLaunchKernelGGL(k_divide, dim3(gb), dim3(256), 0, s, d.w1, d.lx, d.ly, d.lz, G.NX, G.NY,
                       G.NZ);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    st = transform(d, d.w1, d.w0, d.w1, true);
    if (st != rocblas_status_success) {
        *blas_status = (int)st;
        return hipErrorUnknown;
    }
    // Inverse result in w0.
    hipLaunchKernelGGL(k_scatter, dim3(gb), dim3(256), 0, s, d.w0, d.sum, d.h2, P, G);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    return launch_pressure_bc(P, G, false, s);
}

}  // namespace af
