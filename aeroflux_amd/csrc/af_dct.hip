// af_dct.hip -- the direct pressure solve's orthonormal DCT-II transforms by
// FFT, for grids whose three extents are powers of two (8 .. 512).
//
// One 1-D transform along an axis: a workgroup stages a tile of L lines of
// length N in LDS and transforms two real lines as one complex N-point FFT
// (line 2p in the real part, 2p+1 in the imaginary part).
//   DCT-II (Makhoul, IEEE TASSP 28(1), 1980): v[m] = x[2m],
//   v[N-1-m] = x[2m+1]; V = FFT(v); X[k] = s_k Re(e^{-i pi k / 2N} V[k]),
//   s_0 = sqrt(1/N), s_k = sqrt(2/N).  The two packed lines separate as
//   V_a = (Z[k] + conj Z[N-k]) / 2, V_b = -i (Z[k] - conj Z[N-k]) / 2.
//   Inverse: Z[k] = e^{i pi k / 2N} ((Y_a[k] - i Y_a[N-k]) + i (Y_b[k] -
//   i Y_b[N-k])), Y[k] = q_k X[k], q_0 = sqrt(1/N), q_k = sqrt(1/2N),
//   Y[N] = 0; z = unnormalised inverse FFT of Z; x_a[n] = Re z[m(n)],
//   x_b[n] = Im z[m(n)], m the reordering above.
// The FFT is Stockham's autosort form: radix-4 stages, then one radix-2
// stage when log2 N is odd, each stage reading its inputs into registers
// and writing its outputs back to the same LDS array after a barrier.
// The inverse FFT is conj(FFT(conj Z)).
//
// The solve (the pinned system of sparse_darwin.go:19-172, see af_direct.hip)
// runs five such passes over the compact interior array:
//   x forward  (reads -Div from the arena; block partial sums of the rhs)
//   y forward
//   z forward, spectral division (with the pin's -S e0 term removed in
//     the spectral domain), z inverse (block partials of q_0 as a spectral
//     dot product)
//   y inverse
//   x inverse  (writes P's interior plus H^2 S - q_0)
// with fixed-order reductions of the partials in between.
#include "af_dct.h"

#include <cmath>
#include <vector>

#include "af_bodies.h"

namespace af {
namespace {

enum Mode { FWD = 0, FWD_SUM = 1, SOLVE = 2, INV = 3, INV_SHIFT = 4, FWD_DIV = 5 };

template <int N>
struct Shape {
    static constexpr int T = N / 4;                 // threads per FFT
    static constexpr int L = N >= 512 ? 8 : (2048 / N > 16 ? 2048 / N : 16);  // lines per tile
    static constexpr int P = L / 2;                 // line pairs (complex FFTs) per tile
    // 256 threads: each takes one radix-4 butterfly of C = 1024/N pairs at
    // a time, in R rounds.  (512 threads, one round, measured slower at
    // N = 256: the fused solve pass then fits one workgroup per CU.)
    static constexpr int THREADS = 256;
    static constexpr int C = THREADS / T;           // FFTs in flight per workgroup
    static constexpr int R = P / C;                 // rounds of FFTs
    static constexpr int LOG = N == 8 ? 3 : N == 16 ? 4 : N == 32 ? 5 : N == 64 ? 6
                             : N == 128 ? 7 : N == 256 ? 8 : 9;
    static constexpr int NP = N + 1;                // LDS stride of a pair (padded: the
                                                    // strided tile loads hit distinct banks)
    static constexpr int PER = L * N / THREADS;     // tile elements per thread
    static_assert(P % C == 0, "pairs must be a multiple of the FFTs in flight");
    static_assert((1 << LOG) == N, "N must be a power of two in 8..512");
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// Entry i of a table of (re, im) pairs in global memory.
template <typename TP>
__device__ __forceinline__ double2 tab(TP p, int i)
{
    return make_double2(p[2 * i], p[2 * i + 1]);
}

template <int N>
__device__ __forceinline__ int reorder(int n)  // Makhoul's m(n)
{
    return (n & 1) ? N - 1 - (n >> 1) : (n >> 1);
}

// In-place forward FFT of the tile's P lines of N complex values.
template <int N, typename TW>
__device__ void fft_tile(double2 *buf, TW tw)
{
    using S = Shape<N>;
    const int g = threadIdx.x / S::T, j = threadIdx.x % S::T;
    int ns = 1;
#pragma unroll
    for (int s = 0; s < S::LOG / 2; s++) {
        double2 v[S::R][4];
#pragma unroll
        for (int r = 0; r < S::R; r++) {
            const double2 *z = buf + (g + r * S::C) * S::NP;
#pragma unroll
            for (int i = 0; i < 4; i++) v[r][i] = z[j + i * (N / 4)];
        }
        __syncthreads();
        const int k = j % ns;
        const int t = k * (N / (4 * ns));
        const double2 w1 = tab(tw, t), w2 = tab(tw, 2 * t), w3 = tab(tw, 3 * t);
        const int base = (j / ns) * 4 * ns + k;
#pragma unroll
        for (int r = 0; r < S::R; r++) {
            double2 *z = buf + (g + r * S::C) * S::NP;
            const double2 a = v[r][0], b = cmul(v[r][1], w1), c = cmul(v[r][2], w2),
                          d = cmul(v[r][3], w3);
            const double2 s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d);
            const double2 dd = csub(b, d);
            const double2 s3 = make_double2(dd.y, -dd.x);  // (b - d) * (-i)
            z[base] = cadd(s0, s2);
            z[base + ns] = cadd(s1, s3);
            z[base + 2 * ns] = csub(s0, s2);
            z[base + 3 * ns] = csub(s1, s3);
        }
        __syncthreads();
        ns *= 4;
    }
    if (S::LOG & 1) {  // last stage, radix 2, ns = N/2: in place per thread
#pragma unroll
        for (int r = 0; r < S::R; r++) {
            double2 *z = buf + (g + r * S::C) * S::NP;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int jj = j + h * S::T;
                const double2 a = z[jj], b = cmul(z[jj + N / 2], tab(tw, jj));
                z[jj] = cadd(a, b);
                z[jj + N / 2] = csub(a, b);
            }
        }
        __syncthreads();
    }
}

template <int N>
__device__ __forceinline__ double block_sum(double v, double *red)
{
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = Shape<N>::THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

template <int N, int MODE>
__global__ __launch_bounds__(Shape<N>::THREADS) void k_dct(DctPass a)
{
    using S = Shape<N>;
    __shared__ double2 buf[S::P * S::NP];
    __shared__ double red[S::THREADS];
    const int tiles_a = (a.A + S::L - 1) / S::L;
    const int b = blockIdx.x / tiles_a;
    const int a0 = (blockIdx.x % tiles_a) * S::L;
    const auto in = body::gp(a.in);
    const auto out = body::gp(a.out);
    const auto tw = body::gp(a.tw), pw = body::gp(a.pw), wc = body::gp(a.wc);
    const auto iq = body::gp(a.iq);
    const bool inverse = MODE == INV || MODE == INV_SHIFT;
    double acc = 0;

    // Load the tile: coalesced along the line when lines are contiguous,
    // across lines otherwise.  Every load is issued before the first LDS
    // write, so their latencies overlap.
    const bool in_rows = a.in_sn == 1;
    const auto at = in + (int64_t)a0 * a.in_sa + (int64_t)b * a.in_sb;
    double vals[S::PER];
#pragma unroll
    for (int r = 0; r < S::PER; r++) {
        const int i = threadIdx.x + r * S::THREADS;
        const int l = in_rows ? i / N : i % S::L, n = in_rows ? i % N : i / S::L;
        if (MODE == FWD_DIV) {
            // computeDivergence (project.go:42-62) at interior cell
            // (n, a0 + l, b), the operations of body::divergence_cell.
            double dv = 0.0;
            if (a0 + l < a.A) {
                const int g = a.G.g, x = n + g, y = a0 + l + g, z = b + g;
                const auto u = body::gp(a.U.d), v = body::gp(a.V.d), w = body::gp(a.W.d);
                const double du = u[body::idx(a.U, x + 1, y, z)] - u[body::idx(a.U, x, y, z)];
                const double dvv = v[body::idx(a.V, x, y + 1, z)] - v[body::idx(a.V, x, y, z)];
                const double dw = w[body::idx(a.W, x, y, z + 1)] - w[body::idx(a.W, x, y, z)];
                dv = a.dscale * (du + dvv + dw);
                body::gp(a.D.d)[body::idx(a.D, x, y, z)] = dv;
            }
            vals[r] = dv;
        } else {
            vals[r] = a0 + l < a.A ? at[n * a.in_sn + (int64_t)l * a.in_sa] : 0.0;
        }
    }
#pragma unroll
    for (int r = 0; r < S::PER; r++) {
        const int i = threadIdx.x + r * S::THREADS;
        const int l = in_rows ? i / N : i % S::L, n = in_rows ? i % N : i / S::L;
        const double v = a.sign * vals[r];
        if (MODE == FWD_SUM || MODE == FWD_DIV) acc += v;
        const int pos = inverse ? n : reorder<N>(n);
        double *slot = reinterpret_cast<double *>(&buf[(l >> 1) * S::NP + pos]);
        slot[l & 1] = v;
    }
    __syncthreads();

    const int g = threadIdx.x / S::T, j = threadIdx.x % S::T;
    if (!inverse) {
        fft_tile<N>(buf, tw);
        // Separate the packed lines and apply the post-twiddle; each thread
        // owns slots k = j + q T, q = 0..3, of its pairs.
        double2 xo[S::R][4];
#pragma unroll
        for (int r = 0; r < S::R; r++) {
            const double2 *z = buf + (g + r * S::C) * S::NP;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = j + q * S::T;
                const double2 zk = z[k], zr = z[(N - k) & (N - 1)];
                const double2 va = make_double2(0.5 * (zk.x + zr.x), 0.5 * (zk.y - zr.y));
                const double2 vb = make_double2(0.5 * (zk.y + zr.y), 0.5 * (zr.x - zk.x));
                const double2 w = tab(pw, k);
                xo[r][q] = make_double2(w.x * va.x - w.y * va.y, w.x * vb.x - w.y * vb.y);
            }
        }
        __syncthreads();
        if (MODE == SOLVE) {
            // Q = (X - S c0x c0y c0z) / lambda, the zero mode and lines past
            // the extent 0; q0 accumulates Q c0x c0y c0z.
            const double Ssum = *body::gp(a.sum);
            const auto c0x = body::gp(a.c0x), c0z = body::gp(a.c0z);
            const auto lx = body::gp(a.lx), lz = body::gp(a.lz);
            const double c0y = body::gp(a.c0y)[b], ly = body::gp(a.ly)[b];
#pragma unroll
            for (int r = 0; r < S::R; r++) {
                const int p = g + r * S::C;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int kx = a0 + 2 * p + h;
                    const bool live = kx < a.A;
                    const double cxy = live ? c0x[kx] * c0y : 0.0;
                    const double lxy = live ? lx[kx] + ly : 0.0;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int k = j + q * S::T;
                        const double e0 = cxy * c0z[k];
                        const double lam = lxy + lz[k];
                        double x = h ? xo[r][q].y : xo[r][q].x;
                        x = (live && lam > 0) ? (x - Ssum * e0) / lam : 0.0;
                        acc += x * e0;
                        if (h) xo[r][q].y = x; else xo[r][q].x = x;
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < S::R; r++) {
            double2 *z = buf + (g + r * S::C) * S::NP;
#pragma unroll
            for (int q = 0; q < 4; q++) z[j + q * S::T] = xo[r][q];
        }
        __syncthreads();
    }
    if (MODE == SOLVE || inverse) {
        // Pre-twiddle: conj Z[k] from the coefficients at k and N-k.
        double2 zc[S::R][4];
#pragma unroll
        for (int r = 0; r < S::R; r++) {
            const double2 *z = buf + (g + r * S::C) * S::NP;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k = j + q * S::T;
                const double2 yk = z[k];
                const double qk = iq[k];
                double2 yr = make_double2(0.0, 0.0);
                if (k > 0) {
                    const double qr = iq[N - k];
                    yr = z[N - k];
                    yr.x *= qr;
                    yr.y *= qr;
                }
                // (Ya[k] + Yb[N-k]) + i (Yb[k] - Ya[N-k])
                const double2 u = make_double2(qk * yk.x + yr.y, qk * yk.y - yr.x);
                const double2 zz = cmul(tab(wc, k), u);
                zc[r][q] = make_double2(zz.x, -zz.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < S::R; r++) {
            double2 *z = buf + (g + r * S::C) * S::NP;
#pragma unroll
            for (int q = 0; q < 4; q++) z[j + q * S::T] = zc[r][q];
        }
        __syncthreads();
        fft_tile<N>(buf, tw);
    }

    // Store the tile (the inverse's result is conj of the FFT's output).
    const double shift = MODE == INV_SHIFT ? *body::gp(a.shift) : 0.0;
    const bool out_rows = a.out_sn == 1;
    const auto to = out + (int64_t)a0 * a.out_sa + (int64_t)b * a.out_sb;
#pragma unroll
    for (int r = 0; r < S::PER; r++) {
        const int i = threadIdx.x + r * S::THREADS;
        const int l = out_rows ? i / N : i % S::L, n = out_rows ? i % N : i / S::L;
        if (a0 + l >= a.A) continue;
        const double *slot;
        double v;
        if (MODE == SOLVE || inverse) {
            slot = reinterpret_cast<const double *>(&buf[(l >> 1) * S::NP + reorder<N>(n)]);
            v = (l & 1) ? -slot[1] : slot[0];
        } else {
            slot = reinterpret_cast<const double *>(&buf[(l >> 1) * S::NP + n]);
            v = slot[l & 1];
        }
        to[n * a.out_sn + (int64_t)l * a.out_sa] = v + shift;
    }
    if (MODE == FWD_SUM || MODE == FWD_DIV || MODE == SOLVE) {
        const double s = block_sum<N>(acc, red);
        if (threadIdx.x == 0) body::gp(a.partials)[blockIdx.x] = s;
    }
}

// Fixed-order sum of the partials: thread t adds t, t+256, ...; then a tree.
// mode 0: *out0 = sum (S).  mode 1: *out0 = h2 * (*sum) - sum (the shift).
__global__ __launch_bounds__(256) void k_reduce_partials(const double *partials, int n,
                                                         const double *sum, double h2, int mode,
                                                         double *out0)
{
    __shared__ double red[256];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partials[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out0 = mode == 0 ? red[0] : h2 * *sum - red[0];
}

template <int MODE>
hipError_t launch_mode(int n, const DctPass &p, hipStream_t s)
{
    switch (n) {
#define AF_DCT_CASE(NN)                                                                        \
    case NN:                                                                                   \
        hipLaunchKernelGGL((k_dct<NN, MODE>), dim3(dct_blocks(NN, p.A, p.B)),                 \
                           dim3(Shape<NN>::THREADS), 0, s, p);                                 \
        break;
        AF_DCT_CASE(8)
        AF_DCT_CASE(16)
        AF_DCT_CASE(32)
        AF_DCT_CASE(64)
        AF_DCT_CASE(128)
        AF_DCT_CASE(256)
        AF_DCT_CASE(512)
#undef AF_DCT_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int lines_per_tile(int n)
{
    switch (n) {
    case 8: return Shape<8>::L;
    case 16: return Shape<16>::L;
    case 32: return Shape<32>::L;
    case 64: return Shape<64>::L;
    case 128: return Shape<128>::L;
    case 256: return Shape<256>::L;
    case 512: return Shape<512>::L;
    default: return 0;
    }
}

}  // namespace

bool dct_fft_supported(int n)
{
    return lines_per_tile(n) > 0;
}

int dct_blocks(int n, int A, int B)
{
    const int L = lines_per_tile(n);
    return L ? ((A + L - 1) / L) * B : 0;
}

void dct_tables(int n, std::vector<double> *tw, std::vector<double> *pw, std::vector<double> *wc,
                std::vector<double> *iq)
{
    tw->resize(2 * (size_t)n);
    pw->resize(2 * (size_t)n);
    wc->resize(2 * (size_t)n);
    iq->resize((size_t)n);
    for (int t = 0; t < n; t++) {
        // e^{-2 pi i t / n}: the angle pi * (2t) / n with 2t reduced mod 2n
        // exactly (it is already below 2n).
        const double ang = M_PI * (double)(2 * t) / (double)n;
        (*tw)[2 * t] = std::cos(ang);
        (*tw)[2 * t + 1] = -std::sin(ang);
        const double h = M_PI * (double)t / (2.0 * n);
        const double s = t == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
        (*pw)[2 * t] = s * std::cos(h);
        (*pw)[2 * t + 1] = -s * std::sin(h);
        (*wc)[2 * t] = std::cos(h);
        (*wc)[2 * t + 1] = std::sin(h);
        (*iq)[t] = t == 0 ? std::sqrt(1.0 / n) : std::sqrt(1.0 / (2.0 * n));
    }
}

hipError_t launch_dct_pass(int mode, int n, const DctPass &p, hipStream_t s)
{
    switch (mode) {
    case 0: return launch_mode<FWD>(n, p, s);
    case 1: return launch_mode<FWD_SUM>(n, p, s);
    case 2: return launch_mode<SOLVE>(n, p, s);
    case 3: return launch_mode<INV>(n, p, s);
    case 4: return launch_mode<INV_SHIFT>(n, p, s);
    case 5: return launch_mode<FWD_DIV>(n, p, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reduce_partials(const double *partials, int n, const double *sum, double h2,
                                  int mode, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, partials, n, sum, h2, mode,
                       out);
    return hipGetLastError();
}

}  // namespace af
