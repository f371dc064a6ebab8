// af_rbgs_tb.hip -- temporally blocked red-black Gauss-Seidel for gfx950.
//
// One launch applies K consecutive colour passes (gaussSeidelSweep,
// project.go:103-130, colours alternating) to a pressure-like field, reading
// `src` and writing `dst` (ping-pong: a tile reads halo cells that a
// neighbouring tile owns, so updates cannot be written in place).
//
// Work decomposition.  A workgroup of 64 x 16 threads owns a haloed tile of
// 128 (x) x TY (y) columns and streams a z-chunk of the grid plane by plane.
// Each lane holds an x-pair (x, x+1) -- always one cell of each colour, so
// every lane updates exactly one cell per plane per pass -- for R rows.  At
// step t the tile loads plane t, then pass j (j = 1..K) updates plane t-j.
// A colour pass reads only cells of the other colour, which that pass never
// writes, so one copy of each plane suffices: when pass j reaches plane q,
// plane q+1 has just been given pass j-1 and plane q-1 pass j, and the
// cells pass j reads there are the other colour's, unchanged since pass
// j-1.  The column window of K+2 planes lives in registers; in-plane
// neighbours (other colour) are exchanged through LDS, one buffer per pass.
//
// Each pass shrinks the region of correct values by one cell in x, y and z,
// so a tile loads a halo of K cells and stores only its core; a z-chunk
// loads K planes below and above.  Neighbours outside the interior are
// replaced by the cell's own value: applyPressureBC (project.go:139-163)
// runs before every colour pass and sets that ghost to exactly the cell's
// current value, so the results are bit-identical to the reference's
// sequence of BC fills and single passes, for any tile shape.
//
// T = double: p = (sum - h2*rhs) / 6.0, the reference formula.
// T = float:  e = (sum + rhs) * (1/6), the FP32 correction smoother of the
//             mixed-precision solve (rhs = FP32 residual).
#include "af_rbgs_tb.h"

#pragma clang fp contract(off)

namespace af {
namespace {

constexpr int LANES = 64;
constexpr int WAVES = 16;
constexpr int TX = 2 * LANES;  // haloed tile width (x)

template <typename T>
__device__ __forceinline__ T relax(T sum, T rhs, T h2);

template <>
__device__ __forceinline__ double relax<double>(double sum, double rhs, double h2)
{
    return (sum - h2 * rhs) / 6.0;
}

template <>
__device__ __forceinline__ float relax<float>(float sum, float rhs, float)
{
    return (sum + rhs) * (1.0f / 6.0f);
}

template <typename T, int K, int R, bool ZERO>
__global__ __launch_bounds__(LANES *WAVES) void k_rbgs_tb(TbArgs<T> a)
{
    constexpr int TY = WAVES * R;
    // Planes are loaded PD steps ahead of their first use so that the loads
    // of plane t+PD are in flight while step t computes.
    constexpr int PD = 1;
    constexpr int W = K + 2 + PD;  // P window (planes t-K-1 .. t+PD)
    constexpr int WD = W;          // rhs window (same modulus)
    extern __shared__ unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);  // [K][TY][LANES]
    if (a.done && *a.done) return;  // uniform across the workgroup

    const int lane = threadIdx.x;
    const int wave = threadIdx.y;
    const int g = a.g, NX = a.NX, NY = a.NY, NZ = a.NZ;
    const int sy = a.sy, sz = a.sz;

    const int xs = g + (int)blockIdx.x * (TX - 2 * K) - K;  // first haloed x
    const int ys = g + (int)blockIdx.y * (TY - 2 * K) - K;  // first haloed y
    const int zc0 = g + (int)blockIdx.z * a.cz;
    const int zc1 = min(zc0 + a.cz, g + NZ);
    const int zl0 = max(g, zc0 - K);
    const int zl1 = min(g + NZ, zc1 + K);
    const int xe = xs + 2 * lane;

    // Per-thread columns: rows y_r = ys + wave + WAVES*r, cells xe, xe+1.
    bool ld[R][2];   // cell inside the interior (loaded, computed)
    bool st[R][2];   // cell in the stored core
    int off[R];      // y*sy of each row
    int yy[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = ys + wave + WAVES * r;
        yy[r] = y;
        off[r] = y * sy;
        const bool yin = y >= g && y < g + NY;
        const bool ycore = y >= ys + K && y < ys + TY - K;
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int x = xe + s;
            const bool xin = x >= g && x < g + NX;
            ld[r][s] = xin && yin;
            st[r][s] = xin && yin && ycore && x >= xs + K && x < xs + TX - K;
        }
    }

    T pw[W][R][2];
    T dw[WD][R][2];
#pragma unroll
    for (int k = 0; k < W; k++)
#pragma unroll
        for (int r = 0; r < R; r++) pw[k][r][0] = pw[k][r][1] = T(0);
#pragma unroll
    for (int k = 0; k < WD; k++)
#pragma unroll
        for (int r = 0; r < R; r++) dw[k][r][0] = dw[k][r][1] = T(0);

    const T h2 = (T)a.h2;
    const int tend = zl1 + K;
    // Prologue: planes zl0 .. zl0+PD-1 into slots 0 .. PD-1.
#pragma unroll
    for (int p = 0; p < PD; p++) {
        const int t = zl0 + p;
        if (t < zl1) {
            const int zo = t * sz;
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int s = 0; s < 2; s++)
                    if (ld[r][s]) {
                        pw[p][r][s] = ZERO ? T(0) : a.src[xe + s + off[r] + zo];
                        dw[p][r][s] = a.rhs[xe + s + off[r] + zo];
                    }
        }
    }
    for (int t0 = zl0; t0 < tend; t0 += W) {
        // Unrolled by W so every window slot is a compile-time constant.
#pragma unroll
        for (int u = 0; u < W; u++) {
            const int t = t0 + u;
            if (t >= tend) break;  // uniform across the workgroup
            // ---- prefetch plane t+PD (slot of plane t-K-2, no longer read) ----
            if (t + PD < zl1) {
                const int zo = (t + PD) * sz;
                const int sl = (u + PD) % W;
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int s = 0; s < 2; s++) {
                        const int i = xe + s + off[r] + zo;
                        if (ld[r][s]) {
                            pw[sl][r][s] = ZERO ? T(0) : a.src[i];
                            dw[sl][r][s] = a.rhs[i];
                        }
                    }
            }
            // ---- passes j = 1..K on planes t-j ----
#pragma unroll
            for (int j = 1; j <= K; j++) {
                const int q = t - j;
                const int lo = max(g, zc0 - (K - j));
                const int hi = min(g + NZ, zc1 + (K - j));
                if (q < lo || q >= hi) continue;  // uniform
                const int col = (a.color0 + j - 1) & 1;
                const int cq = (u - j + 8 * W) % W;
                const int cm = (u - j - 1 + 8 * W) % W;
                const int cp = (u - j + 1 + 8 * W) % W;
                const int dq = cq;
                T *buf = lds + (j - 1) * TY * LANES;
                // publish this pair's other-colour cell of plane q
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int upd = ((xe + yy[r] + q) & 1) == col ? 0 : 1;
                    buf[(wave + WAVES * r) * LANES + lane] = upd == 0 ? pw[cq][r][1] : pw[cq][r][0];
                }
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int row = wave + WAVES * r;
                    const int y = yy[r];
                    const int upd = ((xe + yy[r] + q) & 1) == col ? 0 : 1;
                    const int x = xe + upd;
                    const T c = upd == 0 ? pw[cq][r][0] : pw[cq][r][1];
                    const T other = upd == 0 ? pw[cq][r][1] : pw[cq][r][0];
                    const T lnb = buf[row * LANES + max(lane - 1, 0)];
                    const T rnb = buf[row * LANES + min(lane + 1, LANES - 1)];
                    const T dnb = buf[max(row - 1, 0) * LANES + lane];
                    const T unb = buf[min(row + 1, TY - 1) * LANES + lane];
                    const T xm = (x == g) ? c : (upd == 1 ? other : lnb);
                    const T xp = (x == g + NX - 1) ? c : (upd == 0 ? other : rnb);
                    const T ym = (y == g) ? c : dnb;
                    const T yp = (y == g + NY - 1) ? c : unb;
                    const T zm = (q == g) ? c : (upd == 0 ? pw[cm][r][0] : pw[cm][r][1]);
                    const T zp = (q == g + NZ - 1) ? c : (upd == 0 ? pw[cp][r][0] : pw[cp][r][1]);
                    const T rhs = upd == 0 ? dw[dq][r][0] : dw[dq][r][1];
                    const T sum = xm + xp + ym + yp + zm + zp;
                    const T nv = relax<T>(sum, rhs, h2);
                    if (upd == 0)
                        pw[cq][r][0] = nv;
                    else
                        pw[cq][r][1] = nv;
                }
                // ---- after the last pass, plane q is final: store the core ----
                if (j == K && q >= zc0 && q < zc1) {
                    const int zo = q * sz;
#pragma unroll
                    for (int r = 0; r < R; r++)
#pragma unroll
                        for (int s = 0; s < 2; s++)
                            if (st[r][s]) a.dst[xe + s + off[r] + zo] = pw[cq][r][s];
                }
            }
            // every read of this step's LDS buffers precedes the next step's writes
            __syncthreads();
        }
    }
}

template <typename T, int K, int R, bool ZERO>
constexpr size_t lds_bytes()
{
    return (size_t)K * WAVES * R * LANES * sizeof(T);
}

template <typename T, int K, int R, bool ZERO>
hipError_t launch_k(const TbArgs<T> &a, hipStream_t s)
{
    constexpr int TY = WAVES * R;
    const dim3 grid((a.NX + TX - 2 * K - 1) / (TX - 2 * K), (a.NY + TY - 2 * K - 1) / (TY - 2 * K),
                    (a.NZ + a.cz - 1) / a.cz);
    constexpr size_t lds = lds_bytes<T, K, R, ZERO>();
    hipLaunchKernelGGL((k_rbgs_tb<T, K, R, ZERO>), grid, dim3(LANES, WAVES), lds, s, a);
    return hipGetLastError();
}

// Kernels whose LDS exceeds the default 64 KiB limit must opt in once.
template <typename T, int K, int R, bool ZERO>
hipError_t allow_lds()
{
    return hipFuncSetAttribute((const void *)k_rbgs_tb<T, K, R, ZERO>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_bytes<T, K, R, ZERO>());
}

template <typename T, int R>
hipError_t allow_all()
{
    hipError_t e;
    if ((e = allow_lds<T, 2, R, true>()) != hipSuccess) return e;
    if ((e = allow_lds<T, 2, R, false>()) != hipSuccess) return e;
    if ((e = allow_lds<T, 4, R, true>()) != hipSuccess) return e;
    if ((e = allow_lds<T, 4, R, false>()) != hipSuccess) return e;
    if ((e = allow_lds<T, 6, R, true>()) != hipSuccess) return e;
    return allow_lds<T, 6, R, false>();
}

template <typename T, int R>
hipError_t dispatch_k(const TbArgs<T> &a, int k, bool zero, hipStream_t s)
{
    switch (k) {
    case 2: return zero ? launch_k<T, 2, R, true>(a, s) : launch_k<T, 2, R, false>(a, s);
    case 4: return zero ? launch_k<T, 4, R, true>(a, s) : launch_k<T, 4, R, false>(a, s);
    case 6: return zero ? launch_k<T, 6, R, true>(a, s) : launch_k<T, 6, R, false>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

template <typename T>
hipError_t launch_rbgs_tb(const TbArgs<T> &a, int k, int rows, bool zero_src, hipStream_t s)
{
    if (rows == 1) return dispatch_k<T, 1>(a, k, zero_src, s);
    if (rows == 2) return dispatch_k<T, 2>(a, k, zero_src, s);
    return hipErrorInvalidValue;
}

template hipError_t launch_rbgs_tb<double>(const TbArgs<double> &, int, int, bool, hipStream_t);

hipError_t rbgs_tb_init()
{
    hipError_t e;
    if ((e = allow_all<double, 1>()) != hipSuccess) return e;
    return allow_all<double, 2>();
}

bool rbgs_tb_supported(int k, int rows)
{
    return (k == 2 || k == 4 || k == 6) && (rows == 1 || rows == 2);
}

}  // namespace af
