// af_xcd.h -- XCD-aware workgroup ordering for gfx950 (8 XCDs, one L2 each).
//
// The dispatcher hands workgroup b of a launch to XCD (b mod 8).  A stencil
// grid numbered x-fastest would therefore scatter neighbouring rows and
// planes across all eight L2 caches, and every XCD would re-fetch the halo
// rows and planes its blocks share.  Here the hardware id is remapped so
// that XCD k runs one contiguous range of the logical block order (x
// fastest, then y, then z): each L2 sees a z-slab marched plane by plane,
// and the rows and planes that neighbouring blocks share are reused from it.
#pragma once

#include <hip/hip_runtime.h>

namespace af {

constexpr int kXcds = 8;

// Logical block index of this workgroup for a 1-D launch of n blocks.
__device__ __forceinline__ int xcd_logical_block(int n)
{
    const int b = blockIdx.x;
    const int q = n / kXcds, r = n % kXcds;
    const int x = b % kXcds;
    const int start = x * q + (x < r ? x : r);
    return start + b / kXcds;
}

// Logical (bx, by, bz) of this workgroup for an nbx * nby * nbz grid
// launched as nbx * nby * nbz one-dimensional blocks.
__device__ __forceinline__ void xcd_block3(int nbx, int nby, int nbz, int &bx, int &by, int &bz)
{
    const int l = xcd_logical_block(nbx * nby * nbz);
    bx = l % nbx;
    const int r = l / nbx;
    by = r % nby;
    bz = r / nby;
}

// The same grid in y-slab order: slabs of `sl` block rows, each walked plane
// by plane (z), then row by row, then x.  A stencil that reads the planes
// above and below keeps them in its XCD's L2 while it works through the
// slab, even when whole planes of every field it touches would not fit;
// sl >= nby is the plane order of xcd_block3.
__device__ __forceinline__ void xcd_block3_slab(int nbx, int nby, int nbz, int sl, int &bx,
                                                int &by, int &bz)
{
    const int l = xcd_logical_block(nbx * nby * nbz);
    sl = sl < nby ? sl : nby;
    const int full = sl * nbz * nbx;  // blocks of a full slab; only the last is partial
    const int s = l / full;
    int r = l - s * full;
    const int rows = (nby - s * sl) < sl ? (nby - s * sl) : sl;
    const int per_z = rows * nbx;
    bz = r / per_z;
    r -= bz * per_z;
    by = s * sl + r / nbx;
    bx = r - (r / nbx) * nbx;
}

}  // namespace af
