// af_noalloc.hip -- MaxDivergence over caller-owned device fields that
// allocates nothing (af_max_divergence_noalloc, include/aeroflux.h).
//
// The engine's af_max_divergence keeps its reduction partials in memory it
// allocates at New.  Callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc) get this
// form instead: the same two kernels (k_maxdiv_partial, then the fixed-tree
// k_reduce_final), with the partials in a file-scope device array of fixed
// size.  Each GPU holds its own copy of the array for the life of the
// process; its device address is looked up once per device with
// hipGetSymbolAddress and only that address is handed to kernels.  Calls
// that overlap on one GPU share the array and race; every partial is written
// by the first kernel before the second reads it, so nothing needs clearing
// between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/aeroflux.h"
#include "af_host.h"
#include "af_internal.h"

namespace {

// 8192 partials (64 KiB): one per 64 x 4 column of (x, y), so any grid with
// ceil(NX/64) * ceil(NY/4) <= 8192, e.g. every grid up to 1024 x 2048 in
// (x, y) (NZ is free).
constexpr int kScratchDoubles = 8192;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's `scratch`, looked up on the first call on
// the device (with that device current) and kept.
int scratch_for_current_device(double **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "no HIP device");
    if (dev < 0 || dev >= kMaxDevices)
        return af::fail(AF_ERR_UNSUPPORTED, "aeroflux: device %d beyond the %d this library serves",
                        dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "look up reduction scratch");
        if (!p) return af::fail(AF_ERR_HIP, "aeroflux: reduction scratch has no device address");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return AF_OK;
}

af::DField view(uint64_t addr, int64_t nx, int64_t ny, int64_t nz)
{
    af::DField d;
    d.d = reinterpret_cast<double *>(static_cast<uintptr_t>(addr));
    d.nx = (int)nx;
    d.ny = (int)ny;
    d.nz = (int)nz;
    d.sy = (int)nx;
    d.sz = (int)(nx * ny);
    return d;
}

}  // namespace

extern "C" int af_max_divergence_noalloc(uint64_t u, uint64_t v, uint64_t w, int64_t nx,
                                         int64_t ny, int64_t nz, int64_t ghost, double h,
                                         uint64_t out, uint64_t stream)
{
    if (!u || !v || !w || !out) return af::fail(AF_ERR_ARG, "aeroflux: null device address");
    if (nx < 1 || ny < 1 || nz < 1 || ghost < 1 || nx > (1 << 30) || ny > (1 << 30) ||
        nz > (1 << 30) || ghost > (1 << 20))
        return af::fail(AF_ERR_ARG, "aeroflux: bad grid %ldx%ldx%ld ghost %ld", (long)nx,
                        (long)ny, (long)nz, (long)ghost);
    if (!(h > 0) || std::isinf(h))
        return af::fail(AF_ERR_ARG, "aeroflux: cell size must be finite and positive, got %g", h);
    const int64_t cx = nx + 2 * ghost, cy = ny + 2 * ghost, cz = nz + 2 * ghost;
    const int64_t cu = (cx + 1) * cy * cz, cv = cx * (cy + 1) * cz, cw = cx * cy * (cz + 1);
    if (cu >= ((int64_t)1 << 31) || cv >= ((int64_t)1 << 31) || cw >= ((int64_t)1 << 31))
        return af::fail(AF_ERR_ARG, "aeroflux: a %ldx%ldx%ld grid exceeds 2^31 elements per field",
                        (long)nx, (long)ny, (long)nz);
    af::Geom G;
    G.NX = (int)nx;
    G.NY = (int)ny;
    G.NZ = (int)nz;
    G.g = (int)ghost;
    const int need = af::max_divergence_partials(G);
    if (need > kScratchDoubles)
        return af::fail(AF_ERR_ARG,
                        "aeroflux: a %ldx%ld (x, y) grid needs %d reduction partials, more than "
                        "the %d this allocation-free form holds; use af_max_divergence",
                        (long)nx, (long)ny, need, kScratchDoubles);
    double *partials = nullptr;
    if (int rc = scratch_for_current_device(&partials)) return rc;
    // Field shapes of NewMAC (grid.go:116-127): U has one more x-plane, V
    // one more y-row, W one more z-plane than the cell-centred fields.
    const af::DField U = view(u, cx + 1, cy, cz);
    const af::DField V = view(v, cx, cy + 1, cz);
    const af::DField W = view(w, cx, cy, cz + 1);
    const double invH = 1.0 / h;  // project.go:216
    hipStream_t s = reinterpret_cast<hipStream_t>(static_cast<uintptr_t>(stream));
    HIP_TRY(af::launch_max_divergence(U, V, W, G, invH, partials,
                                      reinterpret_cast<double *>(static_cast<uintptr_t>(out)), s),
            "max divergence (no allocation)");
    return AF_OK;
}
