// af_snapshot.hip -- arena snapshot and restore.  The reference has no
// on-disk or in-memory snapshot of an engine; the fields it would hold are
// the arena that MAC.Data exposes (grid.go:101-149), copied whole.
//
// Checksum: the sum, modulo 2^64, of mix(w_i XOR (i * 0x9E3779B97F4A7C15))
// over the payload's 64-bit words w_i (i counted from 0 across all eight
// fields), where mix is the SplitMix64 output function (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014).
// A sum is independent of the order the words are visited in, so the
// device computes it with any grid and integer atomics, and a host can
// recompute it word by word.
#include "af_snapshot.h"

#include <cstring>

#include "af_host.h"
#include "af_internal.h"

namespace af {
namespace {

constexpr char kMagic[8] = {'A', 'F', 'S', 'N', 'A', 'P', '\0', '\1'};
constexpr uint32_t kVersion = 1;
static_assert(sizeof(af_snapshot_header) == 128, "snapshot header is 128 bytes");

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_checksum(const unsigned long long *w, int64_t n,
                                                  int64_t first, unsigned long long *acc)
{
    __shared__ unsigned long long sh[256];
    unsigned long long s = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
        s += mix64(w[i] ^ ((unsigned long long)(first + i) * 0x9E3779B97F4A7C15ull));
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc, sh[0]);
}

__global__ void k_clear_u64(unsigned long long *p) { *p = 0; }

// Checksum of the eight fields as they lie in the arena.
int arena_checksum(const af_field *f, unsigned char *base, unsigned long long *acc,
                   hipStream_t s, uint64_t *out)
{
    hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(1), 0, s, acc);
    int64_t first = 0;
    for (int id = 0; id < AF_FIELD_COUNT; id++) {
        const int64_t n = f[id].Count;
        if (n > 0) {
            const int64_t blocks = (n + 255) / 256;
            const unsigned grid = (unsigned)(blocks < 4096 ? blocks : 4096);
            hipLaunchKernelGGL(k_checksum, dim3(grid), dim3(256), 0, s,
                               reinterpret_cast<const unsigned long long *>(base + f[id].Off), n,
                               first, acc);
        }
        first += n;
    }
    HIP_TRY(hipGetLastError(), "snapshot checksum");
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, acc, sizeof(v), hipMemcpyDeviceToHost, s), "snapshot checksum");
    HIP_TRY(hipStreamSynchronize(s), "snapshot checksum");
    *out = (uint64_t)v;
    return AF_OK;
}

int64_t payload_bytes(const af_field *f)
{
    int64_t n = 0;
    for (int id = 0; id < AF_FIELD_COUNT; id++) n += f[id].Count;
    return n * 8;
}

}  // namespace

int64_t snapshot_bytes(const af_field *f)
{
    return (int64_t)sizeof(af_snapshot_header) + payload_bytes(f);
}

int snapshot_save(const af_field *f, unsigned char *base, const af_config &c, void *dst,
                  int64_t bytes, unsigned long long *acc, hipStream_t s)
{
    if (!dst) return fail(AF_ERR_ARG, "aeroflux: null snapshot buffer");
    const int64_t need = snapshot_bytes(f);
    if (bytes < need)
        return fail(AF_ERR_ARG, "aeroflux: snapshot needs %ld bytes, got %ld", (long)need,
                    (long)bytes);
    af_snapshot_header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.Magic, kMagic, sizeof(kMagic));
    h.Version = kVersion;
    h.HeaderBytes = (uint32_t)sizeof(af_snapshot_header);
    h.NX = c.NX; h.NY = c.NY; h.NZ = c.NZ; h.Ghost = c.Ghost;
    h.H = c.H; h.DT = c.DT; h.Viscosity = c.Viscosity; h.Density = c.Density;
    h.PayloadBytes = payload_bytes(f);
    if (int rc = arena_checksum(f, base, acc, s, &h.Checksum)) return rc;
    unsigned char *out = static_cast<unsigned char *>(dst);
    int64_t off = (int64_t)sizeof(h);
    for (int id = 0; id < AF_FIELD_COUNT; id++) {
        HIP_TRY(hipMemcpyAsync(out + off, base + f[id].Off, (size_t)f[id].Count * 8,
                               hipMemcpyDefault, s),
                "snapshot copy");
        off += f[id].Count * 8;
    }
    HIP_TRY(hipMemcpyAsync(out, &h, sizeof(h), hipMemcpyDefault, s), "snapshot header");
    HIP_TRY(hipStreamSynchronize(s), "snapshot");
    return AF_OK;
}

int snapshot_load(const af_field *f, unsigned char *base, const af_config &c, const void *src,
                  int64_t bytes, unsigned long long *acc, hipStream_t s)
{
    if (!src) return fail(AF_ERR_ARG, "aeroflux: null snapshot buffer");
    if (bytes < (int64_t)sizeof(af_snapshot_header))
        return fail(AF_ERR_ARG, "aeroflux: snapshot of %ld bytes is shorter than its header",
                    (long)bytes);
    af_snapshot_header h;
    HIP_TRY(hipMemcpyAsync(&h, src, sizeof(h), hipMemcpyDefault, s), "snapshot header");
    HIP_TRY(hipStreamSynchronize(s), "snapshot header");
    if (std::memcmp(h.Magic, kMagic, sizeof(kMagic)) != 0)
        return fail(AF_ERR_ARG, "aeroflux: not an aeroflux snapshot");
    if (h.Version != kVersion || h.HeaderBytes != sizeof(af_snapshot_header))
        return fail(AF_ERR_ARG, "aeroflux: snapshot version %u (header %u bytes) is not %u",
                    h.Version, h.HeaderBytes, kVersion);
    if (h.NX != c.NX || h.NY != c.NY || h.NZ != c.NZ || h.Ghost != c.Ghost)
        return fail(AF_ERR_ARG,
                    "aeroflux: snapshot grid %ldx%ldx%ld ghost %ld does not match "
                    "%ldx%ldx%ld ghost %ld",
                    (long)h.NX, (long)h.NY, (long)h.NZ, (long)h.Ghost, (long)c.NX, (long)c.NY,
                    (long)c.NZ, (long)c.Ghost);
    const int64_t payload = payload_bytes(f);
    if (h.PayloadBytes != payload)
        return fail(AF_ERR_ARG, "aeroflux: snapshot payload is %ld bytes, the arena's %ld",
                    (long)h.PayloadBytes, (long)payload);
    if (bytes < (int64_t)sizeof(h) + payload)
        return fail(AF_ERR_ARG, "aeroflux: snapshot truncated: %ld of %ld bytes", (long)bytes,
                    (long)(sizeof(h) + payload));
    const unsigned char *in = static_cast<const unsigned char *>(src);
    int64_t off = (int64_t)sizeof(h);
    for (int id = 0; id < AF_FIELD_COUNT; id++) {
        HIP_TRY(hipMemcpyAsync(base + f[id].Off, in + off, (size_t)f[id].Count * 8,
                               hipMemcpyDefault, s),
                "restore copy");
        off += f[id].Count * 8;
    }
    uint64_t sum = 0;
    if (int rc = arena_checksum(f, base, acc, s, &sum)) return rc;
    if (sum != h.Checksum) {
        for (int id = 0; id < AF_FIELD_COUNT; id++)
            HIP_TRY(launch_zero(reinterpret_cast<double *>(base + f[id].Off), f[id].Count, s),
                    "zero after bad snapshot");
        HIP_TRY(hipStreamSynchronize(s), "zero after bad snapshot");
        return fail(AF_ERR_ARG,
                    "aeroflux: snapshot checksum mismatch (%016llx, header %016llx); the "
                    "fields were zeroed",
                    (unsigned long long)sum, (unsigned long long)h.Checksum);
    }
    return AF_OK;
}

}  // namespace af
