// af_dct.h -- FFT-based orthonormal DCT-II passes of the direct pressure
// solve (af_dct.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "af_internal.h"

namespace af {

// One transform pass over lines of length n: element e of line (a, b) is at
// e * sn + a * sa + b * sb (in and out may differ), a < A, b < B.
struct DctPass {
    const double *in = nullptr;
    double *out = nullptr;
    int64_t in_sn = 0, in_sa = 0, in_sb = 0;
    int64_t out_sn = 0, out_sa = 0, out_sb = 0;
    int A = 0, B = 0;
    double sign = 1.0;              // multiplies every loaded value
    const double *tw = nullptr;     // e^{-2 pi i t / n}, t < n, as (re, im) pairs
    const double *pw = nullptr;     // s_k e^{-i pi k / 2n}, (re, im) pairs
    const double *wc = nullptr;     // e^{i pi k / 2n}, (re, im) pairs
    const double *iq = nullptr;     // inverse scale q_k
    double *partials = nullptr;     // one per workgroup (modes 1, 2)
    // Mode 2 (solve along the last axis; a = x index, b = y index):
    const double *lx = nullptr, *ly = nullptr, *lz = nullptr;     // eigenvalues / H^2
    const double *c0x = nullptr, *c0y = nullptr, *c0z = nullptr;  // C(k, 0)
    const double *sum = nullptr;    // S, the rhs sum
    const double *shift = nullptr;  // mode 4: added to every output
    // Mode 5 (x forward from the velocity): computeDivergence's fields and
    // scale; the divergence is stored to D and -D is transformed.
    DField U{}, V{}, W{}, D{};
    Geom G{};
    double dscale = 0;
};

bool dct_fft_supported(int n);        // n a power of two in 8 .. 512
int dct_blocks(int n, int A, int B);  // workgroups of one pass

// Host tables for length n: tw, pw, wc as interleaved (re, im) pairs.
void dct_tables(int n, std::vector<double> *tw, std::vector<double> *pw, std::vector<double> *wc,
                std::vector<double> *iq);

// mode: 0 forward, 1 forward + block sums of the loaded values, 2 forward +
// spectral division + inverse (with block partials of q_0), 3 inverse,
// 4 inverse + *shift, 5 as 1 along x with the values computed as the
// divergence of U, V, W (stored to D on the way).
hipError_t launch_dct_pass(int mode, int n, const DctPass &p, hipStream_t s);

// One workgroup: mode 0 *out = sum of partials[0..n); mode 1
// *out = h2 * (*sum) - sum of partials.  Fixed order.
hipError_t launch_reduce_partials(const double *partials, int n, const double *sum, double h2,
                                  int mode, double *out, hipStream_t s);

}  // namespace af
