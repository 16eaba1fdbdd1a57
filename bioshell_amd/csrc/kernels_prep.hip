// kernels_prep.hip -- single-ensemble launches of the preparation steps
// (bodies and the data flow in prep_device.hpp) and the single-pair
// superposition kernel.
#include "prep_device.hpp"

namespace bs {
namespace {

// Reference: the fits of the S samples onto conformation 0 and their
// rotated sums, one slice of samples per work-group (fit0_slice),
// the mean (thread per coordinate, one wave per work-group to spread the few
// coordinates over more CUs), then R0 (one block).
__global__ __launch_bounds__(256) void fit0_kernel(const double* __restrict__ xyz, int n, int S,
                                                   int L, double* __restrict__ partial)
{
    __shared__ Fit0Lds lds;
    fit0_slice(xyz, n, S, L, blockIdx.x, partial, lds);
}

__global__ __launch_bounds__(64) void mean_final_kernel(const double* __restrict__ partial,
                                                         int n_slices, int S, int L,
                                                         double* __restrict__ ref)
{
    const int v = blockIdx.x * 64 + threadIdx.x;
    if (v >= 3 * L) return;
    mean_final_coord(partial, n_slices, S, L, v, ref);
}

__global__ __launch_bounds__(256) void ref_gram_kernel(double* __restrict__ ref, int L)
{
    __shared__ double part[4][6];
    ref_gram_block(ref, L, part);
}

template <int MODE>
__global__ __launch_bounds__(256) void fit_pack_kernel(const double* __restrict__ xyz,
                                                       const double* __restrict__ ref, int n,
                                                       int n_pad, int L, int n_kblk,
                                                       double* __restrict__ rot,
                                                       double* __restrict__ cent,
                                                       uint4* __restrict__ panels,
                                                       double* __restrict__ consts, int ld)
{
    __shared__ FitPackLds lds;
    const int blk = xcd_contiguous_block(fit_groups(n_pad));
    if (blk < 0) return;
    fit_pack_block<MODE>(xyz, ref, n, L, n_kblk, blk * kFitConfs, rot, cent, panels, consts, ld, lds);
}

// ------------------------------------------------------ single superposition
__global__ __launch_bounds__(64) void superpose_kernel(const double* __restrict__ a,
                                                       const double* __restrict__ b, int L,
                                                       double* __restrict__ out)
{
    const int lane = threadIdx.x;
    double sa[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
    for (int k = lane; k < L; k += 64)
        for (int c = 0; c < 3; ++c) {
            sa[c] += a[3 * k + c];
            sb[c] += b[3 * k + c];
        }
    for (int c = 0; c < 3; ++c) {
        sa[c] = wave_sum(sa[c]) / L;
        sb[c] = wave_sum(sb[c]) / L;
    }
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double ga = 0, gb = 0;
    for (int k = lane; k < L; k += 64) {
        double pa[3], pb[3];
        for (int c = 0; c < 3; ++c) {
            pa[c] = a[3 * k + c] - sa[c];
            pb[c] = b[3 * k + c] - sb[c];
        }
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) m[3 * p + q] += pb[p] * pa[q];
        ga += pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2];
        gb += pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2];
    }
    for (int q = 0; q < 9; ++q) m[q] = wave_sum(m[q]);
    ga = wave_sum(ga);
    gb = wave_sum(gb);
    if (lane == 0) {
        double R[9], lam;
        rotation_from_cov(m, R, &lam);
        for (int q = 0; q < 9; ++q) out[q] = R[q];
        for (int r = 0; r < 3; ++r)
            out[9 + r] = sa[r] - (R[3 * r] * sb[0] + R[3 * r + 1] * sb[1] + R[3 * r + 2] * sb[2]);
        const double r2 = 2.0 * (0.5 * (ga + gb) - lam) / L;
        out[12] = sqrt(r2 > 0.0 ? r2 : 0.0);
    }
}

}  // namespace

void launch_reference(const double* xyz, int n, int L, double* partial, double* ref, hipStream_t s)
{
    const int S = ref_sample_count(n);
    const int slices = ref_slices(S);
    fit0_kernel<<<slices, 256, 0, s>>>(xyz, n, S, L, partial);
    mean_final_kernel<<<(3 * L + 63) / 64, 64, 0, s>>>(partial, slices, S, L, ref);
    ref_gram_kernel<<<1, 256, 0, s>>>(ref, L);
}

void launch_fit_pack(const double* xyz, const double* ref, int n, int n_pad, int L, int n_kblk,
                     GramMode mode, double* rot, double* cent, void* panels, double* consts, int ld,
                     hipStream_t s)
{
    const int blocks = xcd_contiguous_grid(fit_groups(n_pad));
    uint4* p = static_cast<uint4*>(panels);
    if (mode == kGramBF16x3)
        fit_pack_kernel<kGramBF16x3><<<blocks, 256, 0, s>>>(xyz, ref, n, n_pad, L, n_kblk, rot,
                                                            cent, p, consts, ld);
    else
        fit_pack_kernel<kGramFP32><<<blocks, 256, 0, s>>>(xyz, ref, n, n_pad, L, n_kblk, rot, cent,
                                                          p, consts, ld);
}

void launch_superpose(const double* a, const double* b, int L, double* out13, hipStream_t s)
{
    superpose_kernel<<<1, 64, 0, s>>>(a, b, L, out13);
}

}  // namespace bs
