// noalloc.hip -- all-vs-all cRMSD on device buffers without allocating:
// bs_crmsd_all_vs_all_device_noalloc.
//
// The existing device entry point keeps its workspace (prepared panels and
// constants, centroids, rotations, the reference and its partial sums, the
// tile order, the flag bitmap and queue) in buffers that grow on demand.
// This variant takes every one of them from `scratch`, a fixed array in the
// library's own device memory, so that callers which cannot allocate -- code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc -- can run the same kernels.  Each GPU holds its own copy of the
// array (kScratchBytes) from the moment the library's code object is loaded
// there until the process ends.  Its device address is taken once per device,
// when a context is created on it (noalloc_init), and kept; host code passes
// that address to the kernels, never the array's name.
//
// Every call on a GPU uses that GPU's one array: calls that overlap on one
// GPU (two streams, two threads with different contexts) race, and must be
// ordered by the caller.  A call rewrites everything it reads: the flag
// bitmap and its count are cleared by a fill kernel at the start, the other
// regions are written before they are read.
#include <mutex>

#include "engine.hpp"

namespace bs {
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kNoAllocListCap = 1 << 16;           // queued FP64 re-evaluations (512 KiB); more go
                                                   // through the bitmap scan
constexpr int kMaxDevices = 64;
constexpr int kTileRows = 64;
constexpr int kSuper = 8;  // tile-order super-blocks, as build_tile_order

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// The XCD-aware super-blocked tile order of build_tile_order, written on the
// device (no host staging): one work-group walks the super-blocks in order,
// 256 at a time, each thread writing one super-block's tiles after a scan of
// the tile counts.
__global__ __launch_bounds__(256) void tile_order_kernel(int tiles_r, int tiles_c, int triangle,
                                                         uint32_t* __restrict__ out)
{
    __shared__ int wave_tot[4];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int sr_n = (tiles_r + kSuper - 1) / kSuper, sc_n = (tiles_c + kSuper - 1) / kSuper;
    const int total = sr_n * sc_n;
    if (tid == 0) base = 0;
    __syncthreads();
    for (int b0 = 0; b0 < total; b0 += 256) {
        const int b = b0 + tid;
        const int sr = b / sc_n, sc = b - (b / sc_n) * sc_n;
        const int r0 = sr * kSuper, r1 = min(r0 + kSuper, tiles_r);
        const int c0 = sc * kSuper, c1 = min(c0 + kSuper, tiles_c);
        int cnt = 0;
        if (b < total && (!triangle || sc >= sr))
            for (int r = r0; r < r1; ++r) cnt += max(0, c1 - (triangle ? max(r, c0) : c0));
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wave_tot[wv] = incl;
        __syncthreads();
        int off = base + incl - cnt;
        for (int w = 0; w < wv; ++w) off += wave_tot[w];
        if (cnt > 0)
            for (int r = r0; r < r1; ++r)
                for (int c = triangle ? max(r, c0) : c0; c < c1; ++c)
                    out[off++] = ((uint32_t)r << 16) | (uint32_t)c;
        __syncthreads();
        if (tid == 0) base += wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        __syncthreads();
    }
}

}  // namespace

}  // namespace bs

// The array itself (file scope).  In host code this name is only a stand-in
// for the device symbol: its address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[bs::kScratchBytes];

namespace bs {

void noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
    g_status[device] = e;
    if (e == hipSuccess) g_scratch[device] = p;
}

// Workspace bytes of an all-vs-all of n_conf x n_atoms in the scratch layout.
size_t noalloc_scratch_bytes(size_t n_conf, size_t n_atoms)
{
    if (n_conf < 2 || n_atoms < 1 || n_atoms > (1u << 20) || n_conf > (1u << 21)) return 0;
    const int n = (int)n_conf, L = (int)n_atoms;
    const PreparedSizes z = prepared_sizes(n, L);
    const size_t tiles = (size_t)(n + kTileRows - 1) / kTileRows;
    const uint64_t n_out = (uint64_t)n * (n - 1) / 2;
    return align256(z.cent) + align256(z.rot) + align256(z.ref) + align256(z.partial) + align256(z.panels) +
           align256(z.consts) + align256(tiles * (tiles + 1) / 2 * 4) +
           align256(flag_region_bytes(n_out, kNoAllocListCap));
}

}  // namespace bs

using namespace bs;

extern "C" {

bs_status bs_crmsd_all_vs_all_device_noalloc(bs_context* ctx, const double* d_xyz, size_t n_conf,
                                             size_t n_atoms, bs_method method, void* d_out_upper,
                                             int out_f64)
{
    return guard([&] {
        BS_REQUIRE(ctx != nullptr, "context is null");
        BS_REQUIRE(n_atoms >= 1 && n_atoms <= (1u << 20), "n_atoms must be in [1, 2^20] (got %zu)", n_atoms);
        BS_REQUIRE(n_conf <= (1u << 21), "n_conf too large (%zu)", n_conf);
        BS_REQUIRE(method >= BS_FP64_PLAIN && method <= BS_BF16X3_GRAM, "unknown method %d", (int)method);
        if (n_conf < 2) return;
        BS_REQUIRE(d_xyz && d_out_upper, "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const hipStream_t s = ctx->stream;
        const int n = (int)n_conf, L = (int)n_atoms;
        if (!is_gram(method)) {  // the per-pair kernel needs no workspace
            launch_plain_avsa(d_xyz, n, L, method == BS_FP64_PLAIN, d_out_upper, out_f64 != 0, ctx->n_cu, s);
            BS_HIP(hipGetLastError());
            return;
        }
        const size_t need = noalloc_scratch_bytes(n_conf, n_atoms);
        if (need > kScratchBytes)
            fail(BS_ERR_OUT_OF_BUDGET,
                 "%zu x %zu needs %zu bytes of workspace; the no-allocation entry point has %zu "
                 "(bs_crmsd_all_vs_all_device allocates what it needs)",
                 n_conf, n_atoms, need, kScratchBytes);
        char* base = nullptr;
        {
            std::lock_guard<std::mutex> g(g_mu);
            if (ctx->device >= 0 && ctx->device < kMaxDevices) base = static_cast<char*>(g_scratch[ctx->device]);
            if (!base)
                fail(BS_ERR_HIP, "no scratch address for device %d (hipGetSymbolAddress: %s)", ctx->device,
                     hipGetErrorString(ctx->device >= 0 && ctx->device < kMaxDevices ? g_status[ctx->device]
                                                                                     : hipErrorInvalidDevice));
        }
        // carve the regions in a fixed order, each 256-byte aligned
        const PreparedSizes z = prepared_sizes(n, L);
        const int tiles = (n + kTileRows - 1) / kTileRows;
        const int n_tiles = tiles * (tiles + 1) / 2;
        const uint64_t n_out = (uint64_t)n * (n - 1) / 2;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            char* p = base + off;
            off += align256(bytes);
            return p;
        };
        PreparedPtrs p;
        p.cent = reinterpret_cast<double*>(take(z.cent));
        p.rot = reinterpret_cast<double*>(take(z.rot));
        p.ref = reinterpret_cast<double*>(take(z.ref));
        p.partial = reinterpret_cast<double*>(take(z.partial));
        p.panels = take(z.panels);
        p.consts = reinterpret_cast<double*>(take(z.consts));
        uint32_t* order = reinterpret_cast<uint32_t*>(take((size_t)n_tiles * 4));
        uint32_t* flags = reinterpret_cast<uint32_t*>(take(flag_region_bytes(n_out, kNoAllocListCap)));
        BS_REQUIRE(off == need && off <= kScratchBytes, "scratch layout mismatch");

        const PrepView v = prepare_into(p, d_xyz, n, L, mode_of(method), nullptr, s);
        tile_order_kernel<<<1, 256, 0, s>>>(tiles, tiles, 1, order);
        BS_HIP(hipGetLastError());
        GramRun run;
        run.triangle = true;
        run.d_out = d_out_upper;
        run.out_f64 = out_f64 != 0;
        // the region's bytes held other data in earlier calls: zero it first
        run_gram_with(v, v, run, order, n_tiles, flags, kNoAllocListCap, true, ctx->gsc.max_wg, s, nullptr);
    });
}

}  // extern "C"
