// cluster.cpp -- entry points for the RMSD-matrix reductions BioShell's
// clustering tools consume: row minima / maxima and the agglomerative
// clustering loop (kernels_cluster.hip).
#include <algorithm>

#include "engine.hpp"

using namespace bs;

namespace {

void check_n(bs_context* ctx, size_t n)
{
    BS_REQUIRE(ctx != nullptr, "context is null");
    BS_REQUIRE(n >= 1 && n <= (1u << 21), "n must be in [1, 2^21] (got %zu)", n);
}

// Device copy of a host packed triangle in the context's workspace.
const float* upload_upper(bs_context* ctx, const float* upper, size_t n)
{
    const size_t bytes = std::max<size_t>(n * (n - 1) / 2 * 4, 4);
    float* d = static_cast<float*>(ctx->ws_out.ensure(bytes, "rmsd matrix"));
    if (n >= 2)
        BS_HIP(hipMemcpyAsync(d, upper, n * (n - 1) / 2 * 4, hipMemcpyHostToDevice, ctx->stream));
    return d;
}

void cluster_device(bs_context* ctx, const float* d_upper, size_t n, int rule, uint32_t* merge_i,
                    uint32_t* merge_j, float* merge_d)
{
    if (n < 2) return;
    hipStream_t s = ctx->stream;
    const int N = (int)n;
    const int n_part = std::max(1, std::min(4 * ctx->n_cu, (N + 3) / 4));
    DevBuf D, aux;
    D.bind(&ctx->budget);
    aux.bind(&ctx->budget);
    const size_t ld = (n + 63) / 64 * 64;  // 256-byte rows: the scan reads 16-byte groups
    float* dm = static_cast<float*>(D.ensure(n * ld * 4, "clustering matrix"));
    const size_t cand = cluster_candidate_bytes();
    char* a = static_cast<char*>(aux.ensure(n * 4 * 4 + 256 + 256 + (size_t)n_part * cand,
                                            "clustering state"));
    unsigned* sizes = reinterpret_cast<unsigned*>(a);
    unsigned* mi = reinterpret_cast<unsigned*>(a + n * 4);
    unsigned* mj = reinterpret_cast<unsigned*>(a + n * 8);
    float* md = reinterpret_cast<float*>(a + n * 12);
    void* state = a + (n * 16 + 255) / 256 * 256;
    void* part = static_cast<char*>(state) + 256;
    BS_REQUIRE(cluster_state_bytes() <= 256, "merge state too large");
    launch_cluster_init(d_upper, N, (int)ld, dm, sizes, state, s);
    BS_HIP(hipGetLastError());
    // the steps are identical launches: capture a run of them once and replay
    // it (host launch cost, not the GPU, bounds a launch-per-step loop)
    const int per_graph = std::min(128, N - 1);
    struct ExecGuard {  // destroyed once the stream has drained
        hipGraphExec_t exec = nullptr;
        hipStream_t s = nullptr;
        ~ExecGuard()
        {
            if (exec) {
                (void)hipStreamSynchronize(s);
                (void)hipGraphExecDestroy(exec);
            }
        }
    } guard_exec;
    guard_exec.s = s;
    hipGraph_t graph = nullptr;
    BS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        for (int k = 0; k < per_graph; ++k)
            launch_cluster_step(dm, N, (int)ld, state, part, n_part, sizes, rule, mi, mj, md, s);
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    BS_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t ie = hipGraphInstantiate(&guard_exec.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    BS_HIP(ie);
    for (int done = 0; done < N - 1; done += per_graph) BS_HIP(hipGraphLaunch(guard_exec.exec, s));
    BS_HIP(hipMemcpyAsync(merge_i, mi, (n - 1) * 4, hipMemcpyDeviceToHost, s));
    BS_HIP(hipMemcpyAsync(merge_j, mj, (n - 1) * 4, hipMemcpyDeviceToHost, s));
    BS_HIP(hipMemcpyAsync(merge_d, md, (n - 1) * 4, hipMemcpyDeviceToHost, s));
    BS_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

bs_status bs_rmsd_row_stats_device(bs_context* ctx, const float* d_upper, size_t n, float* d_row_min,
                                   float* d_row_max)
{
    return guard([&] {
        check_n(ctx, n);
        BS_REQUIRE((n < 2 || d_upper) && d_row_min && d_row_max, "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        void* scratch = ctx->ws_stats.ensure(row_stats_scratch_bytes((int)n), "row statistics tiles");
        launch_row_stats(d_upper, (int)n, d_row_min, d_row_max, scratch, ctx->stream);
        BS_HIP(hipGetLastError());
    });
}

bs_status bs_rmsd_row_stats(bs_context* ctx, const float* upper, size_t n, float* row_min,
                            float* row_max)
{
    return guard([&] {
        ctx = resolve_context(ctx);
        check_n(ctx, n);
        BS_REQUIRE((n < 2 || upper) && row_min && row_max, "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        if (n < 2) {  // no other element: the reference's initial values
            row_min[0] = 3.402823466e+38f;
            row_max[0] = -3.402823466e+38f;
            return;
        }
        const float* d = upload_upper(ctx, upper, n);
        float* st = static_cast<float*>(ctx->ws_misc.ensure(n * 8, "row statistics"));
        void* scratch = ctx->ws_stats.ensure(row_stats_scratch_bytes((int)n), "row statistics tiles");
        launch_row_stats(d, (int)n, st, st + n, scratch, ctx->stream);
        BS_HIP(hipGetLastError());
        BS_HIP(hipMemcpyAsync(row_min, st, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        BS_HIP(hipMemcpyAsync(row_max, st + n, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bs_status bs_rmsd_members_row_max(bs_context* ctx, const float* upper, size_t n,
                                  const uint32_t* members, size_t m, float* row_max)
{
    return guard([&] {
        ctx = resolve_context(ctx);
        check_n(ctx, n);
        BS_REQUIRE(members && row_max && (n < 2 || upper), "null pointer argument");
        BS_REQUIRE(m >= 1 && m <= n, "m must be in [1, n]");
        for (size_t k = 0; k < m; ++k) BS_REQUIRE(members[k] < n, "member %u out of range", members[k]);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        if (m == 1) {
            row_max[0] = -3.402823466e+38f;
            return;
        }
        const float* d = upload_upper(ctx, upper, n);
        char* w = static_cast<char*>(ctx->ws_misc.ensure(m * 8, "members"));
        BS_HIP(hipMemcpyAsync(w, members, m * 4, hipMemcpyHostToDevice, ctx->stream));
        float* out = reinterpret_cast<float*>(w + m * 4);
        launch_members_row_max(d, (int)n, reinterpret_cast<const unsigned*>(w), (int)m, out, ctx->stream);
        BS_HIP(hipGetLastError());
        BS_HIP(hipMemcpyAsync(row_max, out, m * 4, hipMemcpyDeviceToHost, ctx->stream));
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bs_status bs_hierarchical_clustering_device(bs_context* ctx, const float* d_upper, size_t n,
                                            bs_linkage linkage, uint32_t* merge_i,
                                            uint32_t* merge_j, float* merge_d)
{
    return guard([&] {
        check_n(ctx, n);
        BS_REQUIRE(linkage >= BS_SINGLE_LINK && linkage <= BS_WARDS_METHOD, "unknown linkage %d",
                   (int)linkage);
        if (n < 2) return;
        BS_REQUIRE(d_upper && merge_i && merge_j && merge_d, "null pointer argument");
        BS_REQUIRE(n <= 65536, "the dense clustering matrix is limited to 65536 elements");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        cluster_device(ctx, d_upper, n, (int)linkage, merge_i, merge_j, merge_d);
    });
}

bs_status bs_hierarchical_clustering(bs_context* ctx, const float* upper, size_t n,
                                     bs_linkage linkage, uint32_t* merge_i, uint32_t* merge_j,
                                     float* merge_d)
{
    return guard([&] {
        ctx = resolve_context(ctx);
        check_n(ctx, n);
        BS_REQUIRE(linkage >= BS_SINGLE_LINK && linkage <= BS_WARDS_METHOD, "unknown linkage %d",
                   (int)linkage);
        if (n < 2) return;
        BS_REQUIRE(upper && merge_i && merge_j && merge_d, "null pointer argument");
        BS_REQUIRE(n <= 65536, "the dense clustering matrix is limited to 65536 elements");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const float* d = upload_upper(ctx, upper, n);
        cluster_device(ctx, d, n, (int)linkage, merge_i, merge_j, merge_d);
    });
}

}  // extern "C"
