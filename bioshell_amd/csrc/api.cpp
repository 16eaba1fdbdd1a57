// api.cpp -- the extern "C" entry points declared in include/bioshell_gpu.h.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "engine.hpp"

namespace bs {
const char* last_error_cstr();

// The process-wide default context: device 0, no HBM budget, created by the
// first call that needs it and kept until the process ends (never destroyed,
// so no call can race its teardown; the driver frees its device memory at
// exit).  A failed creation is remembered and reported to every later call.
bs_context* resolve_context(bs_context* ctx)
{
    if (ctx) return ctx;
    static std::once_flag once;
    static bs_context* dflt = nullptr;
    static bs_status dflt_status = BS_OK;
    static std::string dflt_error;
    std::call_once(once, [] {
        dflt_status = bs_context_create(0, 0, &dflt);
        if (dflt_status != BS_OK) dflt_error = bs_last_error();
    });
    if (dflt_status != BS_OK) fail(dflt_status, "default context: %s", dflt_error.c_str());
    return dflt;
}
}  // namespace bs

using namespace bs;

namespace {

void check_ctx(const bs_context* ctx) { BS_REQUIRE(ctx != nullptr, "context is null"); }

void check_shape(size_t n_conf, size_t n_atoms)
{
    BS_REQUIRE(n_atoms >= 1, "n_atoms must be >= 1 (got %zu)", n_atoms);
    BS_REQUIRE(n_atoms <= (1u << 20), "n_atoms too large (%zu)", n_atoms);
    BS_REQUIRE(n_conf <= (1u << 21), "n_conf too large (%zu)", n_conf);
}

bool valid_method(bs_method m) { return m >= BS_FP64_PLAIN && m <= BS_BF16X3_GRAM; }

}  // namespace

namespace bs {
void context_retain(bs_context* ctx) { ctx->refs.fetch_add(1, std::memory_order_relaxed); }

void context_release(bs_context* ctx)
{
    if (ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)guard([&] {
        DeviceGuard dg(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->prep.reset();
        ctx->qws.reset();
        for (auto& r : ctx->ring) r.destroy();
        ctx->ring.clear();
        if (ctx->h_counts) (void)hipHostFree(ctx->h_counts);
        ctx->h_counts = nullptr;
        ctx->gsc.reset();
        for (auto& b : ctx->stream_blk) b.reset();
        for (int k = 0; k < 2; ++k) {
            if (ctx->h_blk[k]) (void)hipHostFree(ctx->h_blk[k]);
            if (ctx->ev_compute[k]) (void)hipEventDestroy(ctx->ev_compute[k]);
            if (ctx->ev_copy[k]) (void)hipEventDestroy(ctx->ev_copy[k]);
        }
        if (ctx->h_blk_counts) (void)hipHostFree(ctx->h_blk_counts);
        if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
        ctx->ws_xyz.reset();
        ctx->ws_out.reset();
        ctx->ws_misc.reset();
        ctx->ws_stats.reset();
        ctx->batch.release();
        (void)hipStreamDestroy(ctx->stream);
    });
    delete ctx;
}
}  // namespace bs

extern "C" {

const char* bs_last_error(void) { return bs::last_error_cstr(); }

const char* bs_version(void) { return "bioshell-mi355x 0.1.0 (gfx950)"; }

int bs_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

bs_status bs_context_create(int device, size_t hbm_budget_bytes, bs_context** out)
{
    return guard([&] {
        BS_REQUIRE(out != nullptr, "out is null");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) fail(BS_ERR_NO_DEVICE, "no HIP device visible");
        BS_REQUIRE(device >= 0 && device < n, "device %d out of range (%d visible)", device, n);
        DeviceGuard dg(device);
        hipDeviceProp_t prop;
        BS_HIP(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            fail(BS_ERR_UNSUPPORTED, "device %d is %s; this library is built for gfx950", device,
                 prop.gcnArchName);
        noalloc_init(device);  // the no-allocation entry point's scratch address on this device
        auto ctx = std::make_unique<bs_context>();
        ctx->device = device;
        ctx->n_cu = prop.multiProcessorCount;
        ctx->gsc.max_wg = 2 * ctx->n_cu;
        ctx->budget.set_cap(hbm_budget_bytes);
        BS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        BS_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        for (auto& b : ctx->stream_blk) b.bind(&ctx->budget);
        ctx->prep.bind(&ctx->budget);
        ctx->gsc.bind(&ctx->budget);
        ctx->ws_xyz.bind(&ctx->budget);
        ctx->ws_out.bind(&ctx->budget);
        ctx->ws_misc.bind(&ctx->budget);
        ctx->ws_stats.bind(&ctx->budget);
        ctx->batch.bind(&ctx->budget);
        *out = ctx.release();
    });
}

void bs_context_destroy(bs_context* ctx)
{
    if (ctx) context_release(ctx);
}

bs_status bs_context_synchronize(bs_context* ctx)
{
    return guard([&] {
        check_ctx(ctx);
        BS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

void* bs_context_stream(bs_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

bs_status bs_context_last_stats(const bs_context* ctx, bs_run_stats* out)
{
    return guard([&] {
        check_ctx(ctx);
        BS_REQUIRE(out != nullptr, "out is null");
        *out = bs_run_stats{};
        if (ctx->ring_used == 0) return;
        const size_t k = (ctx->ring_next + ctx->ring.size() - 1) % ctx->ring.size();
        *out = ctx->ring[k].read();
    });
}

bs_status bs_context_history(const bs_context* ctx, bs_run_stats* out, size_t max_calls,
                             size_t* n_out)
{
    return guard([&] {
        check_ctx(ctx);
        BS_REQUIRE(n_out != nullptr && (out != nullptr || max_calls == 0), "null pointer argument");
        const size_t n = std::min(max_calls, ctx->ring_used);
        const size_t R = ctx->ring.size();
        for (size_t q = 0; q < n; ++q) {
            const size_t k = (ctx->ring_next + R - n + q) % R;
            out[q] = ctx->ring[k].read();
        }
        *n_out = n;
    });
}

size_t bs_context_bytes_in_use(const bs_context* ctx) { return ctx ? ctx->budget.in_use() : 0; }

bs_status bs_default_context(bs_context** out)
{
    return guard([&] {
        BS_REQUIRE(out != nullptr, "out is null");
        *out = nullptr;
        *out = resolve_context(nullptr);
    });
}

bs_status bs_crmsd_pair_f64(bs_context* ctx, const double* a_xyz, const double* b_xyz,
                            size_t n_atoms, double* out)
{
    return guard([&] {
        BS_REQUIRE(a_xyz && b_xyz && out, "null pointer argument");
        check_shape(2, n_atoms);
        ctx = resolve_context(ctx);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        hipStream_t s = ctx->stream;
        double* d = static_cast<double*>(ctx->ws_xyz.ensure(n_atoms * 6 * 8, "pair coordinates"));
        double* o = static_cast<double*>(ctx->ws_out.ensure(64, "pair output"));
        BS_HIP(hipMemcpyAsync(d, a_xyz, n_atoms * 24, hipMemcpyHostToDevice, s));
        BS_HIP(hipMemcpyAsync(d + 3 * n_atoms, b_xyz, n_atoms * 24, hipMemcpyHostToDevice, s));
        avsa_device(ctx, d, 2, (int)n_atoms, BS_FP64_PLAIN, o, true);
        BS_HIP(hipMemcpyAsync(out, o, 8, hipMemcpyDeviceToHost, s));
        BS_HIP(hipStreamSynchronize(s));
    });
}

bs_status bs_superpose_f64(bs_context* ctx, const double* a_xyz, const double* b_xyz,
                           size_t n_atoms, double* rot9, double* trans3, double* rmsd)
{
    return guard([&] {
        BS_REQUIRE(a_xyz && b_xyz && rot9 && trans3, "null pointer argument");
        check_shape(2, n_atoms);
        ctx = resolve_context(ctx);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        hipStream_t s = ctx->stream;
        double* d = static_cast<double*>(ctx->ws_xyz.ensure(n_atoms * 6 * 8, "pair coordinates"));
        double* o = static_cast<double*>(ctx->ws_out.ensure(16 * 8, "superposition output"));
        BS_HIP(hipMemcpyAsync(d, a_xyz, n_atoms * 24, hipMemcpyHostToDevice, s));
        BS_HIP(hipMemcpyAsync(d + 3 * n_atoms, b_xyz, n_atoms * 24, hipMemcpyHostToDevice, s));
        launch_superpose(d, d + 3 * n_atoms, (int)n_atoms, o, s);
        BS_HIP(hipGetLastError());
        double h[13];
        BS_HIP(hipMemcpyAsync(h, o, sizeof(h), hipMemcpyDeviceToHost, s));
        BS_HIP(hipStreamSynchronize(s));
        std::memcpy(rot9, h, 9 * 8);
        std::memcpy(trans3, h + 9, 3 * 8);
        if (rmsd) *rmsd = h[12];
    });
}

bs_status bs_superpose_ensemble(bs_context* ctx, const double* xyz, size_t n_conf, size_t n_atoms,
                                const double* ref_xyz, size_t ref_index, double* rot9, double* trans3,
                                double* rmsd, double* superposed)
{
    return guard([&] {
        check_shape(n_conf, n_atoms);
        BS_REQUIRE(n_conf >= 1, "n_conf must be >= 1");
        BS_REQUIRE(xyz, "null coordinates");
        BS_REQUIRE(ref_xyz || ref_index < n_conf, "ref_index %zu out of range (%zu conformations)",
                   ref_index, n_conf);
        ctx = resolve_context(ctx);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        hipStream_t s = ctx->stream;
        const size_t L = n_atoms, cb = n_conf * L * 24;
        char* in = static_cast<char*>(ctx->ws_xyz.ensure(cb + L * 24, "superposition input"));
        auto rup = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t o_tr = rup(n_conf * 72), o_rm = o_tr + rup(n_conf * 24), o_sp = o_rm + rup(n_conf * 8);
        char* ob = static_cast<char*>(ctx->ws_out.ensure(o_sp + (superposed ? cb : 0), "superposition output"));
        double* misc = static_cast<double*>(
            ctx->ws_misc.ensure(superpose_scratch_bytes((int)n_conf, (int)L), "superposition scratch"));
        BS_HIP(hipMemcpyAsync(in, xyz, cb, hipMemcpyHostToDevice, s));
        const double* d_ref = reinterpret_cast<const double*>(in) + ref_index * 3 * L;
        if (ref_xyz) {
            BS_HIP(hipMemcpyAsync(in + cb, ref_xyz, L * 24, hipMemcpyHostToDevice, s));
            d_ref = reinterpret_cast<const double*>(in + cb);
        }
        double* d_rot = reinterpret_cast<double*>(ob);
        double* d_tr = reinterpret_cast<double*>(ob + o_tr);
        double* d_rm = reinterpret_cast<double*>(ob + o_rm);
        double* d_sp = superposed ? reinterpret_cast<double*>(ob + o_sp) : nullptr;
        launch_superpose_ensemble(reinterpret_cast<const double*>(in), (int)n_conf, (int)L, d_ref, misc,
                                  d_rot, d_tr, d_rm, d_sp, s);
        BS_HIP(hipGetLastError());
        if (rot9) BS_HIP(hipMemcpyAsync(rot9, d_rot, n_conf * 72, hipMemcpyDeviceToHost, s));
        if (trans3) BS_HIP(hipMemcpyAsync(trans3, d_tr, n_conf * 24, hipMemcpyDeviceToHost, s));
        if (rmsd) BS_HIP(hipMemcpyAsync(rmsd, d_rm, n_conf * 8, hipMemcpyDeviceToHost, s));
        if (superposed) BS_HIP(hipMemcpyAsync(superposed, d_sp, cb, hipMemcpyDeviceToHost, s));
        BS_HIP(hipStreamSynchronize(s));
    });
}

bs_status bs_superpose_ensemble_device(bs_context* ctx, const double* d_xyz, size_t n_conf, size_t n_atoms,
                                       const double* d_ref, size_t ref_index, double* d_rot9,
                                       double* d_trans3, double* d_rmsd, double* d_superposed)
{
    return guard([&] {
        check_ctx(ctx);
        check_shape(n_conf, n_atoms);
        BS_REQUIRE(n_conf >= 1, "n_conf must be >= 1");
        BS_REQUIRE(d_xyz, "null coordinates");
        BS_REQUIRE(d_ref || ref_index < n_conf, "ref_index %zu out of range (%zu conformations)", ref_index,
                   n_conf);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        const size_t L = n_atoms;
        double* misc = static_cast<double*>(
            ctx->ws_misc.ensure(superpose_scratch_bytes((int)n_conf, (int)L), "superposition scratch"));
        launch_superpose_ensemble(d_xyz, (int)n_conf, (int)L, d_ref ? d_ref : d_xyz + ref_index * 3 * L, misc,
                                  d_rot9, d_trans3, d_rmsd, d_superposed, ctx->stream);
        BS_HIP(hipGetLastError());
    });
}

bs_status bs_crmsd_all_vs_all_device(bs_context* ctx, const double* d_xyz, size_t n_conf,
                                     size_t n_atoms, bs_method method, void* d_out_upper,
                                     int out_f64)
{
    return guard([&] {
        check_ctx(ctx);
        check_shape(n_conf, n_atoms);
        BS_REQUIRE(valid_method(method), "unknown method %d", (int)method);
        BS_REQUIRE(n_conf < 2 || (d_xyz && d_out_upper), "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        avsa_device(ctx, d_xyz, (int)n_conf, (int)n_atoms, method, d_out_upper, out_f64 != 0);
    });
}

bs_status bs_crmsd_all_vs_all_row_stats_device(bs_context* ctx, const double* d_xyz, size_t n_conf,
                                               size_t n_atoms, bs_method method, float* d_out_upper,
                                               float* d_row_min, float* d_row_max)
{
    return guard([&] {
        check_ctx(ctx);
        check_shape(n_conf, n_atoms);
        BS_REQUIRE(valid_method(method), "unknown method %d", (int)method);
        BS_REQUIRE(n_conf >= 2, "row statistics need at least 2 conformations");
        BS_REQUIRE(d_xyz && d_out_upper && d_row_min && d_row_max, "null pointer argument");
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        avsa_device(ctx, d_xyz, (int)n_conf, (int)n_atoms, method, d_out_upper, false, d_row_min,
                    d_row_max);
    });
}

bs_status bs_crmsd_all_vs_all(bs_context* ctx, const double* xyz, size_t n_conf, size_t n_atoms,
                              bs_method method, void* out_upper, int out_f64)
{
    return guard([&] {
        check_shape(n_conf, n_atoms);
        BS_REQUIRE(valid_method(method), "unknown method %d", (int)method);
        if (n_conf < 2) return;
        BS_REQUIRE(xyz && out_upper, "null pointer argument");
        ctx = resolve_context(ctx);
        std::lock_guard<std::mutex> lk(ctx->mu);
        DeviceGuard dg(ctx->device);
        hipStream_t s = ctx->stream;
        const size_t in_bytes = n_conf * n_atoms * 3 * 8;
        const size_t pairs = n_conf * (n_conf - 1) / 2;
        const size_t out_bytes = pairs * (out_f64 ? 8 : 4);
        double* d = static_cast<double*>(ctx->ws_xyz.ensure(in_bytes, "ensemble coordinates"));
        void* o = ctx->ws_out.ensure(out_bytes, "rmsd matrix");
        BS_HIP(hipMemcpyAsync(d, xyz, in_bytes, hipMemcpyHostToDevice, s));
        avsa_device(ctx, d, (int)n_conf, (int)n_atoms, method, o, out_f64 != 0);
        BS_HIP(hipMemcpyAsync(out_upper, o, out_bytes, hipMemcpyDeviceToHost, s));
        BS_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
