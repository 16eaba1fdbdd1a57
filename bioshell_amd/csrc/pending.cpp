// pending.cpp -- entry points whose implementation lands in later commits.
#include "engine.hpp"

using namespace bs;

extern "C" {

bs_status bs_crmsd_all_vs_all_streamed(bs_context*, const double*, size_t, size_t, bs_method,
                                       size_t, float*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
bs_status bs_crmsd_batch(bs_context*, const double* const*, const size_t*, const size_t*, size_t,
                         bs_method, float* const*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
bs_status bs_autotune(bs_context*, bs_method, const char*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
bs_status bs_autotune_load(bs_context*, const char*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
bs_status bs_library_create(bs_context*, const double*, size_t, size_t, bs_library**)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
void bs_library_destroy(bs_library*) {}
size_t bs_library_size(const bs_library*) { return 0; }
size_t bs_library_atoms(const bs_library*) { return 0; }
bs_status bs_query_topk(bs_context*, bs_library*, const double*, size_t, size_t, uint32_t*, float*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
bs_status bs_query_all(bs_context*, bs_library*, const double*, size_t, float*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
bs_status bs_server_create(bs_context*, bs_library* const*, size_t, size_t, int, bs_server**)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
void bs_server_destroy(bs_server*) {}
bs_status bs_server_submit(bs_server*, size_t, const double*, size_t, size_t, uint32_t*, float*,
                           uint64_t*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
bs_status bs_server_poll(bs_server*, uint64_t) { return BS_ERR_UNSUPPORTED; }
bs_status bs_server_wait(bs_server*, uint64_t) { return BS_ERR_UNSUPPORTED; }
bs_status bs_server_cancel(bs_server*, uint64_t) { return BS_ERR_UNSUPPORTED; }
bs_status bs_server_latency_ms(bs_server*, uint64_t, double*) { return BS_ERR_UNSUPPORTED; }

}  // extern "C"
