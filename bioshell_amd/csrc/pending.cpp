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
}  // extern "C"
