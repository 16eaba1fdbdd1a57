// pending.cpp -- entry points whose implementation lands in later commits.
#include "engine.hpp"

using namespace bs;

extern "C" {

bs_status bs_crmsd_all_vs_all_streamed(bs_context*, const double*, size_t, size_t, bs_method,
                                       size_t, float*)
{
    return guard([] { fail(BS_ERR_UNSUPPORTED, "not implemented yet"); });
}
}  // extern "C"
