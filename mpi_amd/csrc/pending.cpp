// pending.cpp -- entry points declared in include/hz.h whose kernels are
// not built yet: they refuse with HZ_ERR_UNSUPPORTED.
#include "hz_internal.h"

using namespace hz;

extern "C" {
int hz_batch_create(hz_plan *const *, int, hz_batch **)
{
    return set_error(HZ_ERR_UNSUPPORTED, "hz_batch_create: not built yet");
}
int hz_batch_solve(hz_batch *, const double *const *, const double *const *, const double *const *,
                   double *const *, hz_stream_t)
{
    return set_error(HZ_ERR_UNSUPPORTED, "hz_batch_solve: not built yet");
}
int hz_batch_destroy(hz_batch *) { return HZ_OK; }
}
