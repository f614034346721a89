// placeholder for the batched large-N path
#include "internal.hpp"
namespace trhepd {
int run_batched(trhepd_ctx *ctx, const Table &, int nb, int, const int *, const double2 *, double2 *, double *, int,
                const double *, int *)
{
    return fail(ctx, TRHEPD_E_ARG, "batched path not built yet (N=%d)", nb);
}
}  // namespace trhepd
