// mixed.hip -- mixed-precision iterative refinement (filled in below).
#include "se_common.h"
#include "se_op.h"

namespace se {
int pcg_mixed(se_op* op, const double* d_f, double* d_x, const se_solve_opts& o, se_solve_info& info) {
    (void)op; (void)d_f; (void)d_x; (void)o; (void)info;
    return fail(SE_ERR_UNSUPPORTED, "mixed precision not built yet");
}
}  // namespace se
