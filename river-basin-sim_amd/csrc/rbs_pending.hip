// rbs_pending.hip -- entry points whose kernels are not built yet.  Each fails
// loudly with RBS_E_STATE rather than computing anything on another path.
#include "rbs_device.hpp"

namespace rbs {
}  // namespace rbs

using namespace rbs;
extern "C" {
int rbs_batch_create(int32_t, rbs_batch **) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_destroy(rbs_batch *) { return RBS_OK; }
int rbs_batch_add(rbs_batch *, rbs_state *, const rbs_run_params *, double, int32_t *) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_remove(rbs_batch *, int32_t) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_advance(rbs_batch *, int32_t, int32_t *) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_member_active(rbs_batch *, int32_t, int32_t *) { set_error("not built"); return RBS_E_STATE; }
}
