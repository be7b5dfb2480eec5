// rbs_pending.hip -- entry points whose kernels are not built yet.  Each fails
// loudly with RBS_E_STATE rather than computing anything on another path.
#include "rbs_device.hpp"

namespace rbs {
int hllc_num_blocks(const rbs_state *s) { return 1 + 0 * s->nx; }
int hllc_launch_step(rbs_state *, bool, double, double, double, double) {
  set_error("HLLC kernels are not built in this library");
  return RBS_E_STATE;
}
int inflow_launch(rbs_state *, bool, double, double) {
  set_error("point-source kernels are not built in this library");
  return RBS_E_STATE;
}
}  // namespace rbs

using namespace rbs;
extern "C" {
int rbs_hllc_step(rbs_state *, double, double, double, double) { return hllc_launch_step(nullptr, false, 0, 0, 0, 0); }
int rbs_set_inflows(rbs_state *, int32_t, const int32_t *, const int32_t *, const int32_t *,
                    const double *, const double *) { set_error("not built"); return RBS_E_STATE; }
int rbs_inflow_apply(rbs_state *, double, double, double *) { set_error("not built"); return RBS_E_STATE; }
int rbs_exner_activate(rbs_state *, double) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_create(int32_t, rbs_batch **) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_destroy(rbs_batch *) { return RBS_OK; }
int rbs_batch_add(rbs_batch *, rbs_state *, const rbs_run_params *, double, int32_t *) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_remove(rbs_batch *, int32_t) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_advance(rbs_batch *, int32_t, int32_t *) { set_error("not built"); return RBS_E_STATE; }
int rbs_batch_member_active(rbs_batch *, int32_t, int32_t *) { set_error("not built"); return RBS_E_STATE; }
}
