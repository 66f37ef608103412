// chkn_grad_off.cpp -- the gradient entry points of include/chkn.h in the FP32 flavours.
//
// The hand-written adjoint (chkn_adjoint.hip, chkn_grad.cpp) is built for FP64 only: its
// flat device arrays are doubles and its checks are against FP64 finite differences.  In
// the FP32 flavours every gradient entry point fails with CHKN_ERR_STATE and says so.
#include <string>

#include "../../include/chkn.h"
#include "chkn_internal.h"
#include "chkn_runtime.h"

namespace chkn {
namespace rt {

void free_grad(chkn_sim*) {}

namespace {
int off(chkn_sim* s)
{
    return set_error(s, CHKN_ERR_STATE, "gradients are built for the FP64 flavour only");
}
}  // namespace

}  // namespace rt
}  // namespace chkn

using namespace chkn::rt;

extern "C" {

int chkn_grad_forward(chkn_sim* sim, int, int, const double*, double*) { return off(sim); }
int chkn_grad_backward(chkn_sim* sim, const double*, double*, double*) { return off(sim); }
long long chkn_state_size(chkn_sim* sim)
{
    off(sim);
    return -1;
}
long long chkn_grad_generation(chkn_sim* sim)
{
    off(sim);
    return -1;
}
int chkn_grad_set_tape_budget(chkn_sim* sim, long long) { return off(sim); }
int chkn_grad_set_gather_form(chkn_sim* sim, int) { return off(sim); }
int chkn_grad_gather_form(chkn_sim* sim) { return off(sim); }
int chkn_grad_resident_segments(chkn_sim* sim)
{
    off(sim);
    return -1;
}
int chkn_num_flow_states(chkn_sim* sim) { return sim ? (int)sim->inflow_w.size() / 6 : -1; }
int chkn_read_conserved_dev(chkn_sim* sim, double*) { return off(sim); }
int chkn_write_conserved_dev(chkn_sim* sim, const double*) { return off(sim); }
int chkn_set_inflow_state(chkn_sim* sim, int, const double*) { return off(sim); }
int chkn_get_inflow_state(chkn_sim* sim, int, double*) { return off(sim); }
int chkn_boundary_face_areas(chkn_sim* sim, int, int, double*) { return off(sim); }
int chkn_flux_vjp_points(int, int, const double*, double*) { return off(nullptr); }
int chkn_flux_vjp_points_noalloc(int, int, const double*, double*) { return off(nullptr); }

}  // extern "C"
