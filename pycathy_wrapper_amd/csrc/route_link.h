// route_link.h -- a routing handle (cathy_gpu_route_*) as the flow handle's
// coupled step uses it: its substep launched on the caller's stream over
// device buffers, for one member.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>

namespace cg {

struct RouteLink {
  std::shared_ptr<void> keep;  // holds the routing handle while linked
  void *route = nullptr;
  int nrow = 0, ncol = 0, ngrid = 0, ncell = 0, numres = 0;
  bool typed = false;  // ALTEZZE cell types other than 0 were set
};

// Looks the routing handle up (throws Error on an unknown handle).
RouteLink route_link(const int64_t *handle);

// One ROUTE + ALTEZZE substep (route.f, altezze.f) of one member on stream
// st.  state: SURFACE_WATER_SN, Q_IN_KK_SN, Q_OUT_KK_SN_1, Q_OUT_KK_SN_2,
// VOLUME_KK_SN, Q_IN_KKP1_SN, Q_OUT_KKP1_SN_1, Q_OUT_KKP1_SN_2,
// VOLUME_KKP1_SN, H_WATER_KKP1_SN, each ngrid doubles by I_BASIN - 1, in
// that order, contiguous.  cu_ak[2]: CU_MAX (out), AK_MAX (in, replaced
// where a branch sets it).  err[1]: 0.
void route_substep(const RouteLink &l, hipStream_t st, double deltat, double dx, double dy, double *state,
                   double *cu_ak, int *err);

}  // namespace cg
