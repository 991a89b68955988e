// Timing-driven net weighting on the device: the weighting step of
// timing_cpp.update_net_weights (dreamplace/ops/timing/src/
// net_weighting_scheme.h:106-190), given what a static timer reports.
//
//   ADAMS  (:106-148): per net of degree <= ignore_net_degree,
//          criticality = 0.5 * criticality + (0.5 if the net lies on a
//          reported critical path), weight *= 1 + criticality;
//   LILITH (:150-188): when wns < 0, per net
//          nc = slack < 0 ? max(0, slack / wns) : 0,
//          criticality = (1 + criticality)^decay * (1 + nc)^(1 - decay) - 1
//          (every net); then per net of degree <= ignore_net_degree
//          weight = min(weight * (1 + criticality), max_net_weight).
// One thread per net; the arithmetic is in T with the reference's order.
#include "dp_common.h"

namespace dp {

template <typename T>
__global__ void __launch_bounds__(256) k_net_weighting(int scheme, int64_t num_nets,
                                                       const T* __restrict__ slack,
                                                       const uint8_t* __restrict__ critical,
                                                       T wns, const int* __restrict__ start,
                                                       int ignore_deg, T decay, T max_w, T* crit,
                                                       T* w) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= num_nets) return;
  const int deg = start[e + 1] - start[e];
  if (scheme == DP_NET_WEIGHTING_ADAMS) {
    if (deg > ignore_deg) return;
    T c = crit[e] * T(0.5);
    if (critical[e]) c += T(0.5);
    crit[e] = c;
    w[e] *= (T(1) + c);
    return;
  }
  T c = crit[e];
  if (wns < T(0)) {
    const T s = slack[e];
    T nc = T(0);
    if (s < T(0)) {
      const T q = s / wns;
      nc = T(0) < q ? q : T(0);
    }
    c = pow(T(1) + c, decay) * pow(T(1) + nc, T(1) - decay) - T(1);
    crit[e] = c;
  }
  if (deg > ignore_deg) return;
  T nw = w[e] * (T(1) + c);
  if (nw > max_w) nw = max_w;
  w[e] = nw;
}

}  // namespace dp

using namespace dp;

extern "C" int dp_net_weighting(int dtype, int scheme, int64_t num_nets, const void* net_slack,
                                const uint8_t* net_critical, double wns,
                                const int32_t* netpin_start, int ignore_net_degree, double decay,
                                double max_net_weight, void* net_criticality, void* net_weights,
                                void* stream) {
  DP_ARG(scheme == DP_NET_WEIGHTING_ADAMS || scheme == DP_NET_WEIGHTING_LILITH,
         "dp_net_weighting: unknown scheme %d", scheme);
  DP_ARG(netpin_start && net_criticality && net_weights, "dp_net_weighting: null pointer");
  DP_ARG(scheme != DP_NET_WEIGHTING_ADAMS || net_critical,
         "dp_net_weighting: ADAMS needs the per-net critical flags");
  DP_ARG(scheme != DP_NET_WEIGHTING_LILITH || net_slack,
         "dp_net_weighting: LILITH needs the per-net slacks");
  DP_ARG(num_nets >= 0 && num_nets < ((int64_t)1 << 31), "dp_net_weighting: bad num_nets");
  if (num_nets == 0) return DP_OK;
  hipStream_t s = S(stream);
  return DP_DISPATCH(dtype, "dp_net_weighting", [&] {
    hipLaunchKernelGGL(k_net_weighting<T>, dim3((unsigned)ceil_div(num_nets, 256)), dim3(256), 0,
                       s, scheme, num_nets, (const T*)net_slack, net_critical, (T)wns,
                       netpin_start, ignore_net_degree, (T)decay, (T)max_net_weight,
                       (T*)net_criticality, (T*)net_weights);
    DP_LAUNCHED("k_net_weighting");
    return DP_OK;
  });
}
