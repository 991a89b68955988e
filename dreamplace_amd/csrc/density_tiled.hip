// LDS-binned density scatter (DP_SCATTER_TILED).
#include "dp_density.h"

namespace dp {

size_t tiled_ws_bytes(int64_t num_bins) {
  (void)num_bins;
  return 0;
}

template <typename T>
int scatter_tiled(const TriArgs<T>& a, const int64_t ranges[2][2], char* tiled_ws,
                  unsigned long long* map, int* qcount, int* qlist, int64_t qcap, hipStream_t s) {
  (void)a; (void)ranges; (void)tiled_ws; (void)map; (void)qcount; (void)qlist; (void)qcap; (void)s;
  set_error("dp_density_map: DP_SCATTER_TILED is not built in this version");
  return DP_ERR_UNSUPPORTED;
}

template int scatter_tiled<float>(const TriArgs<float>&, const int64_t[2][2], char*,
                                  unsigned long long*, int*, int*, int64_t, hipStream_t);
template int scatter_tiled<double>(const TriArgs<double>&, const int64_t[2][2], char*,
                                   unsigned long long*, int*, int*, int64_t, hipStream_t);

}  // namespace dp
