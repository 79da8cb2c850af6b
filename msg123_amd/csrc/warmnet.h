// warmnet.h -- internal interface between the Newton driver and the
// warm-start network (warmnet.hip).  Device pointers, the network's stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct msg123_net;

namespace msg123 {

int net_n(const msg123_net *net);
hipStream_t net_stream(const msg123_net *net);
// features x [N][4] bf16, h_lin = h_n + r (h_n - h_nm1) and the increment
// scale (device scalar) from the last two heads
int net_features(msg123_net *net, const double *hn, const double *hm, const double *top,
                 const float *lnks, double r, double L, double rech_depth, uint16_t *x,
                 double *hlin, double *scale_dev);
// out = h_lin + scale * net(x) / L in cells unsaturated at h_n (h_n below the
// cell top), h_lin elsewhere: the saturated zone is elliptic, the residual
// there is very sensitive to head errors, and linear extrapolation is already
// accurate in it
int net_guess(msg123_net *net, const uint16_t *x, const double *hlin, const double *scale_dev,
              double L, const double *hn, const double *top, double *out);

}  // namespace msg123
