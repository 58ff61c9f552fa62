// Bath correlation integrals on the device: eta(tau) of a power-law spectral
// density for many times (and baths) in one launch, the input of every
// influence tensor (oqupy/bath_correlations.py:518-591, 667-759).
//
//   eta(tau) = int_0^inf J(w) / w^2 [ (1 - cos w tau) coth(w / 2T)
//                                     + i (sin w tau - w tau) ] dw,
//   J(w) = 2 alpha w^zeta w_c^(1-zeta) X(w),
//
// X the cutoff (exponential exp(-w/w_c), Gaussian exp(-(w/w_c)^2) or hard
// w < w_c), coth = 1 at T = 0. The integral is a composite 16-point
// Gauss-Legendre rule over uniform panels of [0, W] (W = 40 w_c, 6.5 w_c or
// w_c: the neglected tail is below 1e-17 of the integrand's scale), each panel
// at most 1 / tau_max wide so that a panel spans less than one radian of
// the oscillation. 1 - cos and sin x - x are evaluated in forms without
// cancellation (2 sin^2(x/2) and a series for |x| < 0.1).
#include <hip/hip_runtime.h>

#include "pte_device.h"
#include "pte_eta.h"

namespace pte {

__device__ __forceinline__ double sin_minus_x(double x) {
  if (fabs(x) < 0.1) {
    const double x2 = x * x;
    return x * x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 / 362880.0)));
  }
  return sin(x) - x;
}

// one workgroup per time
__global__ void __launch_bounds__(256)
eta_kernel(const EtaBath* baths, const int* tau_bath, const double* taus, int ntau,
           const double* glx, const double* glw, double2* out) {
  __shared__ double slot[2][4], sloti[2][4];
  int parity = 0;
  const int i = blockIdx.x;
  if (i >= ntau) return;
  const EtaBath b = baths[tau_bath[i]];
  const double tau = taus[i];
  const double pre = 2.0 * b.alpha * pow(b.wc, 1.0 - b.zeta);
  const long nodes = (long)b.npan * kEtaNodes;
  double2 acc = make_double2(0.0, 0.0);
  for (long e = threadIdx.x; e < nodes; e += 256) {
    const int p = (int)(e / kEtaNodes), q = (int)(e % kEtaNodes);
    const double w = b.width * (p + 0.5 * (glx[q] + 1.0));
    double cut;
    if (b.cutoff == ETA_CUTOFF_EXPONENTIAL) cut = exp(-w / b.wc);
    else if (b.cutoff == ETA_CUTOFF_GAUSSIAN) cut = exp(-(w / b.wc) * (w / b.wc));
    else cut = (w < b.wc) ? 1.0 : 0.0;
    const double jw2 = pre * pow(w, b.zeta - 2.0) * cut;    // J(w) / w^2
    const double x = w * tau;
    const double s = sin(0.5 * x);
    const double coth = b.temp > 0.0 ? 1.0 / tanh(w / (2.0 * b.temp)) : 1.0;
    const double f = 0.5 * b.width * glw[q] * jw2;
    acc.x += f * (2.0 * s * s * coth);
    acc.y += f * sin_minus_x(x);
  }
  const double2 r = block_sum256c(acc, slot, sloti, parity);
  if (threadIdx.x == 0) out[i] = r;
}

}  // namespace pte
