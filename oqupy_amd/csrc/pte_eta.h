// eta(tau) of power-law baths on the device (see pte_eta.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace pte {

enum EtaCutoff { ETA_CUTOFF_EXPONENTIAL = 0, ETA_CUTOFF_GAUSSIAN = 1, ETA_CUTOFF_HARD = 2 };

constexpr int kEtaNodes = 16;   // Gauss-Legendre points per panel

struct EtaBath {
  double alpha, zeta, wc, temp;
  double width;    // panel width
  int npan;        // panels over [0, W]
  int cutoff;      // EtaCutoff
};

__global__ void eta_kernel(const EtaBath* baths, const int* tau_bath, const double* taus,
                           int ntau, const double* glx, const double* glw, double2* out);

}  // namespace pte
