// WRF-Chem input adapter: the reference's dev_convert (mgpu.cu:513-646)
// over a box of grid cells, writing the solver's structure-of-arrays inputs.
//
// Per cell (mgpu.cu:537-645):
//   M      = dens2con_a * rho                       dens2con_a = 1e-3/mwair*N_A
//   H2O    = dens2con_w * qv * rho                  dens2con_w = 1e-3/mwh2o*N_A
//   es     = 1000*0.6112*exp(17.67*(T-273.15)/(T-29.65)), qvs = es/(P-es)
//   rh     = MIN(MAX(qv/qvs, 0), 1)
//   rc_n2o5 = 1/(3.6e4*exp(-(rh/0.28)^2.8 + 300))   as written (mgpu.cu:552)
//          or 1/(3.6e4*exp(-(rh/0.28)^2.8) + 300)   with the parenthesis fixed
//   J_s    = J_wrf,s / 60 (1/min -> 1/s) for the 21 rates the mechanism uses
//   conc_s = 1e-6 * dens2con_a * rho * MAX(x_s, 0)  (ppmv -> molec/cm3), x_s
//            from the chemistry or radical array per the shuffle table
// Inputs are WRF arrays [field][ny][nz][nx] (the reference's 48 x 20 x 48
// with a 4-cell halo in x and y); the box cell (j, k, i) reads input
// (j + oy, k + oz, i + ox) and writes output p = (j*bz + k)*bx + i.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "radm2.h"
#include "radm2_solver.h"

namespace radm2 {
namespace wrf {

constexpr double DENS2CON_A = 1.0e-3 * (1.0 / gen::WRF_MWAIR) * gen::WRF_NAVGDRO;
constexpr double DENS2CON_W = 1.0e-3 * (1.0 / gen::WRF_MWH2O) * gen::WRF_NAVGDRO;

__global__ void k_wrf_convert(radm2_wrf_grid g, const double* __restrict__ phy,
                              const double* __restrict__ moist, const double* __restrict__ ph,
                              const double* __restrict__ chem, const double* __restrict__ rad,
                              int rc_fixed, int64_t ld, double* __restrict__ conc,
                              double* __restrict__ temp, double* __restrict__ cair,
                              double* __restrict__ h2o, double* __restrict__ jrate,
                              double* __restrict__ rc) {
  const int64_t n = (int64_t)g.bx * g.by * g.bz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = (int)(p % g.bx);
  const int64_t t = p / g.bx;
  const int k = (int)(t % g.bz);
  const int j = (int)(t / g.bz);
  const int64_t fs = (int64_t)g.nx * g.ny * g.nz;
  const int64_t q = ((int64_t)(j + g.oy) * g.nz + (k + g.oz)) * g.nx + (i + g.ox);
  const double P = phy[q], T = phy[fs + q], rho = phy[2 * fs + q];
  const double qv = moist[fs + q];
  cair[p] = DENS2CON_A * rho;
  h2o[p] = DENS2CON_W * qv * rho;
  temp[p] = T;
  const double es = 1000. * 0.6112 * exp(17.67 * (T - 273.15) / (T - 29.65));
  const double qvs = es / (P - es);
  const double rh = dmin_ref(dmax_ref(qv / qvs, 0.0), 1.0);
  rc[p] = rc_fixed ? 1.0 / (3.6e4 * exp(-pow(rh / 0.28, 2.8)) + 300.0)
                   : 1.0 / (3.6e4 * exp(-pow(rh / 0.28, 2.8) + 300.0));
  const double conv = 1.0E-6 * DENS2CON_A * rho;
  for (int s = 0; s < RADM2_NJ; s++) jrate[s * ld + p] = ph[s * fs + q] / 60.0;
#pragma unroll
  for (int s = 0; s < RADM2_NVAR; s++) {
    const double* a = gen::WRF_SRC[s] ? rad : chem;
    conc[s * ld + p] = conv * dmax_ref(a[gen::WRF_FIELD[s] * fs + q], 0.0);
  }
}

}  // namespace wrf
}  // namespace radm2

extern "C" int radm2_wrf_convert(void* stream, const radm2_wrf_grid* g, const double* phy,
                                 const double* moist, const double* ph, const double* chem,
                                 const double* rad, int rc_fixed, int64_t ld, double* conc,
                                 double* temp, double* cair, double* h2o, double* jrate,
                                 double* rc_n2o5) {
  if (!g || !phy || !moist || !ph || !chem || !rad || !conc || !temp || !cair || !h2o || !jrate ||
      !rc_n2o5)
    return RADM2_E_ARG;
  if (g->bx <= 0 || g->by <= 0 || g->bz <= 0 || g->ox < 0 || g->oy < 0 || g->oz < 0 ||
      g->ox + g->bx > g->nx || g->oy + g->by > g->ny || g->oz + g->bz > g->nz)
    return RADM2_E_ARG;
  const int64_t n = (int64_t)g->bx * g->by * g->bz;
  if (ld < n || (rc_fixed != 0 && rc_fixed != 1)) return RADM2_E_ARG;
  hipLaunchKernelGGL(radm2::wrf::k_wrf_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, *g, phy, moist, ph, chem, rad, rc_fixed, ld, conc, temp,
                     cair, h2o, jrate, rc_n2o5);
  return hipGetLastError() == hipSuccess ? RADM2_OK : RADM2_E_HIP;
}
