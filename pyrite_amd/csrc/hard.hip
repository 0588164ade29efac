// hard.hip -- table-side entry points of the coupled radiative mode.
//
// hr_radiative_lut evaluates the exact per-element soft and hard moments
// (radiative.hpp) at every node of the transport LUT's ln-E grid, one thread
// per (layer, node): the step interpolates these columns like the other LUT
// columns, the reference's prepared moment lookup
// (src/pyrite/montecarlo/transport/_jit_moment_lookup_device.py:10-70) placed
// on this build's transport grid.  hr_radiative_sample exposes the event's
// photon sampler on its own, for checks against the CPU restatement.
#include "radiative.hpp"

namespace hr {
namespace {

rad::Tables tables_of(const hr_hard_events &h, int32_t max_el, const int32_t *nel) {
    rad::Tables t;
    t.incident = h.rad_incident;
    t.nominal = h.rad_nominal;
    t.top = h.rad_top;
    t.chi = h.rad_chi;
    t.z = h.rad_z;
    t.ncm3 = h.rad_ncm3;
    t.nt = h.rad_nt;
    t.nel = nel;
    t.plasma = h.rad_plasma;
    t.max_el = max_el;
    t.max_t = h.rad_max_t;
    t.n_red = h.rad_n_red;
    t.cutoff_eV = h.rad_cutoff_eV;
    return t;
}

__global__ void __launch_bounds__(256) radiative_lut_kernel(rad::Tables t, int32_t n_layers,
                                                            const double *__restrict__ E_keV, int64_t n_E,
                                                            double *__restrict__ soft, double *__restrict__ rate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n_layers * n_E) return;
    const int L = (int)(i / n_E);
    const double E_eV = E_keV[i % n_E] * 1.0e3;
    soft[i] = rad::layer_moment(rad::SOFT, t, L, E_eV) * 1.0e-3;  // eV/A -> keV/A
    rate[i] = rad::layer_moment(rad::HARD_RATE, t, L, E_eV);
}

__global__ void __launch_bounds__(256) radiative_sample_kernel(rad::Tables t, int32_t slot, double plasma,
                                                               int64_t n, const double *__restrict__ E_eV,
                                                               const double *__restrict__ u,
                                                               double *__restrict__ k) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    k[i] = rad::sample_photon(t, slot, E_eV[i], u[i], plasma);
}

hr_status check_tables(const hr_hard_events *h, const char *who) {
    HR_REQUIRE(h != nullptr, "%s: null tables", who);
    HR_REQUIRE(h->rad_incident && h->rad_nominal && h->rad_top && h->rad_chi && h->rad_z && h->rad_ncm3 &&
                   h->rad_nt && h->rad_plasma,
               "%s: null radiative table pointer", who);
    HR_REQUIRE(h->rad_max_t >= 2 && h->rad_n_red >= 2 && h->rad_cutoff_eV > 0.0,
               "%s: radiative tables need >= 2 incident rows and reduced nodes and a positive cutoff", who);
    return HR_OK;
}

}  // namespace
}  // namespace hr

extern "C" hr_status hr_radiative_lut(const hr_hard_events *h, int32_t n_layers, int32_t max_el,
                                      const int32_t *L_nel, const double *E_grid_keV, int64_t n_E, double *soft_out,
                                      double *rate_out, void *stream) {
    HR_TRACE("hr_radiative_lut");
    using namespace hr;
    hr_status st = check_tables(h, "hr_radiative_lut");
    if (st != HR_OK) return st;
    HR_REQUIRE(n_layers >= 1 && max_el >= 1 && n_E >= 0, "hr_radiative_lut: bad sizes (n_layers=%d max_el=%d)",
               n_layers, max_el);
    if (n_E == 0) return HR_OK;
    HR_REQUIRE(L_nel && E_grid_keV && soft_out && rate_out, "hr_radiative_lut: null element counts, grid or output");
    HR_REQUIRE((int64_t)n_layers * n_E < (int64_t)1 << 40, "hr_radiative_lut: too many nodes");
    const int64_t total = (int64_t)n_layers * n_E;
    hipLaunchKernelGGL(radiative_lut_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, tables_of(*h, max_el, L_nel), n_layers, E_grid_keV, n_E, soft_out,
                       rate_out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_radiative_sample(const hr_hard_events *h, int32_t slot, double plasma_eV, int64_t n,
                                         const double *E_eV, const double *u, double *k_out, void *stream) {
    HR_TRACE("hr_radiative_sample");
    using namespace hr;
    hr_status st = check_tables(h, "hr_radiative_sample");
    if (st != HR_OK) return st;
    HR_REQUIRE(slot >= 0 && n >= 0 && plasma_eV >= 0.0, "hr_radiative_sample: negative slot, count or plasma energy");
    if (n == 0) return HR_OK;
    HR_REQUIRE(E_eV && u && k_out, "hr_radiative_sample: null input or output");
    hipLaunchKernelGGL(radiative_sample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, tables_of(*h, 1, nullptr), slot, plasma_eV, n, E_eV, u, k_out);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
