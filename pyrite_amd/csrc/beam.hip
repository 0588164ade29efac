// beam.hip -- incident electron states on the entrance face, on the device.
//
// Follows PyRITE's ungrooved analytic beam entry
// (src/pyrite/montecarlo/transport/beam_entry.py:23-126 initial_beam_positions,
// :347-359 primary_directions; transverse.py:169-224 sample_transverse;
// geometry.py:330-394 project_beam_entry): per-electron standard normals are
// counter addressed inside a child-stream namespace
// (kinematics.py:267-312 counter_uniforms / counter_normals), so electron e's
// entry state depends on (seed, e) alone and any block [start, start+n) equals
// that slice of one large draw.  At 1e8 electrons the states are generated
// where they are used instead of being built and copied from the host.
#include <math.h>

#include "hr_common.hpp"

namespace hr {
namespace {

constexpr double kU52 = 1.0 / 4503599627370496.0;  // 2^-52
constexpr double kSqrt2 = 1.4142135623730951;

// ((splitmix64(k_e + PHI (c+1)) >> 12) + 1/2) 2^-52, k_e = splitmix64(root + PHI (e+1))
__device__ __forceinline__ double counter_uniform(uint64_t root, uint64_t e, uint64_t c) {
    const uint64_t key = splitmix64(root + kGolden * (e + 1ULL));
    const uint64_t z = splitmix64(key + kGolden * (c + 1ULL));
    return ((double)(z >> 12) + 0.5) * kU52;
}

// Phi^-1(u) = -sqrt(2) erfcinv(2u) (the standard normal quantile, ndtri).
__device__ __forceinline__ double normal_quantile(double u) { return -kSqrt2 * erfcinv(2.0 * u); }

// |z| clamp of the energy-spread normal: E0 (1 + 6 |f|) is a strict upper energy
// (beam_entry.py:244-248)
constexpr double kSpreadClamp = 6.0;

__global__ void __launch_bounds__(256) beam_entry_kernel(hr_beam_args p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const uint64_t e = (uint64_t)(p.start + i);
    double u = 0.0, v = 0.0, xp = 0.0, yp = 0.0;
    if (p.mode == 1) {
        u = normal_quantile(counter_uniform(p.root, e, 0)) * p.sigma_x;
        v = normal_quantile(counter_uniform(p.root, e, 1)) * p.sigma_y;
    } else if (p.mode == 2) {
        // x = sqrt(eps beta) z0, x' = sqrt(eps/beta) (z1 - alpha z0); y from draws 2, 3
        const double z0 = normal_quantile(counter_uniform(p.root, e, 0));
        const double z1 = normal_quantile(counter_uniform(p.root, e, 1));
        const double z2 = normal_quantile(counter_uniform(p.root, e, 2));
        const double z3 = normal_quantile(counter_uniform(p.root, e, 3));
        u = (p.twiss_x[0] * z0) * 1.0e7;
        xp = p.twiss_x[1] * (z1 - p.twiss_x[2] * z0);
        v = (p.twiss_y[0] * z2) * 1.0e7;
        yp = p.twiss_y[1] * (z3 - p.twiss_y[2] * z2);
    }
    double px = u, py = v;
    if (p.projected) {
        // o = u e_x + v e_y; p0 = o + s* b with s* = -o_z / b_z, b the lab beam axis
        const double ox = u * p.e_x[0] + v * p.e_y[0];
        const double oy = u * p.e_x[1] + v * p.e_y[1];
        const double oz = u * p.e_x[2] + v * p.e_y[2];
        const double s = -oz / p.axis[2];
        px = ox + s * p.axis[0];
        py = oy + s * p.axis[1];
    }
    p.pos[3 * i] = px;
    p.pos[3 * i + 1] = py;
    p.pos[3 * i + 2] = 0.0;
    double dx = p.beam_dir[0], dy = p.beam_dir[1], dz = p.beam_dir[2];
    if (p.mode == 2) {
        // (x' b0 + y' b1 + beam_dir) / |.|, b0, b1 the first two columns of the beam-frame basis
        dx = xp * p.basis[0] + yp * p.basis[1] + p.beam_dir[0];
        dy = xp * p.basis[3] + yp * p.basis[4] + p.beam_dir[1];
        dz = xp * p.basis[6] + yp * p.basis[7] + p.beam_dir[2];
        const double nrm = sqrt(dx * dx + dy * dy + dz * dz);
        dx = dx / nrm;
        dy = dy / nrm;
        dz = dz / nrm;
    }
    p.dirs[3 * i] = dx;
    p.dirs[3 * i + 1] = dy;
    p.dirs[3 * i + 2] = dz;
    double E = p.E0_keV;
    if (p.energy_spread_frac != 0.0) {
        const double z = fmin(kSpreadClamp, fmax(-kSpreadClamp, normal_quantile(counter_uniform(p.energy_root, e, 0))));
        E = E * (1.0 + p.energy_spread_frac * z);
    }
    p.E_keV[i] = E;
    p.clock[i] = 0.0;
}

}  // namespace
}  // namespace hr

extern "C" hr_status hr_beam_entry(const hr_beam_args *p, void *stream) {
    HR_TRACE("hr_beam_entry");
    using namespace hr;
    HR_REQUIRE(p != nullptr && p->n >= 0 && p->start >= 0, "hr_beam_entry: bad args");
    if (p->n == 0) return HR_OK;
    HR_REQUIRE(p->mode >= 0 && p->mode <= 2, "hr_beam_entry: mode %d is not 0 (pencil), 1 (spot) or 2 (twiss)",
               p->mode);
    HR_REQUIRE(p->pos && p->dirs && p->E_keV && p->clock, "hr_beam_entry: null output");
    HR_REQUIRE(p->beam_dir[2] > 1e-6, "hr_beam_entry: beam must point into the slab");
    HR_REQUIRE(!p->projected || p->axis[2] > 1e-6, "hr_beam_entry: the projection axis must point into the slab");
    HR_REQUIRE(p->E0_keV > 0.0, "hr_beam_entry: E0 must be positive");
    HR_REQUIRE(std::isfinite(p->energy_spread_frac) && fabs(p->energy_spread_frac) * kSpreadClamp < 1.0,
               "hr_beam_entry: energy_spread_frac=%g could draw a non-positive energy (|f| * 6 must stay below 1; "
               "the Gaussian spread model needs spread << 1)", p->energy_spread_frac);
    hipLaunchKernelGGL(beam_entry_kernel, dim3((unsigned)((p->n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       *p);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
