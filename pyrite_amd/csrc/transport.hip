// transport.hip -- per-electron LUT transport on gfx950.
//
// Physics and draw order follow PyRITE's per-electron LUT core
// (src/pyrite/montecarlo/transport/cores.py:365-1250 and its CUDA twin
// _jit_lut_kernel.py:64-574): screened-Rutherford elastic flights from a
// uniform-ln(E) LUT, spliced Joy-Luo/Berger-Seltzer stopping with the midpoint
// rule, Urban straggling on a per-(flight, substep) key domain, slab or prism
// geometry, and the cap/slot row contract.
//
// Execution model (MI355X-first, not the reference's one-thread-one-electron
// grid): a persistent grid sized to the device's resident wave slots; every
// lane of a wave64 runs one electron at a time and, at a wave-uniform point at
// the top of each step, lanes whose electron has ended claim the next electron
// id from a device counter with one aggregated atomic per wave.  Lanes thus
// stay busy while histories of very different lengths drain, instead of idling
// until the longest history of a static 64-electron wave ends.  Every draw and
// every output slot is a function of the electron id alone, so the claim order
// cannot change any result.
#include <cmath>

#include "transport_step.hpp"

namespace hr {
namespace {

// Claims up to one electron per requesting lane with a single atomic per wave.
__device__ __forceinline__ int64_t claim(unsigned long long *next, bool want, int64_t limit) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const int count = __popcll(mask);
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(next, (unsigned long long)count);
    base = __shfl(base, leader);
    if (!want) return -1;
    const int rank = __popcll(mask & ((1ULL << lane) - 1ULL));
    const int64_t id = (int64_t)base + rank;
    return id < limit ? id : -1;
}

// Optional occupancy target (waves per SIMD) for tuning builds; the default
// lets the compiler keep the whole step in registers.
#ifdef HR_TRANSPORT_WAVES_PER_EU
#define HR_TRANSPORT_OCCUPANCY __attribute__((amdgpu_waves_per_eu(HR_TRANSPORT_WAVES_PER_EU, 8)))
#else
#define HR_TRANSPORT_OCCUPANCY
#endif

template <bool kMixed, bool kHard = false, int kCone = 0>
__global__ void __launch_bounds__(256) HR_TRANSPORT_OCCUPANCY transport_kernel(hr_transport_args a,
                                                        unsigned long long *__restrict__ next) {
    Lane s;
    bool live = false;
    bool drained = false;
    while (true) {
        // Wave-uniform point: lanes without an electron claim the next id.
        const bool want = !live && !drained;
        const int64_t i = claim(next, want, a.e_count);
        if (want) {
            if (i < 0) {
                drained = true;
            } else {
                const int64_t e = a.e_start + i;
                if constexpr (kCone != 0) {
                    a.cone.count[i] = 0;
                    a.cone.killed[i] = 0;
                }
                if (!a.alive[e]) {
                    a.seg_count[i] = 0;
                    a.exit_code[i] = EXIT_NOT_ENTERED;
                    if (a.straggle_on) a.stragg_dE[e] = 0.0;
                } else {
                    load_lane(a, e, s);
                    live = a.max_steps > 0;
                    if (!live) store_lane(a, s, EXIT_STEP_LIMITED);
                }
            }
        }
        if (__ballot(live || !drained) == 0) break;
        if (live) {
            const int code = step_lane<true, kMixed, kHard, kCone>(a, s);
            if (code >= 0) {
                store_lane(a, s, (int8_t)code);
                live = false;
            }
        }
    }
}

// Bounded continuation: lane i of the active population advances its history
// by at most a.cap rows, resuming from the loop-carried record of the previous
// launch.  The record holds every value step_lane carries between steps (open
// optical depth, stream counter, flight/substep/step counters; the straggling
// accumulator stays in stragg_dE), so a history split over any number of
// launches emits the rows of one uninterrupted run, bit for bit.
template <bool kMixed>
__global__ void __launch_bounds__(256) HR_TRANSPORT_OCCUPANCY continue_kernel(hr_transport_args a,
                                                                               hr_continuation c) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.e_count) return;
    const int64_t h = c.history[i];
    if (h < 0) {
        a.seg_count[i] = 0;
        a.exit_code[i] = (int8_t)HR_LANE_EMPTY;
        return;
    }
    const bool fresh = c.steps[i] == 0 && c.nseg[i] == 0;
    if (fresh && !a.alive[i]) {
        a.seg_count[i] = 0;
        a.exit_code[i] = EXIT_NOT_ENTERED;
        if (a.straggle_on) a.stragg_dE[i] = 0.0;
        return;
    }
    Lane s;
    load_lane(a, i, s);
    s.id = h;
    s.tau = c.tau[i];
    s.draws = c.draws[i];
    s.flight = c.flight[i];
    s.substep = c.substep[i];
    s.steps = c.steps[i];
    if (a.straggle_on && !fresh) s.stragg = a.stragg_dE[i];
    int code = -1;
    if (a.max_steps <= 0) {
        code = EXIT_STEP_LIMITED;
    } else {
        while (code < 0 && s.nseg < a.cap) code = step_lane<true, kMixed>(a, s);
    }
    store_lane(a, s, code < 0 ? (int8_t)EXIT_SUSPENDED : (int8_t)code);
    c.tau[i] = s.tau;
    c.draws[i] = s.draws;
    c.flight[i] = s.flight;
    c.substep[i] = s.substep;
    c.steps[i] = s.steps;
    c.nseg[i] += s.nseg;
}

__global__ void reset_counter(unsigned long long *c) { *c = 0ULL; }

// Resident 256-lane blocks per CU of transport_kernel, per variant (0 plain,
// 1 hard-event modes, 2 cone splitting, 3 cone splitting on tabulated angles)
// and device.
int g_occupancy[4][64];
hr_status transport_occupancy(int block, int variant, int &occupancy) {
    int dev = 0;
    HR_CHECK_HIP(hipGetDevice(&dev));
    HR_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
    HR_REQUIRE(variant >= 0 && variant < 4, "transport variant %d out of range", variant);
    if (block == 256 && g_occupancy[variant][dev] > 0) {
        occupancy = g_occupancy[variant][dev];
        return HR_OK;
    }
    int blocks = 0;
    if (variant == 1)
        HR_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, transport_kernel<false, true>, block, 0));
    else if (variant == 2)
        HR_CHECK_HIP(
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, transport_kernel<false, false, 1>, block, 0));
    else if (variant == 3)
        HR_CHECK_HIP(
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, transport_kernel<false, false, 2>, block, 0));
    else
        HR_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, transport_kernel<false>, block, 0));
    occupancy = blocks > 0 ? blocks : 1;
    if (block == 256) g_occupancy[variant][dev] = occupancy;
    return HR_OK;
}

// Pointers and values the cone splitting reads; checked before any launch.
hr_status check_cone(const hr_transport_args &a, const char *who) {
    const hr_cone_split &c = a.cone;
    HR_REQUIRE(c.on == 0 || c.on == 1, "%s: cone.on must be 0 or 1", who);
    if (!c.on) return HR_OK;
    HR_REQUIRE(!a.hard.rad_on && !a.hard.shell_on, "%s: cone splitting is not combined with the hard-event modes",
               who);
    HR_REQUIRE(c.split_key && c.count && c.killed, "%s: cone splitting needs split keys, count and killed", who);
    HR_REQUIRE(c.cap >= 1 && c.pos && c.dir && c.E_keV && c.clock && c.weight && c.flight,
               "%s: cone splitting needs copy slots (cap >= 1)", who);
    const double nn = c.nx * c.nx + c.ny * c.ny + c.nz * c.nz;
    HR_REQUIRE(fabs(nn - 1.0) < 1e-9, "%s: cone direction must be a unit vector", who);
    HR_REQUIRE(c.sin_c > 0.0 && c.cos_c > 0.0 && fabs(c.cos_c * c.cos_c + c.sin_c * c.sin_c - 1.0) < 1e-9,
               "%s: cone half-angle must lie in (0, pi/2)", who);
    HR_REQUIRE(c.target_weight > 0.0 && std::isfinite(c.target_weight), "%s: target_weight must be positive", who);
    HR_REQUIRE(c.min_energy_keV >= 0.0 && std::isfinite(c.min_energy_keV), "%s: min_energy_keV must be >= 0", who);
    HR_REQUIRE(c.angle_scale >= 0.0 && std::isfinite(c.angle_scale), "%s: angle_scale must be >= 0 (0: off)", who);
    return HR_OK;
}

// Pointers and sizes the hard-event modes read; checked before any launch.
hr_status check_hard(const hr_transport_args &a, const char *who) {
    const hr_hard_events &h = a.hard;
    HR_REQUIRE((h.rad_on == 0 || h.rad_on == 1) && (h.shell_on == 0 || h.shell_on == 1),
               "%s: rad_on and shell_on must be 0 or 1", who);
    if (!h.rad_on && !h.shell_on) return HR_OK;
    HR_REQUIRE(a.energy_model == 1, "%s: the hard-event modes need the midpoint energy model", who);
    if (h.rad_on) {
        HR_REQUIRE(h.rad_key && h.lut_rad_soft && h.lut_rad_rate && h.rad_incident && h.rad_nominal && h.rad_top &&
                       h.rad_chi && h.rad_z && h.rad_ncm3 && h.rad_nt && h.rad_plasma,
                   "%s: radiative mode needs its keys, LUT columns and tables", who);
        HR_REQUIRE(h.rad_max_t >= 2 && h.rad_n_red >= 2 && h.rad_cutoff_eV > 0.0,
                   "%s: radiative tables need >= 2 incident rows and reduced nodes and a positive cutoff", who);
        HR_REQUIRE(a.max_el <= 8, "%s: radiative mode takes at most 8 elements per layer (max_el=%d)", who, a.max_el);
        HR_REQUIRE(a.cap == 0 || ((h.seg_rad_k == nullptr) == (h.seg_rad_Z == nullptr)),
                   "%s: seg_rad_k and seg_rad_Z are written together", who);
    }
    if (h.shell_on) {
        HR_REQUIRE(h.lut_sh_rate && h.lut_sh_omega2 && h.lut_sh_ch_rate && h.sh_U && h.sh_W && h.sh_branch &&
                       h.sh_code && h.sh_nch,
                   "%s: shell mode needs its LUT columns and channel tables", who);
        HR_REQUIRE(h.sh_max_ch >= 1 && h.sh_cutoff_eV > 0.0, "%s: shell mode needs channels and a positive cutoff",
                   who);
        HR_REQUIRE(a.cap == 0 || ((h.seg_hard_W == nullptr) == (h.seg_hard_ch == nullptr)),
                   "%s: seg_hard_W and seg_hard_ch are written together", who);
    }
    return HR_OK;
}

}  // namespace
}  // namespace hr

extern "C" hr_status hr_transport_run(const hr_transport_args *args, void *stream) {
    HR_TRACE("hr_transport_run");
    using namespace hr;
    HR_REQUIRE(args != nullptr, "hr_transport_run: null args");
    const hr_transport_args &a = *args;
    HR_REQUIRE(a.e_start >= 0 && a.e_count >= 0 && a.cap >= 0, "hr_transport_run: negative range or cap");
    HR_REQUIRE(a.n_layers >= 1 && a.max_el >= 1 && a.lut_n_energy >= 2, "hr_transport_run: bad table shape");
    HR_REQUIRE(a.energy_model == 0 || a.energy_model == 1, "hr_transport_run: energy_model must be 0 or 1");
    if (a.e_count == 0) return HR_OK;
    HR_REQUIRE(a.stream_key && a.alive && a.clock && a.pos && a.dirs && a.E_keV && a.E_cut &&
                   a.seg_count && a.exit_code && a.L_nel && a.L_top && a.L_bot && a.lut_total_rate &&
                   a.lut_dEds && a.lut_inv_beta && a.lut_cdf && a.lut_alpha,
               "hr_transport_run: null state or table pointer");
    HR_REQUIRE(!a.straggle_on || (a.stragg_dE && a.L_Z && a.L_J && a.L_k && a.L_coeff && a.L_E_cross),
               "hr_transport_run: straggling needs stragg_dE and element tables");
    HR_REQUIRE(a.cap == 0 || (a.seg_dir && a.seg_mid && a.seg_len && a.seg_E && a.seg_t0 && a.seg_id &&
                              a.seg_lay),
               "hr_transport_run: null row pointer");
    HR_REQUIRE(a.cap == 0 || a.energy_model == 0 ||
                   (a.seg_E_end && a.seg_t_end && a.seg_flight && a.seg_substep && a.seg_event),
               "hr_transport_run: midpoint rows need E_end/t_end/flight/substep/event");
    HR_REQUIRE(a.precision == 0 || a.precision == 1, "hr_transport_run: precision must be 0 (fp64) or 1 (mixed)");
    HR_REQUIRE(a.precision == 0 || (a.lut_total_rate_f && a.lut_dEds_f && a.lut_inv_beta_f && a.lut_cdf_f &&
                                    a.lut_alpha_f),
               "hr_transport_run: mixed precision needs the FP32 table copies");
    HR_REQUIRE(a.work != nullptr, "hr_transport_run: null claim counter (work)");
    hr_status st = check_hard(a, "hr_transport_run");
    if (st != HR_OK) return st;
    st = check_cone(a, "hr_transport_run");
    if (st != HR_OK) return st;
    st = elsepa::check(a, "hr_transport_run");
    if (st != HR_OK) return st;
    const bool hard = a.hard.rad_on || a.hard.shell_on;
    const bool cone = a.cone.on != 0;
    int cu = 0, occupancy = 0;
    st = device_cu(cu);
    if (st != HR_OK) return st;
    // Whole wave64s, at most the kernel's launch bound of 256 lanes.
    const int block = a.block_size <= 0 ? 256 : (a.block_size >= 256 ? 256 : pick_block(a.block_size, 256));
    const bool cone_tab = cone && a.elsepa.on;
    st = transport_occupancy(block, hard ? 1 : cone_tab ? 3 : cone ? 2 : 0, occupancy);
    if (st != HR_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, a.work);
    HR_CHECK_HIP(hipGetLastError());
    const int64_t need = (a.e_count + block - 1) / block;
    const int64_t resident = (int64_t)cu * occupancy;
    const int64_t grid = need < resident ? need : resident;
    if (cone_tab && a.precision == 1)
        hipLaunchKernelGGL((transport_kernel<true, false, 2>), dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    else if (cone_tab)
        hipLaunchKernelGGL((transport_kernel<false, false, 2>), dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    else if (cone && a.precision == 1)
        hipLaunchKernelGGL((transport_kernel<true, false, 1>), dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    else if (cone)
        hipLaunchKernelGGL((transport_kernel<false, false, 1>), dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    else if (hard && a.precision == 1)
        hipLaunchKernelGGL((transport_kernel<true, true>), dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    else if (hard)
        hipLaunchKernelGGL((transport_kernel<false, true>), dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    else if (a.precision == 1)
        hipLaunchKernelGGL(transport_kernel<true>, dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    else
        hipLaunchKernelGGL(transport_kernel<false>, dim3((unsigned)grid), dim3(block), 0, s, a, a.work);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}

extern "C" hr_status hr_transport_continue(const hr_transport_args *args, const hr_continuation *cont,
                                           void *stream) {
    HR_TRACE("hr_transport_continue");
    using namespace hr;
    HR_REQUIRE(args != nullptr && cont != nullptr, "hr_transport_continue: null args");
    const hr_transport_args &a = *args;
    const hr_continuation &c = *cont;
    HR_REQUIRE(a.e_start == 0 && a.e_count >= 0, "hr_transport_continue: lanes are [0, e_count), e_start must be 0");
    HR_REQUIRE(a.cap >= 1 && a.cap <= (1 << 30), "hr_transport_continue: cap (rows per launch) %lld outside [1, 2^30]",
               (long long)a.cap);
    HR_REQUIRE(a.e_count <= (1LL << 40), "hr_transport_continue: too many lanes");
    HR_REQUIRE(a.n_layers >= 1 && a.max_el >= 1 && a.lut_n_energy >= 2, "hr_transport_continue: bad table shape");
    HR_REQUIRE(a.energy_model == 0 || a.energy_model == 1, "hr_transport_continue: energy_model must be 0 or 1");
    if (a.e_count == 0) return HR_OK;
    HR_REQUIRE(a.stream_key && a.alive && a.clock && a.pos && a.dirs && a.E_keV && a.E_cut && a.seg_count &&
                   a.exit_code && a.L_nel && a.L_top && a.L_bot && a.lut_total_rate && a.lut_dEds &&
                   a.lut_inv_beta && a.lut_cdf && a.lut_alpha,
               "hr_transport_continue: null state or table pointer");
    HR_REQUIRE(c.history && c.tau && c.draws && c.nseg && c.flight && c.substep && c.steps,
               "hr_transport_continue: null continuation record");
    HR_REQUIRE(!a.straggle_on || (a.stragg_dE && a.L_Z && a.L_J && a.L_k && a.L_coeff && a.L_E_cross),
               "hr_transport_continue: straggling needs stragg_dE and element tables");
    HR_REQUIRE(a.seg_dir && a.seg_mid && a.seg_len && a.seg_E && a.seg_t0 && a.seg_id && a.seg_lay,
               "hr_transport_continue: null row pointer");
    HR_REQUIRE(a.energy_model == 0 || (a.seg_E_end && a.seg_t_end && a.seg_flight && a.seg_substep && a.seg_event),
               "hr_transport_continue: midpoint rows need E_end/t_end/flight/substep/event");
    HR_REQUIRE(a.precision == 0 || (a.precision == 1 && a.lut_total_rate_f && a.lut_dEds_f && a.lut_inv_beta_f &&
                                    a.lut_cdf_f && a.lut_alpha_f),
               "hr_transport_continue: precision 0, or 1 with the FP32 table copies");
    HR_REQUIRE(!a.cone.on, "hr_transport_continue: cone splitting runs on hr_transport_run only (exit code 6 "
                           "is a suspended lane here)");
    HR_REQUIRE(!a.hard.rad_on && !a.hard.shell_on,
               "hr_transport_continue: the hard-event modes run on hr_transport_run only (their optical depths "
               "and counters are not in the continuation record)");
    const hr_status est = elsepa::check(a, "hr_transport_continue");
    if (est != HR_OK) return est;
    const unsigned grid = (unsigned)((a.e_count + 255) / 256);
    if (a.precision == 1)
        hipLaunchKernelGGL(continue_kernel<true>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, c);
    else
        hipLaunchKernelGGL(continue_kernel<false>, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, c);
    HR_CHECK_HIP(hipGetLastError());
    return HR_OK;
}
